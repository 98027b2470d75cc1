"""GPU parity on unstructured meshes read from Nastran models (tests/golden/
butterfly_*.bdf: a disk meshed as a square core and four blocks to the rim,
so three cells meet at the core's corners; shuffled sparse Nastran ids), through
the library's reader and the C ABI, against the oracle built from the same
arrays.  These meshes are not lattices: the structured-mesh fast paths
(multigrid, matrix-free classes) refuse them or do not apply, and the general
path (element kernels, segmented scatter, CG) carries them.

Bars: CSR pattern bit-exact against the oracle's (6-variable MAST pattern
restricted to the active variables); K and R at a state within 1e-12 of the
row's largest entry; compliance 1e-9 and the gradient 1e-7 after solves to
1e-10 (hex8, densities from the seed) and 1e-9 (the thin plate).
"""
import os

import numpy as np
import pytest

from helpers import rel_err_floor, restrict_csr

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _oracle_problem(oracle, m, kind, params):
    con = np.zeros(m.n_dofs, np.uint8)
    con[m.constrained] = 1
    P = oracle.Problem(kind, arrays=dict(xyz=m.xyz, conn=m.conn, n_vars=6, node_base=m.node_base,
                                         constrained=con, params=params, point_load=m.point_loads))
    d = m.desc.contents
    P.set(E=d.E, nu=d.nu)
    if kind == oracle.PLATE:
        P.set(kappa=d.kappa, off=0.0, pressure=0.0)
    return P


@pytest.mark.parametrize("name", ["butterfly_hex", "butterfly_quad"])
def test_nastran_model_assembly_and_design(oracle, mastg, name):
    print(f"[nastran] {name}", flush=True)
    m = mastg.NastranModel(os.path.join(GOLD, name + ".bdf"))
    plate = name.endswith("quad")
    kind = oracle.PLATE if plate else oracle.CANTILEVER
    nav = 5 if plate else 3
    params = m.params.copy() if plate else np.random.default_rng(8).uniform(0.1, 1.0, m.n_elems)
    P = _oracle_problem(oracle, m, kind, params)
    ctx = mastg.Context(m.desc)
    ctx.set_params(params)
    # pattern: the oracle's MAST-exact pattern restricted to the active variables
    keep = np.sort(np.concatenate([m.node_base + v for v in range(nav)]))
    orp, ocol = P.csr()
    X = np.random.default_rng(3).standard_normal(m.n_dofs) * 1e-4
    X[m.constrained] = 0.0
    R0, K0 = P.assemble(X)
    rrp, rcol, rK = restrict_csr(orp, ocol, K0, keep)
    rp, col = ctx.csr_pattern()
    assert np.array_equal(rp, rrp) and np.array_equal(col, rcol)
    R, J = ctx.assemble(X)
    assert rel_err_floor(J, rK, rrp, floor=1e-2) <= 1e-12
    assert np.max(np.abs(R - R0)[keep]) <= 1e-12 * np.abs(R0[keep]).max()
    # the lattice-only solvers refuse the unstructured mesh
    with pytest.raises(mastg.MastgError):
        ctx.design_eval(params, solver=mastg.SOLVER_MG)
    rtol = 1e-9 if plate else 1e-10
    ref = P.solve_elastic(rtol * 1e-2)
    solvers = [mastg.SOLVER_CG_JACOBI] + ([] if plate else [mastg.SOLVER_MIXED])
    for sv in solvers:
        r = ctx.design_eval(params, solver=sv, rtol=rtol)
        ec = abs(r["objective"] - ref["compliance"]) / abs(ref["compliance"])
        eg = rel_err_floor(r["grad"], ref["dc"], floor=1e-6)
        print(f"  solver {sv}: C {r['objective']:.12g} (oracle {ref['compliance']:.12g}), rel {ec:.1e}, "
              f"grad {eg:.1e}, {r['iters']} iterations", flush=True)
        assert ec <= 1e-9 and eg <= 1e-7
    ctx.close()
    m.close()
