"""The MAST-exact 6-variable structural matrix (mastg_export_csr_full) against
the oracle's assembly in MAST's own variable layout (mast_exact: all 6
structural variables per node, the 1e-20 rotational diagonal of the hex8
element, solid_element_3d.cpp:413-417, the 1e-8 theta_z diagonal of the plate
element, structural_element_2d.cpp:1215-1219, libMesh's constraint convention
on every constrained DOF, rotations included).

Bars: pattern bit-exact; values within 1e-12 of max(|K_ij|, 1e-2 x the row's
largest entry) -- the rotational rows hold only their diagonal, so those are
held to 1e-12 of themselves; for both hex8 assembly paths (fused row kernel,
element kernel + scatter), a distorted hex8 mesh (no congruence classes) and
the plate.
"""
import numpy as np
import pytest

from helpers import ctx_from_oracle, rel_err_floor

pytestmark = pytest.mark.gpu


def _check(ctx, P, X):
    ctx.assemble(X)
    rp, col, val = ctx.export_csr_full()
    orp, ocol = P.csr()
    _, K = P.assemble(X)
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    err = rel_err_floor(val, K, orp, floor=1e-2)
    print(f"  n_rows {len(rp) - 1}, nnz {len(col)}, max rel err {err:.2e}", flush=True)
    assert err <= 1e-12
    return rp, val


@pytest.mark.parametrize("mesh", [(2, 1, 1), (4, 2, 2), (6, 3, 3)])
def test_hex8_full_matrix(oracle, mastg, mesh):
    print(f"[export 6-var] hex8 {mesh}", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, *mesh, seed=3, lo=0.3, hi=1.0, mast_exact=True)
    ctx = ctx_from_oracle(mastg, P)
    _, con = P.dofmap()
    X = np.random.default_rng(2).standard_normal(P.n_dofs) * 1e-4
    X[con == 1] = 0.0
    rp, val = _check(ctx, P, X)
    # a rotational row of an unconstrained node holds only its tiny diagonal
    base, _ = P.dofmap()
    free = [n for n in range(len(base)) if not con[base[n] + 3]]
    d = base[free[0]] + 3
    row = val[rp[d]:rp[d + 1]]
    assert np.count_nonzero(row) == 1 and 0 < row.max() < 1e-5
    ctx.set_option(mastg.OPT_ROW_ASSEMBLY, 0)      # the element kernel + scatter path
    _check(ctx, P, X)
    ctx.close()


def test_hex8_full_matrix_distorted(oracle, mastg):
    print("[export 6-var] distorted hex8 5x3x3", flush=True)
    P0 = oracle.Problem(oracle.CANTILEVER, 5, 3, 3, seed=4, lo=0.2, hi=1.0, mast_exact=True)
    xyz, conn = P0.mesh()
    lo, hi = xyz.min(0), xyz.max(0)
    inner = np.all((xyz > lo + 1e-12) & (xyz < hi - 1e-12), axis=1)
    xyz = xyz.copy()
    xyz[inner] += 0.2 / 5 * np.random.default_rng(6).uniform(-1, 1, (int(inner.sum()), 3))
    base, con = P0.dofmap()
    arrays = dict(xyz=xyz, conn=conn, n_vars=6, node_base=base, constrained=con, params=P0.params(),
                  point_load=P0.point_load())
    P = oracle.Problem(oracle.CANTILEVER, arrays=arrays)
    desc = mastg.desc_from_arrays(mastg.HEX8_ELASTIC, xyz, conn, 6, node_dof_base=base,
                                  constrained_dofs=np.nonzero(con)[0], point_loads=P0.point_load(),
                                  E=72e9, nu=0.33, penalty=3.0)
    ctx = mastg.Context(desc)
    ctx.set_params(P0.params())
    X = np.random.default_rng(3).standard_normal(P.n_dofs) * 1e-4
    X[con == 1] = 0.0
    _check(ctx, P, X)
    ctx.close()


def test_plate_full_matrix(oracle, mastg):
    print("[export 6-var] plate 6x5", flush=True)
    P = oracle.Problem(oracle.PLATE, 6, 5, seed=2, mast_exact=True)
    ctx = ctx_from_oracle(mastg, P)
    _, con = P.dofmap()
    X = np.random.default_rng(5).standard_normal(P.n_dofs) * 1e-4
    X[con == 1] = 0.0
    _check(ctx, P, X)
    ctx.close()


def test_export_refusals(oracle, mastg):
    # the active-variable layout has no 6-variable matrix
    P = oracle.Problem(oracle.CANTILEVER, 2, 1, 1, seed=1, lo=0.3, hi=1.0)
    ctx = ctx_from_oracle(mastg, P)
    with pytest.raises(mastg.MastgError):
        ctx.export_csr_full()
    ctx.close()
    # no assembly yet
    P = oracle.Problem(oracle.CANTILEVER, 2, 1, 1, seed=1, lo=0.3, hi=1.0, mast_exact=True)
    ctx = ctx_from_oracle(mastg, P)
    with pytest.raises(mastg.MastgError):
        ctx.export_csr_full()
    ctx.close()
