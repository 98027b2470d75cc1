"""GPU parity of the design-space steps (density filter, volume, optimality-
criteria update) against the oracle's numpy restatement (oracle/design.py:
MAST::FilterBase's all-pairs filter map, example_6's volume, the OC step by
bisection), through the C ABI, and a short SIMP loop (filter -> design
evaluation -> chain rule -> OC) against the same loop on the oracle.

Bars: filter and its transpose 1e-14 of the largest entry; volume and its
gradient 1e-14; the OC step 1e-8 (its multiplier comes from a 7-pass
multi-section to ~1e-10 relative against the oracle's bisection to 1e-15);
the SIMP loop's compliances 1e-6 per iteration over 5 iterations.
"""
import numpy as np
import pytest

import design as D
from helpers import ctx_from_oracle

pytestmark = pytest.mark.gpu


def _desc(mastg, P):
    xyz, conn = P.mesh()
    phys = {1: mastg.HEX8_ELASTIC, 2: mastg.QUAD4_MINDLIN, 3: mastg.HEX8_HEAT}[P.kind]
    return mastg.desc_from_arrays(phys, xyz, conn, P.n_vars), xyz, conn


@pytest.mark.parametrize("kind,mesh,r_h,distort", [("hex8", (8, 4, 4), 1.5, 0.0), ("hex8", (6, 5, 4), 2.2, 0.2),
                                                    ("plate", (10, 8), 2.5, 0.0)])
def test_filter_and_volume(oracle, mastg, kind, mesh, r_h, distort):
    print(f"[design filter] {kind} {mesh} r = {r_h} h, distortion {distort}", flush=True)
    P = oracle.Problem(oracle.CANTILEVER if kind == "hex8" else oracle.PLATE, *mesh, seed=2, lo=0.3, hi=1.0)
    desc, xyz, conn = _desc(mastg, P)
    if distort:
        # move interior nodes (a general hex8 mesh; the design space does not need a lattice)
        rng = np.random.default_rng(4)
        h = 1.0 / mesh[0]
        xyz = xyz + distort * h * rng.uniform(-0.5, 0.5, xyz.shape)
        desc = mastg.desc_from_arrays(mastg.HEX8_ELASTIC, xyz, conn, P.n_vars)
    h = 1.0 / mesh[0]
    Dg = mastg.Design(desc, r_h * h)
    c = D.centroids(xyz, conn)
    W = D.filter_matrix(c, r_h * h)
    assert Dg.nnz == int(np.count_nonzero(np.sqrt(((c[:, None] - c[None]) ** 2).sum(-1)) <= r_h * h))
    x = np.random.default_rng(1).uniform(0.0, 1.0, P.n_elems)
    y = np.random.default_rng(2).standard_normal(P.n_elems)
    fx = Dg.filter(x)
    fty = Dg.filter(y, transpose=True)
    assert np.max(np.abs(fx - W @ x)) <= 1e-14 * np.abs(W @ x).max()
    assert np.max(np.abs(fty - W.T @ y)) <= 1e-14 * np.abs(W.T @ y).max()
    assert np.allclose(Dg.filter(np.ones(P.n_elems)), 1.0, rtol=0, atol=1e-14)
    Ve = D.hex8_volumes(xyz, conn) if kind == "hex8" else D.quad4_areas(xyz, conn)
    v, dv = Dg.volume(x, want_grad=True)
    assert abs(v - np.dot(Ve, x)) <= 1e-14 * abs(v) * 10
    assert np.max(np.abs(dv - Ve)) <= 1e-14 * Ve.max()
    assert abs(Dg.v_total - Ve.sum()) <= 1e-14 * Ve.sum() * 10
    Dg.close()


def test_oc_update(oracle, mastg):
    print("[design OC] hex8 16x8x8", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, 16, 8, 8, seed=2, lo=0.3, hi=1.0)
    desc, xyz, conn = _desc(mastg, P)
    Dg = mastg.Design(desc, 1.5 / 16)
    rng = np.random.default_rng(5)
    rho = rng.uniform(0.05, 0.95, P.n_elems)
    dc = -rng.uniform(1e-3, 1e2, P.n_elems)
    Ve = D.hex8_volumes(xyz, conn)
    for vf, move, rmin in [(0.4, 0.2, 1e-3), (0.35, 0.3, 0.01), (0.8, 0.5, 0.0)]:
        new, lam = Dg.oc_update(rho, dc, Ve, vf, move, rmin)
        ref, lam0 = D.oc_update(rho, dc, Ve, Ve.sum(), vf, move, rmin)
        err = np.max(np.abs(new - ref))
        print(f"  vf {vf}: lambda {lam:.12e} (oracle {lam0:.12e}), max |rho - oracle| {err:.2e}", flush=True)
        assert err <= 1e-8
        assert np.dot(Ve, new) <= vf * Ve.sum() * (1 + 1e-12)
        assert abs(np.dot(Ve, new) - vf * Ve.sum()) <= 1e-8 * Ve.sum()
    # a target the move limit cannot reach: every density takes its lower bound
    new, _ = Dg.oc_update(rho, dc, Ve, 0.1, 0.05, 0.01)
    assert np.array_equal(new, np.maximum(0.01, rho - 0.05))
    Dg.close()


def test_simp_loop_against_oracle(oracle, mastg):
    nx, ny, nz = 12, 6, 6
    vf, move, rmin, r = 0.4, 0.2, 0.01, 1.5 / 12
    print(f"[design SIMP loop] hex8 {nx}x{ny}x{nz}, 5 iterations", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, nx, ny, nz, seed=1, lo=0.3, hi=1.0)
    xyz, conn = P.mesh()
    ctx = ctx_from_oracle(mastg, P)
    desc, _, _ = _desc(mastg, P)
    Dg = mastg.Design(desc, r)
    W = D.filter_matrix(D.centroids(xyz, conn), r)
    Ve = D.hex8_volumes(xyz, conn)
    x_g = np.full(P.n_elems, vf)
    x_o = x_g.copy()
    for it in range(5):
        # GPU: filter, multigrid design evaluation, chain rule, OC
        rt = Dg.filter(x_g)
        res = ctx.design_eval(rt, solver=mastg.SOLVER_MG)
        dcx = Dg.filter(res["grad"], transpose=True)
        _, dvr = Dg.volume(rt, want_grad=True)
        dvx = Dg.filter(dvr, transpose=True)
        x_g, _ = Dg.oc_update(x_g, dcx, dvx, vf, move, rmin)
        # oracle loop
        P.set_params(W @ x_o)
        ref = P.solve_elastic(1e-10)
        x_o, _ = D.oc_update(x_o, W.T @ ref["dc"], W.T @ Ve, Ve.sum(), vf, move, rmin)
        ec = abs(res["objective"] - ref["compliance"]) / ref["compliance"]
        print(f"  iteration {it}: C {res['objective']:.10g} (oracle {ref['compliance']:.10g}, rel {ec:.1e}), "
              f"max |x - oracle x| {np.max(np.abs(x_g - x_o)):.1e}", flush=True)
        assert ec <= 1e-6
    assert np.max(np.abs(x_g - x_o)) <= 1e-5
    ctx.close()
    Dg.close()
