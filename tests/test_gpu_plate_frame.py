"""Mindlin plates outside one z = const plane through the C ABI against the
oracle: every element whose nodes do not share one z is computed in its own
frame (GeomElem::_init_local_elem_2d, src/mesh/geom_elem.cpp:460-511) and
goes to the global system through T (StructuralElementBase::transform_*,
src/elasticity/structural_element_base.cpp:1568-1660), all six variables
active, MAST's 1e-8 on the local theta_z diagonal in the Jacobian.

Geometries: the 6x5 plate rotated rigidly in space (every element framed),
a plate folded along a line (flat and framed elements side by side), and a
plate with warped elements (nodes moved off the plane: the surface map with
a local z).  Bars: the 6-variable matrix within 1e-12 of max(|K_ij|, 1e-2 x
the row's largest entry); residual at a random state within 1e-12 of its
largest entry; per-element sensitivity dots within 1e-12; dR/dp against the
oracle's residual differences within 1e-10 of the largest entry for E (R is
linear in E), 1e-8 for nu (step 1e-6), and for the pressure against the
oracle's residual at zero state over p (the load alone) within 1e-12;
compliance within 1e-9 and dC/dh within 1e-7 after CG to 1e-10.
"""
import numpy as np
import pytest

from helpers import ctx_from_oracle, rel_err_floor

pytestmark = pytest.mark.gpu

MAT = dict(E=70e9, nu=0.3, kappa=5.0 / 6.0, off=0.0, pressure=1e3)


def _rotation(axis, angle):
    a = np.asarray(axis, float)
    a /= np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * K @ K


def _geometry(xyz, kind):
    if kind == "rotated":
        return (xyz - 0.3) @ _rotation([1, 2, 0.5], 0.7).T + np.array([0.1, -0.2, 0.4])
    out = xyz.copy()
    if kind == "folded":
        # fold at x = 0.5 by 40 degrees: x > 0.5 rises along the fold
        x0 = 0.5 + 1e-9
        t = np.maximum(out[:, 0] - x0, 0.0)
        out[:, 0] = np.minimum(out[:, 0], x0) + t * np.cos(0.7)
        out[:, 2] = t * np.sin(0.7)
        return out
    # warped: interior nodes off the plane by up to 5% of the mesh size
    lo, hi = xyz.min(0), xyz.max(0)
    inner = np.all((xyz[:, :2] > lo[:2] + 1e-12) & (xyz[:, :2] < hi[:2] - 1e-12), axis=1)
    out[inner, 2] += 0.05 * (hi[0] - lo[0]) / 6 * np.random.default_rng(3).uniform(-1, 1, int(inner.sum()))
    return out


def _problem(oracle, kind, mesh=(6, 5), **mat):
    P0 = oracle.Problem(oracle.PLATE, *mesh, seed=2, mast_exact=True)
    xyz, conn = P0.mesh()
    base, con = P0.dofmap()
    P = oracle.Problem(oracle.PLATE, arrays=dict(xyz=_geometry(xyz, kind), conn=conn, n_vars=6, node_base=base,
                                                 constrained=con, params=P0.params()))
    m = dict(MAT)
    m.update(mat)
    P.set(**m)
    return P


def _ctx(oracle, mastg, kind, mesh=(6, 5)):
    P = _problem(oracle, kind, mesh)
    return P, ctx_from_oracle(mastg, P)


@pytest.mark.parametrize("kind", ["rotated", "folded", "warped"])
def test_plate_frame_matrix_and_residual(oracle, mastg, kind):
    print(f"[plate frame K, R] {kind} 6x5", flush=True)
    P, ctx = _ctx(oracle, mastg, kind)
    _, con = P.dofmap()
    rng = np.random.default_rng(5)
    X = rng.standard_normal(P.n_dofs) * 1e-4
    X[con == 1] = 0.0
    R0, K0 = P.assemble(X)
    R, _ = ctx.assemble(X)
    eR = np.max(np.abs(R - R0)) / np.abs(R0).max()
    rp, col, val = ctx.export_csr_full()
    orp, ocol = P.csr()
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    eK = rel_err_floor(val, K0, orp, floor=1e-2)
    lam = rng.standard_normal(P.n_dofs)
    eg = rel_err_floor(ctx.sensitivity(X, lam), P.sens_dots(X, lam))
    print(f"  R {eR:.1e}, K {eK:.1e}, sensitivity dots {eg:.1e}", flush=True)
    assert eR <= 1e-12 and eK <= 1e-12 and eg <= 1e-12
    ctx.close()


@pytest.mark.parametrize("kind", ["rotated", "folded"])
def test_plate_frame_global_dRdp(oracle, mastg, kind):
    print(f"[plate frame dR/dp] {kind} 6x5", flush=True)
    P, ctx = _ctx(oracle, mastg, kind)
    _, con = P.dofmap()
    X = np.random.default_rng(7).standard_normal(P.n_dofs) * 1e-4
    X[con == 1] = 0.0
    # linear in E (exact differences), smooth in nu (step 1e-6)
    for param, name, step, bar in (("E", "E", 1e-3, 1e-10), ("NU", "nu", 1e-6, 1e-8)):
        g = ctx.sensitivity_rhs(getattr(mastg, "PARAM_" + param), X)
        d = step * MAT[name]
        Rp, _ = _problem(oracle, kind, **{name: MAT[name] + d}).assemble(X, want_K=False)
        Rm, _ = _problem(oracle, kind, **{name: MAT[name] - d}).assemble(X, want_K=False)
        fd = (Rp - Rm) / (2 * d)
        err = np.max(np.abs(g - fd)) / np.abs(fd).max()
        print(f"  {param}: {err:.1e}", flush=True)
        assert err <= bar
    # the pressure enters R only through the load: dR/dp = R(X = 0) / p
    g = ctx.sensitivity_rhs(mastg.PARAM_PRESSURE, X)
    R0, _ = P.assemble(np.zeros(P.n_dofs), want_K=False)
    err = np.max(np.abs(g - R0 / MAT["pressure"])) / np.abs(R0 / MAT["pressure"]).max()
    print(f"  PRESSURE: {err:.1e}", flush=True)
    assert err <= 1e-12
    ctx.close()


@pytest.mark.parametrize("kind,mesh", [("rotated", (6, 5)), ("folded", (12, 8)), ("warped", (6, 5))])
def test_plate_frame_design_eval(oracle, mastg, kind, mesh):
    print(f"[plate frame design eval] {kind} {mesh}", flush=True)
    P, ctx = _ctx(oracle, mastg, kind, mesh)
    ref = P.solve_elastic(1e-12)
    r = ctx.design_eval(P.params())
    ec = abs(r["objective"] - ref["compliance"]) / abs(ref["compliance"])
    eg = rel_err_floor(r["grad"], ref["dc"], floor=1e-6)
    print(f"  C {ec:.1e}, grad {eg:.1e}, {r['iters']} iterations", flush=True)
    assert ec <= 1e-9 and eg <= 1e-7
    # no multigrid for six active variables
    with pytest.raises(mastg.MastgError):
        ctx.design_eval(P.params(), solver=mastg.SOLVER_MG)
    ctx.close()


def test_plate_frame_needs_six_variables(oracle, mastg):
    P = _problem(oracle, "folded")
    xyz, conn = P.mesh()
    desc = mastg.desc_from_arrays(mastg.QUAD4_MINDLIN, xyz, conn, 5, E=70e9, nu=0.3, kappa=5.0 / 6.0)
    with pytest.raises(mastg.MastgError):
        mastg.Context(desc)
