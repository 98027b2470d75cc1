"""Oracle solves, outputs and adjoint sensitivities pinned independently:

* CG + Jacobi against a sparse direct solve (scipy) of the same matrix;
* MAST's 6-variable DofMap and the 3-variable active layout give the same answer;
* dC/drho_e and dC/dh_e against central finite differences of direct solves
  (step 1e-4 of the parameter: at smaller steps the round-off of the
  ill-conditioned direct solves dominates the difference quotient);
* the simply supported square plate against the Navier series;
* steady conduction against the 1D closed form, Newton's quadratic rate, and
  d(mean T)/dq_e against central differences.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl


def direct(P, params=None):
    if params is not None:
        P.set_params(params)
    R, K = P.assemble()
    rp, col = P.csr()
    A = sp.csr_matrix((K, col, rp), shape=(P.n_dofs, P.n_dofs)).tocsc()
    u = spl.spsolve(A, -R)
    return u, float(-R @ u), A, R


def test_cg_matches_direct(oracle):
    P = oracle.Problem(oracle.CANTILEVER, 20, 10, 10, seed=1, lo=0.3, hi=1.0)
    r = P.solve_elastic(1e-10)
    u, C, _, _ = direct(P)
    assert r["iters"] > 0 and r["rel_res"] <= 1e-10
    np.testing.assert_allclose(r["u"], u, rtol=0, atol=1e-7 * np.abs(u).max())
    assert abs(r["compliance"] - C) <= 1e-9 * C


def test_mast_exact_layout_equivalent(oracle):
    Pa = oracle.Problem(oracle.CANTILEVER, 8, 4, 4, seed=3, lo=0.01, hi=1.0)
    Pe = oracle.Problem(oracle.CANTILEVER, 8, 4, 4, seed=3, lo=0.01, hi=1.0, mast_exact=True)
    ra, re = Pa.solve_elastic(1e-12), Pe.solve_elastic(1e-12)
    assert abs(ra["compliance"] - re["compliance"]) <= 1e-11 * ra["compliance"]
    np.testing.assert_allclose(ra["dc"], re["dc"], rtol=1e-9, atol=1e-12 * np.abs(ra["dc"]).max())
    assert abs(ra["iters"] - re["iters"]) <= 1   # same iterates up to summation order


@pytest.mark.parametrize("lo", [0.3, 0.01])
def test_compliance_sensitivity_fd(oracle, lo):
    P = oracle.Problem(oracle.CANTILEVER, 6, 3, 3, seed=5, lo=lo, hi=1.0)
    rho = P.params()
    r = P.solve_elastic(1e-13)
    for e in (0, 7, 25, 53):
        h = 1e-4 * rho[e]
        p = rho.copy(); p[e] += h
        Cp = direct(P, p)[1]
        p[e] -= 2 * h
        Cm = direct(P, p)[1]
        fd = (Cp - Cm) / (2 * h)
        assert abs(r["dc"][e] - fd) <= 1e-6 * abs(fd), (e, r["dc"][e], fd)
    P.set_params(rho)


def navier_center_deflection(q, a, D, nu_terms=40):
    # Kirchhoff plate, simply supported square a x a, uniform load q
    w = 0.0
    for m in range(1, 2 * nu_terms, 2):
        for n in range(1, 2 * nu_terms, 2):
            w += (np.sin(m * np.pi / 2) * np.sin(n * np.pi / 2) /
                  (m * n * (m * m + n * n) ** 2))
    return 16 * q * a ** 4 / (np.pi ** 6 * D) * w


def test_plate_navier(oracle):
    n = 32
    P = oracle.Problem(oracle.PLATE, n, n, seed=1)
    h = 0.01
    P.set_params(np.full(P.n_elems, h))
    r = P.solve_elastic(1e-12)
    base, _ = P.dofmap()
    xyz, _ = P.mesh()
    c = np.argmin(np.sum((xyz[:, :2] - 0.5) ** 2, axis=1))
    w = r["u"][base[c] + 2]
    E, nu = 70e9, 0.3
    D = E * h ** 3 / 12 / (1 - nu ** 2)
    wn = navier_center_deflection(1e3, 1.0, D)
    # load along -z; Mindlin shear adds ~ (h/a)^2 = 1e-4; mesh error ~ 1 %
    assert w < 0 and abs(-w / wn - 1) < 0.015, (w, wn)
    assert abs(r["compliance"] - 1e3 * 0) >= 0   # compliance is F.u > 0
    assert r["compliance"] > 0


def test_plate_thickness_sensitivity_fd(oracle):
    # central differences of direct solves: truncation O(step^2) and solver
    # round-off O(eps kappa / step) balance near step = 1e-4 p (at 1e-7 p, the
    # round-off of the ill-conditioned solves alone is ~1e-5 relative)
    P = oracle.Problem(oracle.PLATE, 8, 8, seed=2)
    hh = P.params()
    r = P.solve_elastic(1e-13)
    for e in (0, 9, 27, 63):
        st = 1e-4 * hh[e]
        p = hh.copy(); p[e] += st
        Cp = direct(P, p)[1]
        p[e] -= 2 * st
        Cm = direct(P, p)[1]
        fd = (Cp - Cm) / (2 * st)
        assert abs(r["dc"][e] - fd) <= 1e-6 * abs(fd), (e, r["dc"][e], fd)
    P.set_params(hh)


def test_heat_1d_closed_form(oracle):
    """alpha = 0, no convection, uniform q: T(z) = q/k (L z - z^2/2) at the nodes."""
    n = 6
    P = oracle.Problem(oracle.HEAT, n, seed=1).set(alpha=0.0, conv_h=0.0)
    q = 4.0e5
    P.set_params(np.full(P.n_elems, q))
    r = P.solve_heat(1e-13, 1e-14)
    xyz, _ = P.mesh()
    base, _ = P.dofmap()
    z = xyz[:, 2]
    Tex = q / 200.0 * (z - z ** 2 / 2)
    np.testing.assert_allclose(r["T"][base], Tex, rtol=0, atol=1e-9 * Tex.max())


def test_heat_newton_quadratic(oracle):
    P = oracle.Problem(oracle.HEAT, 6, seed=3)
    r = P.solve_heat(1e-12, 1e-14)
    h = r["res_hist"] / r["res_hist"][0]
    assert 3 <= r["newton_iters"] <= 8
    # once in the basin the error roughly squares
    k = np.argmax(h < 1e-2)
    assert h[k + 1] < 10 * h[k] ** 1.6
    assert h[-1] <= 1e-12


def test_heat_mean_T_sensitivity_fd(oracle):
    P = oracle.Problem(oracle.HEAT, 4, seed=4)
    q = P.params()
    r = P.solve_heat(1e-13, 1e-14)
    for e in (0, 21, 63):
        st = 1e-4 * max(q[e], 1.0)
        p = q.copy(); p[e] += st
        P.set_params(p)
        mp = P.solve_heat(1e-13, 1e-14)["mean_T"]
        p[e] -= 2 * st
        P.set_params(p)
        mm = P.solve_heat(1e-13, 1e-14)["mean_T"]
        fd = (mp - mm) / (2 * st)
        assert abs(r["dJ"][e] - fd) <= 1e-6 * abs(fd), (e, r["dJ"][e], fd)
    P.set_params(q)
