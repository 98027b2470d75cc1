"""GPU results at the full benchmark sizes against the FP64 oracle.

* C2 hex8 SIMP cantilever 128x64x64 (1,635,075 active DOF): the oracle's
  complete FP64 design evaluation (element loop, row-search scatter, CG +
  Jacobi to 1e-10, compliance, dC/drho_e; about a minute on 16 host cores) on
  exactly the context's inputs, against the multigrid design evaluation and
  the mixed-precision path: compliance within 1e-9 and all 524,288 dC/drho_e
  within 1e-7 (relative, floor 1e-6 of the largest).

* C3 Mindlin plate: dC/dh_e by the adjoint against central finite
  differences with step 1e-7 h_e on 64 seeded elements.  One element's
  thickness moves C by ~1e-12 relative at this step, so C(h + d) - C(h - d)
  cannot be formed by subtracting two compliances.  It is formed by the exact
  identity
      C(h+d) - C(h-d) = F.u+ - F.u- = -u+^T (K(h+d) - K(h-d)) u-
  (K symmetric, F independent of h; K(h+d) - K(h-d) is non-zero only in the
  element's own block, from the oracle's element routine at h_e +- d): the
  central difference quotient itself, evaluated without cancellation.
  - At 32^2 and 64^2, u+ and u- are exact: direct sparse solves (with one
    step of refinement) of the oracle's assembled FP64 matrices, independent
    of the GPU.  The oracle's adjoint at the exact state is held to them at
    1e-7 (measured 6e-9 at 32^2, 1.5e-8 at 64^2) and the GPU's design-
    evaluation gradient (multigrid CG to 1e-9, the plate's FP64 floor at these
    sizes) at 1e-5.
  - K(h+d) - K(h-d) comes from the element's exact thickness polynomial
    (membrane and shear ~ h, bending ~ h^3): subtracting the two element
    matrices at d = 1e-7 h loses 9 digits, which the elements whose
    displacement is nearly rigid (|K| |u|^2 ~ 1e10 u^T K u) turn into errors
    of 1e-5 and more.
  - At 512^2 the thin plate's FP64 floor (||b - K u|| / ||b|| bottoms out
    near 2e-8) stops solves at 1e-7, and two solves that stop after different
    numbers of iterations differ by up to a few 1e-4 of the small
    displacements next to the supported edges, more than the bar.  There u,
    u+ and u- are the GPU's multigrid iterates after the same fixed number of
    iterations (a consistent discrete state), and the central difference is
    held at 1e-5 to the library's sensitivity pass at u (mastg_sensitivity,
    lambda = u); the design evaluation's gradient is held to that one at its
    solve accuracy, 1e-3 of the largest entry.
"""
import numpy as np
import pytest

from helpers import ctx_from_oracle, rel_err_floor

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(900)
def test_c2_full_size_against_oracle(oracle, mastg):
    print("[full size vs oracle] C2 hex8 128x64x64: oracle FP64 design evaluation (CPU, ~1 min)",
          flush=True)
    P = oracle.Problem(oracle.CANTILEVER, 128, 64, 64, seed=1, lo=0.01, hi=1.0)
    ctx = ctx_from_oracle(mastg, P)
    rho = P.params()
    # the library's own C2 workload has the same densities
    w = mastg.Workload(1, 128, 64, 64, seed=1, lo=0.01, hi=1.0)
    assert np.array_equal(w.params, rho)
    rm = ctx.design_eval(rho, solver=mastg.SOLVER_MG)
    rx = ctx.design_eval(rho, solver=mastg.SOLVER_MIXED)
    print(f"  GPU multigrid C={rm['objective']:.12g} ({rm['iters']} it), mixed "
          f"C={rx['objective']:.12g} ({rx['iters']} it)", flush=True)
    ref = P.solve_elastic(1e-10)
    C0, g0 = ref["compliance"], ref["dc"]
    print(f"  oracle C={C0:.12g} ({ref['iters']} CG+Jacobi it, rel res {ref['rel_res']:.2e})", flush=True)
    summary = []
    for name, r in (("multigrid", rm), ("mixed", rx)):
        ec = abs(r["objective"] - C0) / abs(C0)
        eg = rel_err_floor(r["grad"], g0, floor=1e-6)
        print(f"  {name}: compliance rel err {ec:.2e}, dC/drho max rel err {eg:.2e} over "
              f"{g0.size} elements", flush=True)
        summary.append(f"{name} C {ec:.1e} grad {eg:.1e}")
        assert r["rel_res"] <= 1.5e-10
        assert ec <= 1e-9, (name, ec)
        assert eg <= 1e-7, (name, eg)
    print("  C2 vs oracle: " + "; ".join(summary), flush=True)
    ctx.close()


def _element_dK(oracle, Xe, he, d):
    """K_e(h + d) - K_e(h - d) on the element's u v w tx ty (var-major) rows,
    without the rounding of subtracting two nearly equal matrices: with no
    offset the Mindlin element matrix is a h + b h^3 (membrane and shear ~ h,
    bending ~ h^3), so a and b come from the oracle's element routine at
    h = 1 and 2, the cubic form is checked at h = 3, and the difference is
    2 d a + (6 h^2 d + 2 d^3) b."""
    K = [oracle.quad4_mindlin(Xe, 70e9, 0.3, t, 0.0, 5.0 / 6.0)[1][:20, :20] for t in (1.0, 2.0, 3.0)]
    b = (K[1] - 2.0 * K[0]) / 6.0
    a = K[0] - b
    assert np.max(np.abs(3.0 * a + 27.0 * b - K[2])) <= 1e-13 * np.abs(K[2]).max()
    return 2.0 * d * a + (6.0 * he * he * d + 2.0 * d ** 3) * b


def _fd_identity(oracle, xyz, conn, base, h, e, d, up, um):
    dofs = np.array([base[conn[e, a]] + v for v in range(5) for a in range(4)])
    dK = _element_dK(oracle, xyz[conn[e]], h[e], d)
    return -float(up[dofs] @ dK @ um[dofs]) / (2 * d)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("n", [32, 64])
def test_c3_adjoint_vs_central_fd_exact_solves(oracle, mastg, n):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    print(f"[FD, exact solves] Mindlin plate {n}x{n}: 64 seeded elements, step 1e-7 h_e", flush=True)
    P = oracle.Problem(oracle.PLATE, n, n, seed=1)
    ctx = ctx_from_oracle(mastg, P)
    h = P.params()
    xyz, conn = P.mesh()
    base, _ = P.dofmap()
    rp, col = P.csr()
    r = ctx.design_eval(h, solver=mastg.SOLVER_MG, rtol=1e-9)
    g = r["grad"]

    def exact_state(hh):
        P.set_params(hh)
        R, K = P.assemble()
        A = sp.csr_matrix((K, col, rp), shape=(P.n_dofs, P.n_dofs)).tocsc()
        lu = spl.splu(A)
        u = lu.solve(-R)
        u += lu.solve(-R - A @ u)            # one step of refinement
        return u

    u0 = exact_state(h)
    g0 = -P.sens_dots(u0, u0)                # the oracle's adjoint at the exact state
    worst = worst_oracle = 0.0
    for e in np.sort(np.random.default_rng(2024).choice(P.n_elems, 64, replace=False)):
        d = 1e-7 * h[e]
        hp = h.copy(); hp[e] += d
        hm = h.copy(); hm[e] -= d
        fd = _fd_identity(oracle, xyz, conn, base, h, e, d, exact_state(hp), exact_state(hm))
        worst = max(worst, abs(fd - g[e]) / abs(g[e]))
        worst_oracle = max(worst_oracle, abs(fd - g0[e]) / abs(g0[e]))
    P.set_params(h)
    print(f"  64 elements: max rel err vs central FD of exact solves: GPU adjoint {worst:.2e}, "
          f"oracle adjoint {worst_oracle:.2e}", flush=True)
    assert worst_oracle <= 1e-7, worst_oracle
    assert worst <= 1e-5, worst
    ctx.close()


@pytest.mark.timeout(900)
def test_c3_adjoint_vs_central_fd_64_elements(oracle, mastg):
    n = 512
    print(f"[full size vs FD] C3 Mindlin plate {n}x{n}: 64 seeded elements, step 1e-7 h_e", flush=True)
    P = oracle.Problem(oracle.PLATE, n, n, seed=1)
    ctx = ctx_from_oracle(mastg, P)
    h = P.params()
    xyz, conn = P.mesh()
    base, _ = P.dofmap()
    rtol = 1e-7
    r = ctx.design_eval(h, solver=mastg.SOLVER_MG, rtol=rtol)
    n_it = int(r["iters"])
    elems = np.sort(np.random.default_rng(2024).choice(P.n_elems, 64, replace=False))
    R0, _ = ctx.assemble(want_J=False)
    F = -R0                                  # R(0) = -F: the pressure load

    def state(hh):
        # the same number of multigrid iterations for every design (no
        # tolerance stop: the iterate is a smooth function of the design)
        ctx.set_params(hh)
        R, _ = ctx.assemble(want_J=True)
        u, st = ctx.solve(-R, solver=mastg.SOLVER_MG, rtol=1e-30, max_iters=n_it, raise_on_error=False)
        assert st["rc"] in (0, -3), st
        counts.add(st["iters"])
        assert len(counts) == 1, counts        # every state after the same iterations
        return u, float(np.dot(F, u))

    counts = set()

    u0, _ = state(h)
    g = -ctx.sensitivity(u0, u0)              # dC/dh = -u . dR/dh (self-adjoint)
    e_eval = float(np.max(np.abs(r["grad"] - g)) / np.abs(g).max())
    print(f"  {next(iter(counts))} iterations; design evaluation gradient vs the sensitivity pass at the fixed-iteration "
          f"state: {e_eval:.2e} of max |dC/dh|", flush=True)
    assert e_eval <= 1e-3
    worst = 0.0
    worst_naive = 0.0
    for k, e in enumerate(elems):
        d = 1e-7 * h[e]
        hp = h.copy(); hp[e] += d
        hm = h.copy(); hm[e] -= d
        up, Cp = state(hp)
        um, Cm = state(hm)
        fd = _fd_identity(oracle, xyz, conn, base, h, e, d, up, um)
        naive = (Cp - Cm) / (2 * d)
        err = abs(fd - g[e]) / abs(g[e])
        worst = max(worst, err)
        worst_naive = max(worst_naive, abs(naive - g[e]) / abs(g[e]))
        if k % 16 == 0:
            print(f"  element {e}: adjoint {g[e]:.10e}, central FD {fd:.10e} (rel err {err:.2e}); "
                  f"naive C+ - C- quotient {naive:.3e}", flush=True)
    print(f"  64 elements: max rel err adjoint vs central FD {worst:.2e} "
          f"(naive subtraction of compliances: {worst_naive:.2e})", flush=True)
    assert worst <= 1e-5, worst
    ctx.close()
