"""GPU parity of the Mindlin quad4 plate path against the FP64 oracle (C ABI).

Bars as for hex8: CSR pattern bit-exact (MAST's 6-variable DofMap restricted to
u v w tx ty, and the 5-variable layout); K within 1e-12 relative to
max(|entry|, 1e-2 x row max); residual and sensitivity dots within 1e-12;
compliance within 1e-9 and dC/dh within 1e-7 after CG to 1e-10; the adjoint
dC/dh against central differences of GPU solves (step 1e-4 h_e) within 1e-5.
"""
import numpy as np
import pytest

from helpers import ctx_from_oracle, rel_err_floor, restrict_csr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mesh", [(1, 1), (2, 2), (3, 2), (16, 16)])
@pytest.mark.parametrize("exact", [False, True])
def test_plate_csr_and_K(oracle, mastg, mesh, exact):
    print(f"[plate K] mesh={mesh} mast_exact={exact}", flush=True)
    P = oracle.Problem(oracle.PLATE, *mesh, seed=3, mast_exact=exact)
    ctx = ctx_from_oracle(mastg, P)
    rp, col = ctx.csr_pattern()
    orp, ocol = P.csr()
    R0, K0 = P.assemble()
    if exact:
        base, _ = P.dofmap()
        keep = np.sort(np.concatenate([base + v for v in range(5)]))
        orp, ocol, K0 = restrict_csr(orp, ocol, K0, keep)
        R0 = R0[keep]
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    R, J = ctx.assemble()
    act = ctx.active_map() >= 0
    assert rel_err_floor(J, K0, orp, floor=1e-2) <= 1e-12
    np.testing.assert_allclose(R[act], R0, rtol=0, atol=1e-12 * np.abs(R0).max())
    ctx.close()


def test_plate_residual_and_sensitivity_dots(oracle, mastg):
    print("[plate R(X), sens dots] 6x5", flush=True)
    P = oracle.Problem(oracle.PLATE, 6, 5, seed=8)
    ctx = ctx_from_oracle(mastg, P)
    rng = np.random.default_rng(2)
    _, con = P.dofmap()
    X = rng.standard_normal(P.n_dofs) * 1e-3
    X[con == 1] = 0
    R0, _ = P.assemble(X)
    R, _ = ctx.assemble(X)
    assert np.max(np.abs(R - R0)) <= 1e-12 * np.abs(R0).max()
    lam = rng.standard_normal(P.n_dofs)
    g0 = P.sens_dots(X, lam)
    g = ctx.sensitivity(X, lam)
    assert rel_err_floor(g, g0) <= 1e-12
    ctx.close()


@pytest.mark.parametrize("n", [8, 24])
def test_plate_design_eval(oracle, mastg, n):
    print(f"[plate design eval] {n}x{n}", flush=True)
    P = oracle.Problem(oracle.PLATE, n, n, seed=1)
    ref = P.solve_elastic(1e-10)
    ctx = ctx_from_oracle(mastg, P)
    r = ctx.design_eval(P.params())
    assert r["rel_res"] <= 1.5e-10
    assert abs(r["objective"] - ref["compliance"]) <= 1e-9 * abs(ref["compliance"])
    assert rel_err_floor(r["grad"], ref["dc"], floor=1e-6) <= 1e-7
    ctx.close()


def test_plate_mixed_refused_fp32_operator_accurate(oracle, mastg):
    """The mixed-precision solver is refused for plates (thin-plate bending
    ~h^3 against membrane ~h puts the Jacobian's conditioning beyond an FP32
    operator); the FP32 node-block operator itself (5x5 blocks) is still
    within FP32 rounding of the FP64 product."""
    print("[plate mixed refused, FP32 operator] 24x24", flush=True)
    P = oracle.Problem(oracle.PLATE, 24, 24, seed=1)
    ctx = ctx_from_oracle(mastg, P)
    with pytest.raises(mastg.MastgError) as e:
        ctx.design_eval(P.params(), solver=mastg.SOLVER_MIXED)
    assert e.value.code == -1
    ctx.assemble()
    x = np.random.default_rng(1).standard_normal(P.n_dofs)
    y64 = ctx.jacobian_product(x, op=0)
    y32 = ctx.jacobian_product(x, op=2)
    assert np.max(np.abs(y32 - y64)) <= 1e-6 * np.max(np.abs(y64))
    ctx.close()


def test_plate_node_block_fp64_operator(oracle, mastg):
    """The FP64 node-block product (the plate multigrid CG's operator: the
    CSR values read with one column index per 5x5 block) equals the oracle's
    CSR product and the GPU CSR product up to summation order, including the
    constrained rows (unit diagonal)."""
    print("[plate FP64 node-block operator] 20x16", flush=True)
    P = oracle.Problem(oracle.PLATE, 20, 16, seed=3)
    ctx = ctx_from_oracle(mastg, P)
    ctx.assemble()
    P.assemble()
    act = ctx.active_map() >= 0
    x = np.random.default_rng(4).standard_normal(P.n_dofs)
    y_csr = ctx.jacobian_product(x, op=0)
    y_nb = ctx.jacobian_product(x, op=4)
    # the oracle's own product (it keeps the drilling DOFs) on the active entries
    y_ref = P.spmv(np.where(act, x, 0.0))
    scale = np.max(np.abs(y_csr))
    print(f"  vs GPU CSR {np.max(np.abs(y_nb - y_csr)) / scale:.2e}, vs oracle "
          f"{np.max(np.abs(y_nb - y_ref)[act]) / scale:.2e}", flush=True)
    assert np.max(np.abs(y_nb - y_csr)) <= 1e-13 * scale
    assert np.max(np.abs(y_nb - y_ref)[act]) <= 1e-12 * scale
    with pytest.raises(mastg.MastgError):
        ctx.jacobian_product(x, op=5)
    ctx.close()


def test_plate_adjoint_vs_fd_on_gpu(mastg):
    print("[plate FD on GPU] 12x12", flush=True)
    # the FP64 floor of ||b - K u|| / ||b|| on this thin plate (cond ~2e9) is
    # ~1e-11 (a dense FP64 solve reaches 9.9e-12), so the solves ask 2e-11
    w = mastg.Workload(2, 12, 12, seed=5)
    ctx = mastg.Context(w.desc)
    h = w.params.copy()
    r = ctx.design_eval(h, rtol=2e-11)
    for e in (0, 17, 70, 143):
        st = 1e-4 * h[e]
        hp = h.copy(); hp[e] += st
        Cp = ctx.design_eval(hp, rtol=2e-11, want_grad=False)["objective"]
        hp[e] -= 2 * st
        Cm = ctx.design_eval(hp, rtol=2e-11, want_grad=False)["objective"]
        fd = (Cp - Cm) / (2 * st)
        assert abs(r["grad"][e] - fd) <= 1e-5 * abs(fd), (e, r["grad"][e], fd)
    ctx.close()
