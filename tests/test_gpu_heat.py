"""GPU parity of the hex8 steady-conduction path against the FP64 oracle (C ABI).

The element Jacobian includes the dk/dT term of k(T) = k0 (1 + alpha T)
(heat_conduction_elem_base.cpp:124-142), so J is non-symmetric: the linear and
adjoint solves use BiCGStab + Jacobi, the adjoint through the transposed values.
Bars: CSR pattern bit-exact; R and J at a temperature state within 1e-12
(J relative to max(|entry|, 1e-2 x row max)); Newton-converged mean temperature
within 1e-9 and d(mean T)/dq_e within 1e-7 of the oracle's.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from helpers import ctx_from_oracle, rel_err_floor

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 2, 3, 6])
def test_heat_csr_R_J_at_state(oracle, mastg, n):
    print(f"[heat K] {n}^3", flush=True)
    P = oracle.Problem(oracle.HEAT, n, seed=2)
    ctx = ctx_from_oracle(mastg, P)
    rp, col = ctx.csr_pattern()
    orp, ocol = P.csr()
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    _, con = P.dofmap()
    for T in (np.zeros(P.n_dofs), 300 + 50 * np.random.default_rng(n).standard_normal(P.n_dofs)):
        T[con == 1] = 0.0
        R0, K0 = P.assemble(T)
        R, J = ctx.assemble(T)
        assert rel_err_floor(J, K0, orp, floor=1e-2) <= 1e-12
        np.testing.assert_allclose(R, R0, rtol=0, atol=1e-12 * np.abs(R0).max())
    ctx.close()


def test_heat_transpose_solve(oracle, mastg):
    print("[heat A^T x = b] 5^3", flush=True)
    P = oracle.Problem(oracle.HEAT, 5, seed=3)
    ctx = ctx_from_oracle(mastg, P)
    _, con = P.dofmap()
    T = 200 + 100 * np.random.default_rng(1).random(P.n_dofs)
    T[con == 1] = 0
    R, J = ctx.assemble(T)
    rp, col = ctx.csr_pattern()
    A = sp.csr_matrix((J, col, rp), shape=(P.n_dofs, P.n_dofs))
    b = np.random.default_rng(2).standard_normal(P.n_dofs)
    b[con == 1] = 0
    for transpose in (False, True):
        x, st = ctx.solve(b, solver=mastg.SOLVER_BICGSTAB, rtol=1e-12, transpose=transpose)
        M = A.T if transpose else A
        res = np.linalg.norm(b - M @ x) / np.linalg.norm(b)
        assert res <= 2e-12 and abs(res - st["rel_res"]) <= 1e-3 * res + 1e-15
    ctx.close()


@pytest.mark.parametrize("n", [4, 8])
def test_heat_design_eval(oracle, mastg, n):
    print(f"[heat Newton + adjoint] {n}^3", flush=True)
    P = oracle.Problem(oracle.HEAT, n, seed=5)
    ref = P.solve_heat(1e-12, 1e-14)
    ctx = ctx_from_oracle(mastg, P)
    r = ctx.design_eval(P.params(), rtol=1e-12)
    assert 2 <= r["newton_iters"] <= 10
    assert abs(r["objective"] - ref["mean_T"]) <= 1e-9 * abs(ref["mean_T"])
    assert rel_err_floor(r["grad"], ref["dJ"], floor=1e-6) <= 1e-7
    # mean temperature output entry point at the oracle's state
    mt = ctx.output(mastg.OUTPUT_MEAN_T, ref["T"])
    assert abs(mt - ref["mean_T"]) <= 1e-12 * abs(ref["mean_T"])
    ctx.close()


# surface radiation on boundary ids 4 and 5 (x = 0 and z = 1; libMesh build_cube
# ids), on top of the convection of ids 1..5: HeatConductionElementBase::
# surface_radiation_residual (heat_conduction_elem_base.cpp:1402-1463)
SB = 5.670374419e-8
RAD_ORACLE = dict(rad_sigma=SB, rad_emiss=0.8, rad_Tamb=250.0, rad_T0=20.0, rad_mask=0x30)
RAD_DESC = dict(rad_sigma=SB, rad_emissivity=0.8, rad_tamb=250.0, rad_tzero=20.0, rad_bid_mask=0x30)


@pytest.mark.parametrize("n", [1, 2, 4])
def test_heat_radiation_R_J_at_state(oracle, mastg, n):
    print(f"[heat radiation K] {n}^3", flush=True)
    P = oracle.Problem(oracle.HEAT, n, seed=7).set(**RAD_ORACLE)
    ctx = ctx_from_oracle(mastg, P, **RAD_DESC)
    rp, _ = ctx.csr_pattern()
    _, con = P.dofmap()
    for T in (np.zeros(P.n_dofs), 600 + 200 * np.random.default_rng(n).standard_normal(P.n_dofs)):
        T[con == 1] = 0.0
        R0, K0 = P.assemble(T)
        R, J = ctx.assemble(T)
        assert rel_err_floor(J, K0, rp, floor=1e-2) <= 1e-12
        np.testing.assert_allclose(R, R0, rtol=0, atol=1e-12 * np.abs(R0).max())
    ctx.close()


@pytest.mark.parametrize("n,solver", [(4, "bicgstab"), (8, "mg")])
def test_heat_radiation_design_eval(oracle, mastg, n, solver):
    print(f"[heat radiation Newton + adjoint] {n}^3 {solver}", flush=True)
    P = oracle.Problem(oracle.HEAT, n, seed=5).set(**RAD_ORACLE)
    ref = P.solve_heat(1e-12, 1e-14)
    ctx = ctx_from_oracle(mastg, P, **RAD_DESC)
    sv = mastg.SOLVER_MG if solver == "mg" else mastg.SOLVER_CG_JACOBI
    r = ctx.design_eval(P.params(), rtol=1e-12, solver=sv)
    print(f"  mean T {r['objective']:.12g} (oracle {ref['mean_T']:.12g}), Newton {r['newton_iters']} "
          f"(oracle {ref['newton_iters']})", flush=True)
    assert 2 <= r["newton_iters"] <= 15
    assert abs(r["objective"] - ref["mean_T"]) <= 1e-9 * abs(ref["mean_T"])
    assert rel_err_floor(r["grad"], ref["dJ"], floor=1e-6) <= 1e-7
    # radiation moved the answer: the same problem without it differs
    P0 = oracle.Problem(oracle.HEAT, n, seed=5)
    assert abs(P0.solve_heat(1e-12, 1e-14)["mean_T"] - ref["mean_T"]) > 1e-3 * abs(ref["mean_T"])
    ctx.close()


@pytest.mark.parametrize("solver", ["bicgstab", "mg"])
def test_heat_distorted_mesh(oracle, mastg, solver):
    """Cells of many shapes (interior nodes displaced): the element kernel forms
    each element's own geometry instead of its congruence class's; R and J at
    a state and the Newton + adjoint design evaluation against the oracle."""
    print(f"[heat distorted] 6^3, {solver}", flush=True)
    P0 = oracle.Problem(oracle.HEAT, 6, seed=4)
    xyz, conn = P0.mesh()
    lo, hi = xyz.min(0), xyz.max(0)
    inner = np.all((xyz > lo + 1e-12) & (xyz < hi - 1e-12), axis=1)
    xyz = xyz.copy()
    xyz[inner] += 0.25 / 6 * np.random.default_rng(8).uniform(-1, 1, (int(inner.sum()), 3))
    base, con = P0.dofmap()
    sides = P0.bsides()
    P = oracle.Problem(oracle.HEAT, arrays=dict(xyz=xyz, conn=conn, n_vars=1, node_base=base, constrained=con,
                                                params=P0.params(), sides=sides))
    P.set(k0=200.0, alpha=1e-3, conv_h=50.0, conv_Tinf=300.0, conv_mask=0x3E)
    desc = mastg.desc_from_arrays(mastg.HEX8_HEAT, xyz, conn, 1, node_dof_base=base,
                                  constrained_dofs=np.nonzero(con)[0], sides=sides, k0=200.0, alpha=1e-3,
                                  conv_h=50.0, conv_tinf=300.0, conv_bid_mask=0x3E)
    ctx = mastg.Context(desc)
    ctx.set_params(P0.params())
    T = 300 + 50 * np.random.default_rng(3).standard_normal(P.n_dofs)
    T[con == 1] = 0.0
    R0, K0 = P.assemble(T)
    R, J = ctx.assemble(T)
    orp, _ = P.csr()
    assert rel_err_floor(J, K0, orp, floor=1e-2) <= 1e-12
    np.testing.assert_allclose(R, R0, rtol=0, atol=1e-12 * np.abs(R0).max())
    ref = P.solve_heat(1e-12, 1e-14)
    sv = mastg.SOLVER_MG if solver == "mg" else mastg.SOLVER_BICGSTAB
    r = ctx.design_eval(P0.params(), solver=sv, rtol=1e-12)
    ec = abs(r["objective"] - ref["mean_T"]) / abs(ref["mean_T"])
    eg = rel_err_floor(r["grad"], ref["dJ"], floor=1e-6)
    print(f"  mean T {ec:.1e}, grad {eg:.1e}, {r['newton_iters']} Newton steps", flush=True)
    assert ec <= 1e-9 and eg <= 1e-7
    ctx.close()
