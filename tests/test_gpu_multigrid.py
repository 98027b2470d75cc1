"""Multigrid-preconditioned CG (hex8 lattice meshes) on the GPU.

* answers at the FP64 paths' bars against the oracle (true ||b - A x||/||b||
  <= 1e-10 with the FP64 CSR, compliance 1e-9, dC/drho 1e-7);
* the iteration count equals (+-1) that of a scipy restatement of the same
  V-cycle on the oracle's matrix (tests/mg_restatement.py), which only holds
  when the Galerkin operators, transfers, constraint handling, Chebyshev
  coefficients and Gershgorin bounds all agree;
* graphs and direct launches give bitwise equal results; repeated solves are
  bitwise reproducible; meshes that are not lattices are refused;
* the Mindlin plate (5 variables, bilinear transfers, power-iteration
  Chebyshev bounds) reaches the same bars.
"""
import numpy as np
import pytest

from helpers import ctx_from_oracle, rel_err_floor
import mg_restatement as MR

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mesh,lo", [((16, 8, 8), 0.01), ((32, 16, 16), 0.01), ((20, 10, 10), 0.3)])
def test_mg_design_eval(oracle, mastg, mesh, lo):
    print(f"[mg design eval] mesh={mesh} lo={lo}", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, *mesh, seed=1, lo=lo, hi=1.0)
    ref = P.solve_elastic(1e-10)
    ctx = ctx_from_oracle(mastg, P)
    r = ctx.design_eval(P.params(), solver=mastg.SOLVER_MG)
    assert r["rel_res"] <= 1.5e-10
    assert abs(r["objective"] - ref["compliance"]) <= 1e-9 * abs(ref["compliance"])
    assert rel_err_floor(r["grad"], ref["dc"], floor=1e-6) <= 1e-7
    assert r["iters"] <= 60 and r["iters"] * 10 < ref["iters"], (r["iters"], ref["iters"])
    ctx.close()


@pytest.mark.parametrize("steps", [None, (3, 3), (2, 4)])
def test_mg_iterations_match_restatement(oracle, mastg, steps):
    """Default smoothing (6 steps on the problem level, 2 on coarse levels) and
    two settings through the options.  The problem level smooths in FP32 (the
    restatement in FP64): iteration counts agree to +-1."""
    print(f"[mg vs scipy restatement] 16x8x8 steps={steps}", flush=True)
    mesh = (16, 8, 8)
    P = oracle.Problem(oracle.CANTILEVER, *mesh, seed=3, lo=0.01, hi=1.0)
    levels, b = MR.build_levels(P, mesh, np.array([1.0 / 16, 0.5 / 8, 0.5 / 8]))
    assert [L["A"].shape[0] for L in levels] == [4131, 675, 135]
    it_ref, _ = MR.pcg_iterations(levels, b, k=steps or (6, 2))
    ctx = ctx_from_oracle(mastg, P)
    if steps:
        ctx.set_option(mastg.OPT_MG_STEPS_FINE, steps[0])
        ctx.set_option(mastg.OPT_MG_STEPS_COARSE, steps[1])
    r = ctx.design_eval(P.params(), solver=mastg.SOLVER_MG)
    assert abs(r["iters"] - it_ref) <= 1, (r["iters"], it_ref)
    ctx.close()


def test_mg_steps_option_range(oracle, mastg):
    P = oracle.Problem(oracle.CANTILEVER, 8, 4, 4, seed=3, lo=0.3, hi=1.0)
    ctx = ctx_from_oracle(mastg, P)
    for bad in (0, 9):
        with pytest.raises(mastg.MastgError):
            ctx.set_option(mastg.OPT_MG_STEPS_FINE, bad)
    for bad in (1, -3, 10001):
        with pytest.raises(mastg.MastgError):
            ctx.set_option(mastg.OPT_MG_CHEB_RATIO, bad)
    ctx.set_option(mastg.OPT_MG_CHEB_RATIO, 0)
    ctx.close()


def test_mg_solve_api_graphs_and_repeat(oracle, mastg):
    print("[mg solve api, graphs vs direct] 16x8x8", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, 16, 8, 8, seed=5, lo=0.01, hi=1.0)
    ctx = ctx_from_oracle(mastg, P)
    R, _ = ctx.assemble()
    x1, s1 = ctx.solve(-R, solver=mastg.SOLVER_MG)
    x2, s2 = ctx.solve(-R, solver=mastg.SOLVER_MG)
    ctx.set_option(mastg.OPT_GRAPHS, 0)
    x3, s3 = ctx.solve(-R, solver=mastg.SOLVER_MG)
    assert s1["rel_res"] <= 1.5e-10
    assert s1["iters"] == s2["iters"] == s3["iters"]
    assert np.array_equal(x1, x2) and np.array_equal(x1, x3)
    # a new assembly (new densities) re-forms the coarse operators
    p = P.params().copy()
    p[::7] *= 0.5
    ctx.set_option(mastg.OPT_GRAPHS, 1)
    r = ctx.design_eval(p, solver=mastg.SOLVER_MG)
    P.set_params(p)
    ref = P.solve_elastic(1e-10)
    assert abs(r["objective"] - ref["compliance"]) <= 1e-9 * ref["compliance"]
    ctx.close()


def test_mg_refuses_non_lattice(oracle, mastg):
    print("[mg refusals]", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, 7, 5, 3, seed=1, lo=0.3, hi=1.0)   # odd counts: no coarse level
    ctx = ctx_from_oracle(mastg, P)
    with pytest.raises(mastg.MastgError) as e:
        ctx.design_eval(P.params(), solver=mastg.SOLVER_MG)
    assert e.value.code == -1
    ctx.close()
    Ph = oracle.Problem(oracle.HEAT, 4, 4, 4, seed=1)     # 125 DOFs: nothing to coarsen
    ch = ctx_from_oracle(mastg, Ph)
    ch.assemble(np.zeros(Ph.n_dofs))
    with pytest.raises(mastg.MastgError) as e:
        ch.solve(np.ones(Ph.n_dofs), solver=mastg.SOLVER_MG)
    assert e.value.code == -1
    ch.close()


def test_mg_design_eval_matrix_free_state(oracle, mastg):
    """The hex8 multigrid design evaluation forms no assembled matrix: CSR
    products refuse until mastg_assemble; a later assembly is for the same
    parameters and the evaluation's results do not depend on it."""
    print("[mg matrix-free state] 16x8x8", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, 16, 8, 8, seed=6, lo=0.01, hi=1.0)
    ctx = ctx_from_oracle(mastg, P)
    r1 = ctx.design_eval(P.params(), solver=mastg.SOLVER_MG)
    x = np.random.default_rng(3).standard_normal(P.n_dofs)
    with pytest.raises(mastg.MastgError):
        ctx.jacobian_product(x, op=0)
    ctx.assemble()
    _, con = P.dofmap()
    xm = np.where(con == 1, 0.0, x)
    y0 = ctx.jacobian_product(xm, op=0)
    y3 = ctx.jacobian_product(x, op=3)
    assert np.max(np.abs(y3 - y0)) <= 1e-12 * np.max(np.abs(y0))
    r2 = ctx.design_eval(P.params(), solver=mastg.SOLVER_MG)
    assert r2["objective"] == r1["objective"] and r2["iters"] == r1["iters"]
    assert np.array_equal(r2["grad"], r1["grad"])
    ctx.close()


@pytest.mark.parametrize("mesh", [(8, 4, 4), (16, 8, 8)])
def test_lattice_operator_equals_csr(oracle, mastg, mesh):
    """The per-node lattice operator (multigrid level 0 and its CG product)
    applies the assembled matrix: equal to the CSR product of the same input
    with its constrained entries zeroed, within 1e-12 of max |y|."""
    print(f"[lattice operator] mesh={mesh}", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, *mesh, seed=4, lo=0.01, hi=1.0)
    ctx = ctx_from_oracle(mastg, P)
    ctx.assemble()
    _, con = P.dofmap()
    x = np.random.default_rng(7).standard_normal(P.n_dofs)
    xm = np.where(con == 1, 0.0, x)
    y0 = ctx.jacobian_product(xm, op=0)
    y3 = ctx.jacobian_product(x, op=3)
    assert np.max(np.abs(y3 - y0)) <= 1e-12 * np.max(np.abs(y0))
    ctx.close()


@pytest.mark.parametrize("n", [16, 24])
def test_mg_plate_design_eval(oracle, mastg, n):
    """Mindlin plate: bilinear transfers of all 5 variables, Galerkin 5x5-block
    coarse operators, Chebyshev bounds from power iteration; answers at the
    plate's FP64 bars and far fewer iterations than Jacobi.  The first call
    on the fresh context is the multigrid design evaluation, which builds the
    hierarchy before the first assembly: the problem level's arguments must
    hold the CSR values and inverse diagonal that the assembly then fills."""
    print(f"[mg plate] {n}x{n}", flush=True)
    P = oracle.Problem(oracle.PLATE, n, n, seed=1)
    ref = P.solve_elastic(1e-10)
    ctx = ctx_from_oracle(mastg, P)
    r = ctx.design_eval(P.params(), solver=mastg.SOLVER_MG)
    assert r["rel_res"] <= 1.5e-10
    assert abs(r["objective"] - ref["compliance"]) <= 1e-9 * abs(ref["compliance"])
    assert rel_err_floor(r["grad"], ref["dc"], floor=1e-6) <= 1e-7
    # thin plates converge slowly under point-Jacobi Chebyshev smoothing; the
    # default steps (4 on the problem level, 2 on coarse levels; fastest per
    # design at 512^2) need ~1/8 of Jacobi-CG's iterations at 16^2
    assert r["iters"] * 6 < ref["iters"], (r["iters"], ref["iters"])
    # deterministic: a second design evaluation is bitwise equal
    r2 = ctx.design_eval(P.params(), solver=mastg.SOLVER_MG)
    assert r2["objective"] == r["objective"] and np.array_equal(r2["grad"], r["grad"])
    ctx.close()


@pytest.mark.parametrize("n", [8, 16])
def test_mg_heat_design_eval(oracle, mastg, n):
    """Nonlinear heat: Newton steps and the adjoint solve with BiCGStab
    right-preconditioned by the multigrid V-cycle (scalar Galerkin levels,
    power-iteration Chebyshev bounds, rebuilt from every Newton Jacobian) give
    the Jacobi path's answers and need far fewer Krylov iterations."""
    print(f"[mg heat] {n}^3", flush=True)
    P = oracle.Problem(oracle.HEAT, n, seed=5)
    ref = P.solve_heat(1e-12, 1e-14)
    ctx = ctx_from_oracle(mastg, P)
    rj = ctx.design_eval(P.params(), rtol=1e-12)
    r = ctx.design_eval(P.params(), rtol=1e-12, solver=mastg.SOLVER_MG)
    assert 2 <= r["newton_iters"] <= 10
    assert abs(r["objective"] - ref["mean_T"]) <= 1e-9 * abs(ref["mean_T"])
    assert rel_err_floor(r["grad"], ref["dJ"], floor=1e-6) <= 1e-7
    assert r["iters"] * 3 < rj["iters"], (r["iters"], rj["iters"])
    ctx.close()
