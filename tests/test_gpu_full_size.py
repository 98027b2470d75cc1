"""Parity at the benchmark sizes through size-independent properties.

The oracle's FP64 solves take minutes at these sizes, so the GPU paths are
checked against the oracle on smaller meshes in the other modules, and here,
at the full configurations, against identities that hold at any size:

* C2 hex8 SIMP cantilever 128x64x64 (E_e = rho_e^3 E, no minimum density):
  K is homogeneous of degree 3 in the densities and the compliance is
  self-adjoint, so sum_e rho_e dC/drho_e = -3 C, C(2 rho) = C(rho) / 8 and
  dC/drho(2 rho) = dC/drho(rho) / 16; the gradient equals a central difference
  of C along a random direction; the multigrid and the plain FP64 path agree
  at the oracle bars; the multigrid solver's lattice operator is symmetric and
  equals the assembled CSR product.
* C3 Mindlin plate 512x512: the multigrid and the Jacobi-preconditioned CG
  agree on compliance and dC/dh; dC/dh equals a central difference of C along
  a random direction.
* C4 nonlinear heat 96^3: the multigrid- and Jacobi-preconditioned Newton +
  adjoint paths agree on mean T and its gradient; the gradient equals a
  central difference along a random direction.

A planted error in any stage (load vector, operator, solve, sensitivity
kernel, adjoint) breaks at least one of these.
"""
import numpy as np
import pytest

from helpers import rel_err_floor

pytestmark = pytest.mark.gpu


def constrained_mask(mastg, w):
    d = w.desc.contents
    m = np.zeros(d.n_nodes * d.n_vars, bool)
    m[np.ctypeslib.as_array(d.constrained_dofs, (d.n_constrained,))] = True
    return m


def test_c2_homogeneity_scaling_and_direction(mastg):
    print("[full size] C2 hex8 128x64x64: identities", flush=True)
    w = mastg.Workload(1, 128, 64, 64, seed=1, lo=0.01, hi=1.0)
    ctx = mastg.Context(w.desc)
    rho = w.params
    r = ctx.design_eval(rho, solver=mastg.SOLVER_MG, rtol=1e-11)
    C, g = r["objective"], r["grad"]
    assert r["rel_res"] <= 1.5e-11
    # degree-3 homogeneity: sum rho dC/drho = -3 C
    assert abs(np.dot(rho, g) + 3.0 * C) <= 1e-9 * abs(C), (np.dot(rho, g), -3 * C)
    # uniform scaling: K(2 rho) = 8 K(rho)
    r2 = ctx.design_eval(2.0 * rho, solver=mastg.SOLVER_MG, rtol=1e-11)
    assert abs(r2["objective"] * 8.0 - C) <= 1e-9 * abs(C)
    assert rel_err_floor(r2["grad"] * 16.0, g, floor=1e-6) <= 1e-7
    # central difference along a random direction (relative perturbations)
    d = rho * np.random.default_rng(11).uniform(-1.0, 1.0, rho.size)
    t = 1e-3
    cp = ctx.design_eval(rho + t * d, solver=mastg.SOLVER_MG, rtol=1e-11, want_grad=False)["objective"]
    cm = ctx.design_eval(rho - t * d, solver=mastg.SOLVER_MG, rtol=1e-11, want_grad=False)["objective"]
    fd = (cp - cm) / (2 * t)
    print(f"  sum rho g + 3C: {abs(np.dot(rho, g) + 3.0 * C) / abs(C):.2e} C; C(2 rho) 8 / C - 1: "
          f"{abs(r2['objective'] * 8.0 - C) / abs(C):.2e}; directional FD "
          f"{abs(fd - np.dot(g, d)) / abs(np.dot(g, d)):.2e}", flush=True)
    # g.d ~ C / 200 (random signs); solve noise ~1e-11 C / t and the O(t^2)
    # truncation stay well below 1e-4 of it
    assert abs(fd - np.dot(g, d)) <= 1e-4 * abs(np.dot(g, d)), (fd, np.dot(g, d))
    ctx.close()


def test_c2_multigrid_equals_plain_and_operator(mastg):
    print("[full size] C2 hex8 128x64x64: multigrid vs plain, lattice operator", flush=True)
    w = mastg.Workload(1, 128, 64, 64, seed=1, lo=0.01, hi=1.0)
    ctx = mastg.Context(w.desc)
    rm = ctx.design_eval(w.params, solver=mastg.SOLVER_MG)
    rp = ctx.design_eval(w.params, solver=mastg.SOLVER_CG_JACOBI)
    assert rm["rel_res"] <= 1.5e-10 and rp["rel_res"] <= 1.5e-10
    assert rm["iters"] * 50 < rp["iters"], (rm["iters"], rp["iters"])
    assert abs(rm["objective"] - rp["objective"]) <= 1e-9 * abs(rp["objective"])
    assert rel_err_floor(rm["grad"], rp["grad"], floor=1e-6) <= 1e-7
    # the lattice operator: symmetric, and equal to the assembled CSR product
    # (the plain design evaluation left the matrix of these densities assembled)
    con = constrained_mask(mastg, w)
    rng = np.random.default_rng(5)
    x = np.where(con, 0.0, rng.standard_normal(con.size))
    y = np.where(con, 0.0, rng.standard_normal(con.size))
    Ax, Ay = ctx.jacobian_product(x, op=3), ctx.jacobian_product(y, op=3)
    scale = np.linalg.norm(Ax) * np.linalg.norm(y)
    assert abs(np.dot(y, Ax) - np.dot(x, Ay)) <= 1e-12 * scale
    A0 = ctx.jacobian_product(x, op=0)
    print(f"  C rel diff {abs(rm['objective'] - rp['objective']) / abs(rp['objective']):.2e}, grad "
          f"{rel_err_floor(rm['grad'], rp['grad'], floor=1e-6):.2e}, symmetry "
          f"{abs(np.dot(y, Ax) - np.dot(x, Ay)) / scale:.2e}, lattice vs CSR "
          f"{np.max(np.abs(Ax - A0)) / np.max(np.abs(A0)):.2e}", flush=True)
    assert np.max(np.abs(Ax - A0)) <= 1e-12 * np.max(np.abs(A0))
    ctx.close()


def test_c3_plate_multigrid_equals_jacobi_and_direction(mastg):
    print("[full size] C3 Mindlin plate 512x512", flush=True)
    w = mastg.Workload(2, 512, 512, 0, seed=1)
    ctx = mastg.Context(w.desc)
    h = w.params
    rm = ctx.design_eval(h, solver=mastg.SOLVER_MG, rtol=1e-7)
    rj = ctx.design_eval(h, solver=mastg.SOLVER_CG_JACOBI, rtol=1e-7)
    assert rm["rel_res"] <= 1e-7 and rj["rel_res"] <= 1e-7
    assert rm["iters"] * 20 < rj["iters"], (rm["iters"], rj["iters"])
    C, g = rm["objective"], rm["grad"]
    assert abs(C - rj["objective"]) <= 1e-8 * abs(C)
    assert rel_err_floor(g, rj["grad"], floor=1e-6) <= 1e-6
    d = h * np.random.default_rng(12).uniform(-1.0, 1.0, h.size)
    t = 1e-3
    # the plate's FP64 floor of ||b - K u|| / ||b|| is ~2e-8 at 512^2: the
    # differenced solves ask the same 1e-7 as the evaluation
    cp = ctx.design_eval(h + t * d, solver=mastg.SOLVER_MG, rtol=1e-7, want_grad=False)["objective"]
    cm = ctx.design_eval(h - t * d, solver=mastg.SOLVER_MG, rtol=1e-7, want_grad=False)["objective"]
    fd = (cp - cm) / (2 * t)
    print(f"  C rel diff vs Jacobi {abs(C - rj['objective']) / abs(C):.2e}, "
          f"grad {rel_err_floor(g, rj['grad'], floor=1e-6):.2e}, "
          f"directional FD {abs(fd - np.dot(g, d)) / abs(np.dot(g, d)):.2e}", flush=True)
    assert abs(fd - np.dot(g, d)) <= 1e-4 * abs(np.dot(g, d)), (fd, np.dot(g, d))
    ctx.close()


def test_c4_heat_multigrid_equals_jacobi_and_direction(mastg):
    print("[full size] C4 heat 96^3", flush=True)
    w = mastg.Workload(3, 96, 96, 96, seed=1)
    ctx = mastg.Context(w.desc)
    q = w.params
    rm = ctx.design_eval(q, solver=mastg.SOLVER_MG)
    rj = ctx.design_eval(q)
    assert 2 <= rm["newton_iters"] <= 10
    assert rm["iters"] * 5 < rj["iters"], (rm["iters"], rj["iters"])
    J, g = rm["objective"], rm["grad"]
    assert abs(J - rj["objective"]) <= 1e-10 * abs(J)
    assert rel_err_floor(g, rj["grad"], floor=1e-6) <= 1e-8
    d = q * np.random.default_rng(13).uniform(-1.0, 1.0, q.size)
    t = 1e-3
    jp = ctx.design_eval(q + t * d, solver=mastg.SOLVER_MG, want_grad=False)["objective"]
    jm = ctx.design_eval(q - t * d, solver=mastg.SOLVER_MG, want_grad=False)["objective"]
    fd = (jp - jm) / (2 * t)
    print(f"  J rel diff vs Jacobi {abs(J - rj['objective']) / abs(J):.2e}, grad "
          f"{rel_err_floor(g, rj['grad'], floor=1e-6):.2e}, directional FD "
          f"{abs(fd - np.dot(g, d)) / abs(np.dot(g, d)):.2e}", flush=True)
    assert abs(fd - np.dot(g, d)) <= 1e-4 * abs(np.dot(g, d)), (fd, np.dot(g, d))
    ctx.close()
