"""Hex8 side pressure (mastg_problem_desc.pres_bid_mask with `pressure`;
StructuralElement3D::surface_pressure_residual, solid_element_3d.cpp:815-876)
through the C ABI against the oracle, on the cantilever block and on a copy
with displaced interior and boundary nodes (warped pressure faces).

Bars: the residual at zero state (= minus the load) within 1e-12 of its
largest entry; compliance 1e-9 and dC/drho 1e-7 after solves to 1e-10 (plain
and multigrid on the block); the pressure sensitivity lambda . dR/dp against
the oracle's residual difference R(p + d) - R(p - d) (R is linear in p)
within 1e-10.
"""
import numpy as np
import pytest

from helpers import rel_err_floor

pytestmark = pytest.mark.gpu

P_SIDES = 0x24          # boundary ids 2 (x = 1 face) and 5 (z = top face)
PRESS = 2.5e5


def _setup(oracle, mastg, mesh, distort):
    P0 = oracle.Problem(oracle.CANTILEVER, *mesh, seed=5, lo=0.3, hi=1.0)
    xyz, conn = P0.mesh()
    if distort:
        # every node but those on the clamped face moves: the pressure faces warp
        rng = np.random.default_rng(9)
        h = 1.0 / mesh[0]
        move = xyz[:, 0] > 1e-12
        xyz = xyz.copy()
        xyz[move] += distort * h * rng.uniform(-1, 1, (int(move.sum()), 3))
    base, con = P0.dofmap()
    sides = P0.bsides()
    P = oracle.Problem(oracle.CANTILEVER, arrays=dict(xyz=xyz, conn=conn, n_vars=P0.n_vars, node_base=base,
                                                      constrained=con, params=P0.params(),
                                                      point_load=P0.point_load(), sides=sides))
    P.set(E=72e9, nu=0.33, penalty=3.0, pressure=PRESS, pres_mask=P_SIDES)
    desc = mastg.desc_from_arrays(mastg.HEX8_ELASTIC, xyz, conn, P0.n_vars, node_dof_base=base,
                                  constrained_dofs=np.nonzero(con)[0], sides=sides, point_loads=P0.point_load(),
                                  E=72e9, nu=0.33, penalty=3.0, pressure=PRESS, pres_bid_mask=P_SIDES)
    ctx = mastg.Context(desc)
    ctx.set_params(P0.params())
    return P, ctx, con


@pytest.mark.parametrize("mesh,distort", [((8, 4, 4), 0.0), ((8, 4, 4), 0.2), ((16, 8, 8), 0.0)])
def test_hex8_pressure(oracle, mastg, mesh, distort):
    print(f"[hex8 pressure] mesh={mesh} distortion={distort}", flush=True)
    P, ctx, con = _setup(oracle, mastg, mesh, distort)
    R0, _ = P.assemble()
    R, _ = ctx.assemble(want_J=False)
    err = np.max(np.abs(R - R0)) / np.abs(R0).max()
    print(f"  load vector rel err {err:.1e}", flush=True)
    assert err <= 1e-12
    ref = P.solve_elastic(1e-12)
    solvers = [mastg.SOLVER_CG_JACOBI] + ([mastg.SOLVER_MG] if distort == 0.0 else [])
    for sv in solvers:
        r = ctx.design_eval(P.params(), solver=sv)
        ec = abs(r["objective"] - ref["compliance"]) / abs(ref["compliance"])
        eg = rel_err_floor(r["grad"], ref["dc"], floor=1e-6)
        print(f"  solver {sv}: C rel err {ec:.1e}, grad {eg:.1e}, {r['iters']} iterations", flush=True)
        assert ec <= 1e-9 and eg <= 1e-7
    # lambda . dR/dp for the pressure against the oracle's residual difference
    lam = np.random.default_rng(4).standard_normal(P.n_dofs)
    lam[con == 1] = 0.0
    X = ref["u"]
    g = ctx.sensitivity_global(mastg.PARAM_PRESSURE, X, lam)
    d = 1e-3 * PRESS
    P.set(pressure=PRESS + d)
    Rp, _ = P.assemble(X)
    P.set(pressure=PRESS - d)
    Rm, _ = P.assemble(X)
    P.set(pressure=PRESS)
    fd = float(np.dot(lam, (Rp - Rm) / (2 * d)))
    print(f"  lambda.dR/dp {g:.12e} (oracle difference {fd:.12e})", flush=True)
    assert abs(g - fd) <= 1e-10 * abs(fd)
    ctx.close()


def test_pressure_refused_without_sides(oracle, mastg):
    P = oracle.Problem(oracle.CANTILEVER, 4, 2, 2, seed=1, lo=0.3, hi=1.0)
    xyz, conn = P.mesh()
    base, con = P.dofmap()
    desc = mastg.desc_from_arrays(mastg.HEX8_ELASTIC, xyz, conn, P.n_vars, node_dof_base=base,
                                  constrained_dofs=np.nonzero(con)[0], E=72e9, nu=0.33, penalty=3.0)
    ctx = mastg.Context(desc)
    ctx.set_params(P.params())
    with pytest.raises(mastg.MastgError):
        ctx.sensitivity_global(mastg.PARAM_PRESSURE, np.zeros(P.n_dofs), np.ones(P.n_dofs))
    ctx.close()
