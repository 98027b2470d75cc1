"""GPU sensitivities with respect to global parameters, through the C ABI.

NonlinearImplicitAssembly::sensitivity_assemble
(src/base/nonlinear_implicit_assembly.cpp:515-686) for a parameter every
element depends on (E, nu, the plate pressure, k0, the convection coefficient,
the emissivity), its adjoint dot (AssemblyBase, src/base/assembly_base.cpp:
594-639) and the forward sensitivity solve (NonlinearSystem::sensitivity_solve,
src/base/nonlinear_system.cpp:742-794).

Pins: dR/dp against central differences of the oracle's own residual in that
parameter at the same state (the residual is linear in E, the pressure, k0,
h and the emissivity, so the difference is exact to rounding; smooth in nu,
step 1e-6); dC/dE = -C/E and dX/dE = -X/E for linear elasticity (K is linear
in E); the forward solve's dX/dnu against central differences of the
oracle's solutions; forward-adjoint duality F . dX/dp = -lambda . dR/dp.
Bars: 1e-10 of the largest entry for the linear parameters, 1e-8 for nu,
1e-6 for differences of solutions.
"""
import numpy as np
import pytest

from helpers import ctx_from_oracle

pytestmark = pytest.mark.gpu

SB = 5.670374419e-8
RAD_ORACLE = dict(rad_sigma=SB, rad_emiss=0.8, rad_Tamb=250.0, rad_T0=20.0, rad_mask=0x30)
RAD_DESC = dict(rad_sigma=SB, rad_emissivity=0.8, rad_tamb=250.0, rad_tzero=20.0, rad_bid_mask=0x30)

# (oracle problem kind, mesh, parameter, oracle scalar name, value, relative step)
CASES = [
    ("hex8", (3, 2, 2), "E", "E", 72e9, 1e-3),
    ("hex8", (3, 2, 2), "NU", "nu", 0.33, 1e-6),
    ("plate", (3, 3), "E", "E", 70e9, 1e-3),
    ("plate", (3, 3), "NU", "nu", 0.3, 1e-6),
    ("plate", (3, 3), "PRESSURE", "pressure", 1e3, 1e-3),
    ("heat", (3,), "K0", "k0", 200.0, 1e-3),
    ("heat", (3,), "CONV_H", "conv_h", 50.0, 1e-3),
    ("heat", (3,), "EMISSIVITY", "rad_emiss", 0.8, 1e-3),
]


def _problem(oracle, kind, mesh, **scalars):
    K = {"hex8": oracle.CANTILEVER, "plate": oracle.PLATE, "heat": oracle.HEAT}[kind]
    P = oracle.Problem(K, *mesh, seed=4, lo=0.3, hi=1.0)
    if kind == "heat":
        P.set(**RAD_ORACLE)
    return P.set(**scalars) if scalars else P


@pytest.mark.parametrize("kind,mesh,param,name,value,step", CASES)
def test_dRdp_against_oracle_differences(oracle, mastg, kind, mesh, param, name, value, step):
    print(f"[global dR/dp] {kind} {mesh} {param}", flush=True)
    P = _problem(oracle, kind, mesh)
    extra = RAD_DESC if kind == "heat" else {}
    ctx = ctx_from_oracle(mastg, P, **extra)
    _, con = P.dofmap()
    rng = np.random.default_rng(7)
    X = rng.standard_normal(P.n_dofs) * (1e-4 if kind != "heat" else 100.0)
    if kind == "heat":
        X += 400.0
    X[con == 1] = 0.0
    g = ctx.sensitivity_rhs(getattr(mastg, "PARAM_" + param), X)
    d = step * value
    Rp, _ = _problem(oracle, kind, mesh, **{name: value + d}).assemble(X, want_K=False)
    Rm, _ = _problem(oracle, kind, mesh, **{name: value - d}).assemble(X, want_K=False)
    fd = (Rp - Rm) / (2 * d)
    act = ctx.active_map() >= 0
    err = np.max(np.abs(g[act] - fd[act])) / np.abs(fd[act]).max()
    print(f"  max |dR/dp - central difference| / max = {err:.2e}", flush=True)
    assert err <= (1e-8 if param == "NU" else 1e-10), err
    assert np.all(g[~act] == 0.0)
    # adjoint dot = lambda . dR/dp
    lam = rng.standard_normal(P.n_dofs)
    v = ctx.sensitivity_global(getattr(mastg, "PARAM_" + param), X, lam)
    assert abs(v - np.dot(lam, g)) <= 1e-12 * np.abs(lam).max() * np.abs(g).sum()
    ctx.close()


def test_compliance_modulus_closed_forms(oracle, mastg):
    print("[global sens] hex8 8x4x4: dC/dE = -C/E, dX/dE = -X/E, duality for nu", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, 8, 4, 4, seed=2, lo=0.3, hi=1.0)
    ctx = ctx_from_oracle(mastg, P)
    R, _ = ctx.assemble()
    u, _ = ctx.solve(-R, rtol=1e-13)
    C = ctx.output(mastg.OUTPUT_COMPLIANCE, u)
    dCdE = -ctx.sensitivity_global(mastg.PARAM_E, u, u)      # self-adjoint: lambda = u
    assert abs(dCdE + C / 72e9) <= 1e-10 * C / 72e9, (dCdE, -C / 72e9)
    dXdE, st = ctx.sensitivity_solve(mastg.PARAM_E, u, rtol=1e-13)
    assert np.max(np.abs(dXdE + u / 72e9)) <= 1e-9 * np.abs(u).max() / 72e9
    # nu: forward solve against central differences of the oracle's solutions,
    # and F . dX/dnu = -u . dR/dnu
    dXdn, _ = ctx.sensitivity_solve(mastg.PARAM_NU, u, rtol=1e-13)
    dn = 1e-5
    up = oracle.Problem(oracle.CANTILEVER, 8, 4, 4, seed=2, lo=0.3, hi=1.0).set(nu=0.33 + dn) \
        .solve_elastic(1e-13)["u"]
    um = oracle.Problem(oracle.CANTILEVER, 8, 4, 4, seed=2, lo=0.3, hi=1.0).set(nu=0.33 - dn) \
        .solve_elastic(1e-13)["u"]
    fd = (up - um) / (2 * dn)
    err = np.max(np.abs(dXdn - fd)) / np.abs(fd).max()
    print(f"  dC/dE rel err {abs(dCdE + C / 72e9) / (C / 72e9):.2e}; dX/dnu vs oracle central "
          f"difference {err:.2e}", flush=True)
    assert err <= 1e-6, err
    F = -R
    dCdn_adj = -ctx.sensitivity_global(mastg.PARAM_NU, u, u)
    assert abs(np.dot(F, dXdn) - dCdn_adj) <= 1e-9 * abs(dCdn_adj)
    ctx.close()


def test_heat_forward_sensitivity_k0(oracle, mastg):
    print("[global sens] heat 4^3 with radiation: dT/dk0 forward vs oracle differences", flush=True)
    P = _problem(oracle, "heat", (4,))
    ref = P.solve_heat(1e-13, 1e-14)
    ctx = ctx_from_oracle(mastg, P, **RAD_DESC)
    T = ref["T"]
    ctx.assemble(T)                       # the Jacobian at the converged state
    dT, _ = ctx.sensitivity_solve(mastg.PARAM_K0, T, rtol=1e-13)
    d = 1e-4 * 200.0
    Tp = _problem(oracle, "heat", (4,), k0=200.0 + d).solve_heat(1e-13, 1e-14)["T"]
    Tm = _problem(oracle, "heat", (4,), k0=200.0 - d).solve_heat(1e-13, 1e-14)["T"]
    fd = (Tp - Tm) / (2 * d)
    err = np.max(np.abs(dT - fd)) / np.abs(fd).max()
    print(f"  dT/dk0 vs oracle central difference {err:.2e}", flush=True)
    assert err <= 1e-6, err
    ctx.close()
