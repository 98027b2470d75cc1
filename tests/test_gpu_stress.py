"""GPU parity of the von Mises stress outputs (MASTG_OUTPUT_KS_STRESS,
_PNORM_STRESS, _RAMP_STRESS) and their adjoint sensitivities against the
oracle (oracle/stress.py over the oracle's quadrature-point stresses),
through the C ABI.

Pins: the value against the oracle's at the same state (1e-12); dKS/dX
against central differences of the oracle's KS along random directions;
Euler's identity for the degree-1 homogeneous von Mises stress,
dKS/dX . u = sum_e rho_e (dKS/drho_e|direct) / p; the direct density
derivative against central differences of the oracle's KS in rho_e at fixed
u; the adjoint total derivative dKS/drho_e = direct_e - lambda . dR/drho_e
(K lambda = dKS/dX) against central differences of the oracle's KS(u(rho))
through its own solves (1e-5, floor 1e-7 of the largest entry) (KSStressStrainOutput, src/elasticity/
ks_stress_output.cpp:53-243; AssemblyBase adjoint sensitivities,
src/base/assembly_base.cpp:594-639).
"""
import numpy as np
import pytest

import stress as S
from helpers import ctx_from_oracle

pytestmark = pytest.mark.gpu


def _ks_oracle(P, u, k, sref):
    sig, jxw = S.element_stresses(P, u)
    return S.ks(S.von_mises(sig), jxw, k, sref)


def test_ks_value_and_derivatives(oracle, mastg):
    print("[KS stress] hex8 6x3x3: value, dKS/dX, direct dKS/drho", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, 6, 3, 3, seed=3, lo=0.3, hi=1.0)
    ctx = ctx_from_oracle(mastg, P)
    ref = P.solve_elastic(1e-12)
    u = ref["u"]
    sig, jxw = S.element_stresses(P, u)
    sref = float(S.von_mises(sig).max())
    k = 8.0
    ctx.set_ks_stress(k, sref)
    v, dv = ctx.output(mastg.OUTPUT_KS_STRESS, u, want_grad=True)
    v0 = _ks_oracle(P, u, k, sref)
    print(f"  KS {v:.15g} (oracle {v0:.15g})", flush=True)
    assert abs(v - v0) <= 1e-12 * max(1.0, abs(v0))
    rng = np.random.default_rng(1)
    _, con = P.dofmap()
    for t in range(3):
        d = rng.standard_normal(P.n_dofs) * np.abs(u).max()
        d[con == 1] = 0.0
        h = 1e-6                      # truncation of the quotient ~3e-8 here (2.9e-6 at 1e-5)
        fd = (_ks_oracle(P, u + h * d, k, sref) - _ks_oracle(P, u - h * d, k, sref)) / (2 * h)
        err = abs(np.dot(dv, d) - fd) / abs(fd)
        print(f"  dKS/dX . d vs central difference: rel err {err:.2e}", flush=True)
        assert err <= 1e-6
    g_dir = ctx.output_sensitivity(mastg.OUTPUT_KS_STRESS, u)
    rho = P.params()
    assert abs(np.dot(dv, u) - np.dot(rho, g_dir) / 3.0) <= 1e-10 * abs(np.dot(dv, u))
    for e in (0, 17, P.n_elems - 1):
        dr = 1e-6 * rho[e]
        rp = rho.copy(); rp[e] += dr
        rm = rho.copy(); rm[e] -= dr
        P.set_params(rp); kp = _ks_oracle(P, u, k, sref)
        P.set_params(rm); km = _ks_oracle(P, u, k, sref)
        P.set_params(rho)
        fd = (kp - km) / (2 * dr)
        assert abs(g_dir[e] - fd) <= 1e-6 * abs(fd) + 1e-12 * np.abs(g_dir).max(), (e, g_dir[e], fd)
    # compliance and mean temperature have no direct density dependence
    assert np.all(ctx.output_sensitivity(mastg.OUTPUT_COMPLIANCE, u) == 0.0)
    ctx.close()


def test_ks_adjoint_total_derivative(oracle, mastg):
    print("[KS stress adjoint] hex8 8x4x4: total dKS/drho vs oracle solves", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, 8, 4, 4, seed=6, lo=0.3, hi=1.0)
    ctx = ctx_from_oracle(mastg, P)
    rho = P.params()
    R, _ = ctx.assemble()
    u, _ = ctx.solve(-R, rtol=1e-13)
    sig, jxw = S.element_stresses(P, u)
    sref = float(S.von_mises(sig).max())
    k = 8.0
    ctx.set_ks_stress(k, sref)
    _, dKdX = ctx.output(mastg.OUTPUT_KS_STRESS, u, want_grad=True)
    lam, _ = ctx.solve(dKdX, rtol=1e-13)               # K symmetric: adjoint = forward solve
    total = ctx.output_sensitivity(mastg.OUTPUT_KS_STRESS, u) - ctx.sensitivity(u, lam)
    for e in (3, 40, 101, P.n_elems - 1):
        dr = 1e-5 * rho[e]
        vals = []
        for s in (1, -1):
            rr = rho.copy(); rr[e] += s * dr
            P.set_params(rr)
            ue = P.solve_elastic(1e-13)["u"]
            vals.append(_ks_oracle(P, ue, k, sref))
        P.set_params(rho)
        fd = (vals[0] - vals[1]) / (2 * dr)
        err = abs(total[e] - fd) / abs(fd)
        print(f"  element {e}: adjoint {total[e]:.10e}, central FD {fd:.10e}, rel err {err:.2e}", flush=True)
        # each oracle solve stops at 1e-13: its error divided by the step
        # floors the quotient near 1e-9 absolute, which only the smallest
        # entries (1e-3 of the largest) notice
        assert abs(total[e] - fd) <= 1e-5 * abs(fd) + 1e-7 * np.abs(total).max()
    ctx.close()


def _agg_oracle(P, u, kind, p, pw, rho, s0):
    sig, jxw = S.element_stresses(P, u)
    svm = S.von_mises(sig)
    return S.pnorm(svm, jxw, p, pw, rho, s0) if kind == "pnorm" else S.ramp(svm, jxw, p, s0)


@pytest.mark.parametrize("kind,p,pw,rho,s0f", [("pnorm", 4.0, 1.0, 1.0, 1.0), ("pnorm", 6.0, 1.0, 200.0, 0.5),
                                              ("ramp", 4.0, 1.0, 1.0, 0.5)])
def test_pnorm_ramp_value_and_derivatives(oracle, mastg, kind, p, pw, rho, s0f):
    """The weighted p-norm (StressStrainOutputBase, stress_output_base.cpp:
    852-1176; rho 200 with s0 half the peak clamps the weight exponent at
    the highest points) and the smooth ramp (smooth_ramp_stress_output.cpp:
    36-241): value against the oracle's (1e-12), state derivative against
    central differences of the oracle's value along random directions, direct
    density derivative against central differences in rho_e at fixed u."""
    print(f"[{kind} stress] hex8 6x3x3 p={p} pw={pw} rho={rho} s0=x{s0f} peak", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, 6, 3, 3, seed=3, lo=0.3, hi=1.0)
    ctx = ctx_from_oracle(mastg, P)
    u = P.solve_elastic(1e-12)["u"]
    sig, _ = S.element_stresses(P, u)
    s0 = s0f * float(S.von_mises(sig).max())
    K = mastg.OUTPUT_PNORM_STRESS if kind == "pnorm" else mastg.OUTPUT_RAMP_STRESS
    with pytest.raises(mastg.MastgError):
        ctx.output(K, u)                             # the reference stress has no default
    ctx.set_stress_aggregation(p, pw, rho, s0)
    v, dv = ctx.output(K, u, want_grad=True)
    v0 = _agg_oracle(P, u, kind, p, pw, rho, s0)
    print(f"  value {v:.15g} (oracle {v0:.15g})", flush=True)
    assert abs(v - v0) <= 1e-12 * abs(v0)
    rng = np.random.default_rng(11)
    _, con = P.dofmap()
    for t in range(3):
        d = rng.standard_normal(P.n_dofs) * np.abs(u).max()
        d[con == 1] = 0.0
        h = 1e-6
        fd = (_agg_oracle(P, u + h * d, kind, p, pw, rho, s0) - _agg_oracle(P, u - h * d, kind, p, pw, rho, s0)) / (2 * h)
        err = abs(np.dot(dv, d) - fd) / abs(fd)
        print(f"  d/dX . d vs central difference: rel err {err:.2e}", flush=True)
        assert err <= 1e-6
    g_dir = ctx.output_sensitivity(K, u)
    rho_e = P.params()
    # the ramp is I - V with V the volume: a step of 1e-6 leaves the quotient
    # ~1e-6 of rounding from that cancellation; 1e-4 keeps truncation ~1e-8
    step = 1e-4 if kind == "ramp" else 1e-6
    for e in (0, 17, P.n_elems - 1):
        dr = step * rho_e[e]
        rp = rho_e.copy(); rp[e] += dr
        rm = rho_e.copy(); rm[e] -= dr
        P.set_params(rp); fp = _agg_oracle(P, u, kind, p, pw, rho, s0)
        P.set_params(rm); fm = _agg_oracle(P, u, kind, p, pw, rho, s0)
        P.set_params(rho_e)
        fd = (fp - fm) / (2 * dr)
        assert abs(g_dir[e] - fd) <= 1e-6 * abs(fd) + 1e-12 * np.abs(g_dir).max(), (e, g_dir[e], fd)
    ctx.close()


def test_pnorm_adjoint_total_derivative(oracle, mastg):
    """The stress constraint of example_6 (p-norm with weight exponent 1,
    examples/structural/example_6/example_6.cpp:515-587, 861-871): its total
    density derivative by the adjoint, direct_e - lambda . dR/drho_e with
    K lambda = dF/dX, against central differences of the oracle's value
    through its own solves."""
    print("[p-norm stress adjoint] hex8 8x4x4: total dF/drho vs oracle solves", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, 8, 4, 4, seed=6, lo=0.3, hi=1.0)
    ctx = ctx_from_oracle(mastg, P)
    rho = P.params()
    R, _ = ctx.assemble()
    u, _ = ctx.solve(-R, rtol=1e-13)
    sig, _ = S.element_stresses(P, u)
    s0 = 0.8 * float(S.von_mises(sig).max())
    p, pw, vr = 6.0, 1.0, 0.5
    ctx.set_stress_aggregation(p, pw, vr, s0)
    _, dFdX = ctx.output(mastg.OUTPUT_PNORM_STRESS, u, want_grad=True)
    lam, _ = ctx.solve(dFdX, rtol=1e-13)
    total = ctx.output_sensitivity(mastg.OUTPUT_PNORM_STRESS, u) - ctx.sensitivity(u, lam)
    for e in (3, 40, 101, P.n_elems - 1):
        dr = 1e-5 * rho[e]
        vals = []
        for sgn in (1, -1):
            rr = rho.copy(); rr[e] += sgn * dr
            P.set_params(rr)
            ue = P.solve_elastic(1e-13)["u"]
            vals.append(_agg_oracle(P, ue, "pnorm", p, pw, vr, s0))
        P.set_params(rho)
        fd = (vals[0] - vals[1]) / (2 * dr)
        err = abs(total[e] - fd) / abs(fd)
        print(f"  element {e}: adjoint {total[e]:.10e}, central FD {fd:.10e}, rel err {err:.2e}", flush=True)
        assert abs(total[e] - fd) <= 1e-5 * abs(fd) + 1e-7 * np.abs(total).max()
    ctx.close()
