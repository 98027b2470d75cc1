"""Oracle (test infrastructure only): the von Mises stress outputs.

Only tests/ use this module, as the checker.  Restated in numpy from
StressStrainOutputBase::Data::von_Mises_stress (src/elasticity/
stress_output_base.cpp:183-195), KSStressStrainOutput::
functional_for_all_elems (src/elasticity/ks_stress_output.cpp:53-93),
StressStrainOutputBase::functional_for_all_elems (the weighted p-norm,
stress_output_base.cpp:852-908) and SmoothRampStressStrainOutput::
functional_for_all_elems (src/elasticity/smooth_ramp_stress_output.cpp:
36-87), over the quadrature-point stresses of the oracle's hex8 element
(StructuralElement3D::calculate_stress, solid_element_3d.cpp:1147-1236).
"""
import numpy as np

import pyoracle as O


def von_mises(s):
    """s: [..., 6] (xx yy zz xy yz zx) -> von Mises stress."""
    return np.sqrt(0.5 * ((s[..., 0] - s[..., 1]) ** 2 + (s[..., 1] - s[..., 2]) ** 2 +
                          (s[..., 2] - s[..., 0]) ** 2) + 3.0 * (s[..., 3] ** 2 + s[..., 4] ** 2 + s[..., 5] ** 2))


def element_stresses(P, u, E=72e9, nu=0.33):
    """Stresses [ne, 8, 6] and JxW [ne, 8] of every hex8 at the state u
    (caller numbering of the oracle problem), E_e = rho_e^p E."""
    xyz, conn = P.mesh()
    base, _ = P.dofmap()
    rho = P.params()
    sig = np.zeros((P.n_elems, 8, 6))
    jxw = np.zeros((P.n_elems, 8))
    for e in range(P.n_elems):
        dofs = np.array([base[conn[e, a]] + v for v in range(3) for a in range(8)])
        sig[e], jxw[e] = O.hex8_stress(xyz[conn[e]], rho[e] ** 3 * E, nu, u[dofs])
    return sig, jxw


def ks(svm, jxw, k, sref):
    """(1/k) ln( sum exp(k (svm - sref) / sref) JxW / sum JxW )."""
    I = np.sum(np.exp(k * (svm - sref) / sref) * jxw)
    return np.log(I / np.sum(jxw)) / k


def pnorm(svm, jxw, p, pw, rho, s0, lim=100.0):
    """s0 (sum (svm/s0)^p w JxW / sum w JxW)^(1/p), w = exp(rho ((svm-s0)/s0)^pw)
    with the exponent clamped at lim (stress_output_base.cpp:885-907)."""
    sp = ((svm - s0) / s0) ** pw
    w = np.where(rho * sp > lim, np.exp(lim), np.exp(np.minimum(rho * sp, lim)))
    I = np.sum((svm / s0) ** p * w * jxw)
    V = np.sum(w * jxw)
    return s0 * (I / V) ** (1.0 / p)


def ramp(svm, jxw, p, s0):
    """sum ((1 + (svm/s0)^p)^(1/p) - 1) JxW, summed as I - V
    (smooth_ramp_stress_output.cpp:73-86)."""
    I = np.sum((1.0 + (svm / s0) ** p) ** (1.0 / p) * jxw)
    return I - np.sum(jxw)
