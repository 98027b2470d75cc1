"""Pins of the oracle's stress restatement: the quadrature-point stresses of
a hex8 under a linear displacement field are the exact constant D eps (the
patch test of calculate_stress), the von Mises value of a uniaxial stress is
its magnitude and of pure shear sqrt(3) tau, and the KS aggregate lies
between the largest and the volume-averaged exponential mean (it tends to the
maximum as k grows).  The weighted p-norm of a uniform field is that stress
for every p and weight, it is a weighted power mean (between the smallest
and the largest value, nondecreasing in p, the maximum as p grows, the plain
volume p-mean without weight), and its clamped weights equal e^100; the
smooth ramp of a uniform field is V ((1 + (s/s0)^p)^(1/p) - 1), it tends to
the volume integral of max(s/s0 - 1, 0) as p grows and exceeds it for every
p."""
import numpy as np

import pyoracle as O
import stress as S


def test_linear_field_gives_exact_constant_stress():
    rng = np.random.default_rng(2)
    X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    X = X * [0.3, 0.4, 0.5] + rng.uniform(-0.03, 0.03, (8, 3))
    G = rng.standard_normal((3, 3)) * 1e-3           # u = G x
    u = np.concatenate([X @ G[v] for v in range(3)])  # var-major
    eps = np.array([G[0, 0], G[1, 1], G[2, 2], G[0, 1] + G[1, 0], G[1, 2] + G[2, 1], G[2, 0] + G[0, 2]])
    D = O.material_D3(72e9, 0.33)
    sig, jxw = O.hex8_stress(X, 72e9, 0.33, u)
    assert np.allclose(sig, (D @ eps)[None, :], rtol=1e-10, atol=1e-6 * np.abs(D @ eps).max())
    assert np.all(jxw > 0)


def test_von_mises_special_cases():
    assert abs(S.von_mises(np.array([5.0, 0, 0, 0, 0, 0])) - 5.0) <= 1e-15
    assert abs(S.von_mises(np.array([0, 0, 0, 2.0, 0, 0])) - 2.0 * np.sqrt(3.0)) <= 1e-14
    assert abs(S.von_mises(np.array([7.0, 7.0, 7.0, 0, 0, 0]))) <= 1e-14   # hydrostatic


def test_ks_bounds():
    rng = np.random.default_rng(5)
    svm = rng.uniform(0.1, 1.0, 200)
    jxw = rng.uniform(0.5, 1.5, 200)
    for k in (1.0, 10.0, 100.0):
        v = S.ks(svm, jxw, k, 1.0)
        assert v <= (svm.max() - 1.0) + 1e-12
        mean = np.sum(jxw * (svm - 1.0)) / jxw.sum()
        assert v >= mean - 1e-12                 # Jensen
    assert abs(S.ks(svm, jxw, 2000.0, 1.0) - (svm.max() - 1.0)) <= 5e-3 * abs(svm.max() - 1.0) + 5e-3


def test_pnorm_pins():
    rng = np.random.default_rng(8)
    jxw = rng.uniform(0.5, 1.5, 300)
    # uniform field: the aggregate is the stress itself, whatever the weights
    for p, pw, rho in ((2.0, 1.0, 1.0), (8.0, 1.0, 5.0), (3.0, 2.0, 0.0)):
        assert abs(S.pnorm(np.full(300, 1.7e8), jxw, p, pw, rho, 2e8) - 1.7e8) <= 1e-14 * 1.7e8
    svm = rng.uniform(0.2, 1.0, 300) * 2e8
    prev = 0.0
    for p in (1.0, 2.0, 4.0, 8.0, 16.0):
        v = S.pnorm(svm, jxw, p, 1.0, 1.0, 2e8)
        assert svm.min() - 1e-6 <= v <= svm.max() + 1e-6
        assert v >= prev - 1e-6
        prev = v
    assert abs(S.pnorm(svm, jxw, 400.0, 1.0, 1.0, 2e8) - svm.max()) <= 2e-2 * svm.max()
    # no weight: the volume-weighted p-mean, by brute force
    v = S.pnorm(svm, jxw, 3.0, 1.0, 0.0, 2e8)
    assert abs(v - (np.sum(svm ** 3 * jxw) / jxw.sum()) ** (1 / 3)) <= 1e-13 * v
    # clamped weights: rho ((s - s0)/s0) > 100 everywhere -> w = e^100 for all,
    # so the weights cancel and the plain p-mean remains
    big = rng.uniform(3.0, 4.0, 300) * 2e8
    v = S.pnorm(big, jxw, 2.0, 1.0, 60.0, 2e8)
    assert abs(v - np.sqrt(np.sum(big ** 2 * jxw) / jxw.sum())) <= 1e-13 * v


def test_ramp_pins():
    rng = np.random.default_rng(9)
    jxw = rng.uniform(0.5, 1.5, 300)
    V = jxw.sum()
    for s, p in ((1.5, 4.0), (0.5, 8.0), (1.0, 2.0)):
        ref = V * ((1 + s ** p) ** (1 / p) - 1)
        assert abs(S.ramp(np.full(300, s * 2e8), jxw, p, 2e8) - ref) <= 1e-12 * max(abs(ref), V)
    svm = rng.uniform(0.2, 2.0, 300) * 2e8
    hinge = np.sum(np.maximum(svm / 2e8 - 1.0, 0.0) * jxw)
    for p in (2.0, 8.0, 32.0):
        assert S.ramp(svm, jxw, p, 2e8) >= hinge - 1e-12
    assert abs(S.ramp(svm, jxw, 400.0, 2e8) - hinge) <= 1e-2 * V
