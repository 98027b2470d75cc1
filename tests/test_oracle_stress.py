"""Pins of the oracle's stress restatement: the quadrature-point stresses of
a hex8 under a linear displacement field are the exact constant D eps (the
patch test of calculate_stress), the von Mises value of a uniaxial stress is
its magnitude and of pure shear sqrt(3) tau, and the KS aggregate lies
between the largest and the volume-averaged exponential mean (it tends to the
maximum as k grows)."""
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
