"""Pins of the oracle's design-space restatement (oracle/design.py) by what
the mathematics fixes: the filter's rows are a partition of unity (constants
pass unchanged), it is symmetric in its unnormalised weights, its support is
the radius ball, and a radius below the element spacing is the identity;
element volumes sum to the block's volume and equal |det A| for an affine
(parallelepiped) element; the OC step meets the volume target, respects its
move limits and bounds, and is monotone in the multiplier."""
import numpy as np
import pytest

import design as D


def _box_mesh(oracle, nx, ny, nz):
    P = oracle.Problem(oracle.CANTILEVER, nx, ny, nz, seed=1, lo=0.3, hi=1.0)
    return P.mesh()


def test_filter_partition_of_unity_and_support(oracle):
    xyz, conn = _box_mesh(oracle, 6, 4, 3)
    c = D.centroids(xyz, conn)
    h = 1.0 / 6
    W = D.filter_matrix(c, 1.5 * h)
    assert np.allclose(W.sum(axis=1), 1.0, rtol=0, atol=1e-15)
    d = np.sqrt(((c[:, None] - c[None]) ** 2).sum(-1))
    assert np.all(W[d > 1.5 * h] == 0.0)
    # an interior element at r = 1.5 h (spacings 1/6, 1/8, 1/6 in x, y, z here)
    # has every neighbour within the ball, and the unnormalised weights are symmetric
    S = 1.0 / W.sum(axis=1)
    A = W / W.sum(axis=1, keepdims=True)
    a = np.where(d <= 1.5 * h, 1.5 * h - d, 0.0)
    assert np.allclose(a, a.T)
    assert np.allclose(W, a / a.sum(axis=1, keepdims=True))
    # a radius below the smallest spacing keeps only the element itself
    assert np.allclose(D.filter_matrix(c, 0.05), np.eye(len(c)))


def test_volumes_box_and_affine(oracle):
    xyz, conn = _box_mesh(oracle, 5, 3, 2)
    V = D.hex8_volumes(xyz, conn)
    assert abs(V.sum() - 1.0 * 0.5 * 0.5) <= 1e-15 * 10
    assert np.allclose(V, (1 / 5) * (0.5 / 3) * (0.5 / 2), rtol=1e-14)
    # affine element: x = A xi-cube
    A = np.array([[1.0, 0.2, 0.1], [0.05, 0.8, 0.3], [0.0, 0.1, 1.2]])
    cube = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                     [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    X = cube @ A.T
    assert abs(D.hex8_volumes(X, np.arange(8)[None])[0] - abs(np.linalg.det(A))) <= 1e-14
    sq = np.array([[0, 0, 0], [2, 0, 0], [2, 3, 0], [0, 3, 0]], float)
    assert abs(D.quad4_areas(sq, np.arange(4)[None])[0] - 6.0) <= 1e-14


def test_oc_step_properties():
    rng = np.random.default_rng(1)
    n = 400
    rho = rng.uniform(0.1, 0.9, n)
    dv = rng.uniform(0.5, 1.5, n)
    dc = -rng.uniform(0.1, 10.0, n)
    vt = dv.sum()
    new, lam = D.oc_update(rho, dc, dv, vt, 0.4, 0.2, 1e-3)
    assert abs(np.dot(dv, new) - 0.4 * vt) <= 1e-9 * vt
    assert np.all(new >= np.maximum(1e-3, rho - 0.2) - 1e-15)
    assert np.all(new <= np.minimum(1.0, rho + 0.2) + 1e-15)
    # interior entries follow rho sqrt(B / lam)
    B = -dc / dv
    free = (new > np.maximum(1e-3, rho - 0.2)) & (new < np.minimum(1.0, rho + 0.2))
    assert np.allclose(new[free], rho[free] * np.sqrt(B[free] / lam), rtol=1e-13)
    # a larger multiplier never adds volume
    n2, _ = D.oc_update(rho, dc, dv, vt, 0.3, 0.2, 1e-3)
    assert np.all(n2 <= new + 1e-15)
