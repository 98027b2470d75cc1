"""Oracle (test infrastructure only): the design-space steps of a SIMP loop.

Only tests/ use this module, as the checker.  Restated in numpy from:

* MAST::FilterBase (src/level_set/filter_base.cpp:35-142, 244-420): the
  filter map of every point holds the points within the radius with weight
  (r - d), normalised by their sum (FilterBase::_init, the all-pairs search);
  compute_filtered_values sums weight x input over the map.  The points here
  are the element centroids (one design variable per element).
* example_6's _evaluate_volume (examples/structural/example_6/example_6.cpp:
  714-791): V = sum_e V_e rho_e, dV/drho_e = V_e.
* the optimality-criteria update of a SIMP design (the standard step; MAST
  drives external optimisers): rho_new = clamp(rho sqrt(B / lambda),
  max(rho_min, rho - m), min(1, rho + m)), B = -dC/drho / dV/drho, lambda by
  bisection on V(rho_new) = vf V_total.
"""
import numpy as np


def centroids(xyz, conn):
    return xyz[conn].mean(axis=1)


def filter_matrix(c, r):
    """Dense filter matrix W (W[e, f] = (r - d_ef) / sum_f (r - d_ef) for d <= r):
    FilterBase::_init's all-pairs search."""
    d = np.sqrt(((c[:, None, :] - c[None, :, :]) ** 2).sum(-1))
    a = np.where(d <= r, r - d, 0.0)
    return a / a.sum(axis=1, keepdims=True)


def hex8_volumes(xyz, conn):
    """Element volumes by 2x2x2 Gauss on det J (exact for the trilinear map)."""
    g = 1.0 / np.sqrt(3.0)
    sx = np.array([-1, 1, 1, -1, -1, 1, 1, -1], float)
    sy = np.array([-1, -1, 1, 1, -1, -1, 1, 1], float)
    sz = np.array([-1, -1, -1, -1, 1, 1, 1, 1], float)
    X = xyz[conn]                                   # [ne, 8, 3]
    V = np.zeros(len(conn))
    for xi in (-g, g):
        for eta in (-g, g):
            for zeta in (-g, g):
                d = np.stack([0.125 * sx * (1 + sy * eta) * (1 + sz * zeta),
                              0.125 * sy * (1 + sx * xi) * (1 + sz * zeta),
                              0.125 * sz * (1 + sx * xi) * (1 + sy * eta)], axis=1)   # [8, 3]
                J = np.einsum("eni,nr->eir", X, d)
                V += np.linalg.det(J)
    return V


def quad4_areas(xyz, conn):
    g = 1.0 / np.sqrt(3.0)
    sx = np.array([-1, 1, 1, -1], float)
    sy = np.array([-1, -1, 1, 1], float)
    X = xyz[conn]
    A = np.zeros(len(conn))
    for xi in (-g, g):
        for eta in (-g, g):
            t0 = np.einsum("eni,n->ei", X, 0.25 * sx * (1 + sy * eta))
            t1 = np.einsum("eni,n->ei", X, 0.25 * sy * (1 + sx * xi))
            A += np.linalg.norm(np.cross(t0, t1), axis=1)
    return A


def oc_update(rho, dc, dv, vol_total, volfrac, move, rho_min, tol=1e-15):
    """Optimality-criteria step with the multiplier found by bisection (in
    log space, to relative width `tol`)."""
    B = np.maximum(0.0, -dc) / dv
    lo_b = np.maximum(rho_min, rho - move)
    hi_b = np.minimum(1.0, rho + move)

    def new(lam):
        return np.minimum(hi_b, np.maximum(lo_b, rho * np.sqrt(B / lam)))

    target = volfrac * vol_total
    l1, l2 = 1e-30, 1e30
    while np.log(l2 / l1) > tol:
        lm = np.sqrt(l1 * l2)
        if np.dot(dv, new(lm)) > target:
            l1 = lm
        else:
            l2 = lm
    return new(l2), l2
