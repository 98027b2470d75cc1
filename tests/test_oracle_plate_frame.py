"""The oracle's plate local frame (GeomElem::_init_local_elem_2d,
src/mesh/geom_elem.cpp:460-511, and the StructuralElementBase transforms,
src/elasticity/structural_element_base.cpp:1568-1660) pinned by rotation
invariance: the 6-variable plate problem rotated rigidly in space (every
element then leaves the x-y plane and is computed in its local frame) has
the flat plate's compliance and gradient, its translations and in-plane
rotations are the flat ones rotated, and the pressure follows the elements'
normals.  The rotation about the normal (theta_z) carries only MAST's 1e-8
diagonal and is left out of the displacement comparison."""
import numpy as np

import pyoracle as O


def rotation(axis, angle):
    a = np.asarray(axis, float)
    a /= np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * K @ K


def test_rotated_plate_matches_flat():
    P0 = O.Problem(O.PLATE, 6, 5, seed=2, mast_exact=True)
    P0.set(pressure=1e3)
    xyz, conn = P0.mesh()
    base, con = P0.dofmap()
    ref = P0.solve_elastic(1e-12)
    for axis, angle in (([1, 2, 0.5], 0.7), ([0, 1, 0], np.pi / 2), ([1, 0, 0], -2.2)):
        R = rotation(axis, angle)
        xyzr = (xyz - 0.3) @ R.T + np.array([0.1, -0.2, 0.4])
        P = O.Problem(O.PLATE, arrays=dict(xyz=xyzr, conn=conn, n_vars=6, node_base=base, constrained=con,
                                           params=P0.params()))
        P.set(E=70e9, nu=0.3, kappa=5.0 / 6.0, off=0.0, pressure=1e3)
        r = P.solve_elastic(1e-12)
        assert abs(r["compliance"] - ref["compliance"]) <= 1e-12 * abs(ref["compliance"])
        assert np.abs(r["dc"] - ref["dc"]).max() <= 1e-10 * np.abs(ref["dc"]).max()
        n = R @ np.array([0.0, 0.0, 1.0])
        scale = np.abs(ref["u"]).max()
        for k in range(len(base)):
            t0, t1 = ref["u"][base[k]:base[k] + 3], r["u"][base[k]:base[k] + 3]
            r0, r1 = ref["u"][base[k] + 3:base[k] + 6], r["u"][base[k] + 3:base[k] + 6]
            assert np.abs(R @ t0 - t1).max() <= 1e-10 * scale
            assert np.abs(R @ r0 - (r1 - np.dot(r1, n) * n)).max() <= 1e-10 * scale
