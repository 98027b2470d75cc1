"""The oracle's hex8 side pressure (StructuralElement3D::surface_pressure_residual,
solid_element_3d.cpp:815-876: f -= int N p n dA, n outward) against what the
geometry fixes independently of its quadrature and of its orientation table:

* the total force on a side is -p times the side's vector area, and the vector
  area of a bilinear quadrilateral is half the cross product of its diagonals
  taken in libMesh's side node order (counter-clockwise seen from outside:
  side 0 = (0,3,2,1), 1 = (0,1,5,4), 2 = (1,2,6,5), 3 = (2,3,7,6),
  4 = (3,0,4,7), 5 = (4,5,6,7)) -- exact also for warped faces;
* on a flat parallelogram face every corner carries a quarter of the force,
  and nodes off the face carry none;
* the vector is linear in p.
"""
import numpy as np

import pyoracle as O

SIDE_NODES = [(0, 3, 2, 1), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7), (4, 5, 6, 7)]
CUBE = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)


def vector_area(X, side):
    a, b, c, d = (X[i] for i in SIDE_NODES[side])
    return 0.5 * np.cross(c - a, d - b)


def test_total_force_is_minus_p_times_vector_area():
    rng = np.random.default_rng(3)
    for trial in range(20):
        X = CUBE * rng.uniform(0.5, 2.0, 3) + 0.15 * rng.uniform(-1, 1, (8, 3))   # warped faces
        p = rng.uniform(-3e5, 3e5)
        for side in range(6):
            f = O.hex8_pressure(X, side, p).reshape(3, 8)
            tot = f.sum(axis=1)
            ref = -p * vector_area(X, side)
            assert np.allclose(tot, ref, rtol=1e-13, atol=1e-13 * abs(p)), (trial, side, tot, ref)
            off = [n for n in range(8) if n not in SIDE_NODES[side]]
            assert np.all(f[:, off] == 0.0)


def test_flat_face_quarters_and_linearity():
    X = CUBE * np.array([2.0, 0.5, 1.5])
    for side in range(6):
        f1 = O.hex8_pressure(X, side, 1.0).reshape(3, 8)
        area = vector_area(X, side)
        for n in SIDE_NODES[side]:
            assert np.allclose(f1[:, n], -0.25 * area, rtol=1e-14, atol=1e-15)
        assert np.allclose(O.hex8_pressure(X, side, -7.5), -7.5 * f1.reshape(-1), rtol=1e-14, atol=0)
