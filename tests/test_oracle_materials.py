"""Oracle material / section matrices against the reference's own golden values.

Golden numbers are the hard-coded truths of the reference tests:
  tests/material/mast_isotropic_material.cpp:647-707 (3D D), :300-420 (2D D)
  tests/material/mast_isotropic_material_sensitivity.cpp:344-392 (2D dD/dE, dD/dnu)
  tests/property/mast_solid_2d_section_element_property_card.cpp:586-906 (A, B, D, S)
(E = 72e9, nu = 0.33, h = 0.06, off = 0.03, kappa = 5/6).
Derivatives the reference does not tabulate are pinned by closed forms (D is
linear in E) and central differences.
"""
import numpy as np
import pytest

E, NU, H, OFF, KAPPA = 72.0e9, 0.33, 0.06, 0.03, 5.0 / 6.0


def close(a, b, rtol=1e-13):
    np.testing.assert_allclose(a, b, rtol=rtol, atol=0)


def test_D3_golden(oracle):
    D = oracle.material_D3(E, NU)
    T = np.zeros((6, 6))
    for i in range(3):
        T[i, i] = 1.066784608580274e11
        T[i + 3, i + 3] = 0.541353383458647e11 / 2.0
    for i, j in [(0, 1), (0, 2), (1, 2)]:
        T[i, j] = T[j, i] = 0.525431225121628e11
    np.testing.assert_allclose(D, T, rtol=1e-14, atol=1.0)
    assert np.all(np.linalg.eigvalsh(D) > 0)


def test_D2_plane_stress_golden(oracle):
    D = oracle.material_D2(E, NU)
    T = np.array([[8.079901245651442e10, 2.666367411064976e10, 0],
                  [2.666367411064976e10, 8.079901245651442e10, 0],
                  [0, 0, 2.706766917293233e10]])
    np.testing.assert_allclose(D, T, rtol=1e-14, atol=1.0)


def test_dD2_goldens(oracle):
    dE = oracle.material_dD2(E, NU, 1.0, 0.0)
    TdE = np.array([[1.122208506340478, 0.370328807092358, 0],
                    [0.370328807092358, 1.122208506340478, 0], [0, 0, 0.375939849624060]])
    np.testing.assert_allclose(dE, TdE, rtol=1e-13, atol=1e-15)
    dnu = oracle.material_dD2(E, NU, 0.0, 1.0)
    Tdnu = np.array([[0.598444037945231e11, 1.005476657087070e11, 0],
                     [1.005476657087070e11, 0.598444037945231e11, 0],
                     [0, 0, -0.203516309570920e11]])
    np.testing.assert_allclose(dnu, Tdnu, rtol=1e-13, atol=1.0)


def test_dD3_closed_form_and_fd(oracle):
    # D is linear in E: dD/dE = D(E=1)
    close(oracle.material_dD3(E, NU, 1.0, 0.0), oracle.material_D3(1.0, NU))
    # dD/dnu against a central difference
    h = 1e-6
    fd = (oracle.material_D3(E, NU + h) - oracle.material_D3(E, NU - h)) / (2 * h)
    an = oracle.material_dD3(E, NU, 0.0, 1.0)
    np.testing.assert_allclose(an, fd, rtol=1e-7, atol=1e-7 * np.abs(an).max())


def test_section_goldens(oracle):
    A, B, Db, S = oracle.section(E, NU, H, OFF, KAPPA)
    np.testing.assert_allclose(
        A, [[4.847940747390865e9, 1.599820446638986e9, 0], [1.599820446638986e9, 4.847940747390865e9, 0],
            [0, 0, 1.624060150375940e9]], rtol=1e-14, atol=1e-3)
    np.testing.assert_allclose(
        B, [[1.454382224217259e8, 0.479946133991696e8, 0], [0.479946133991696e8, 1.454382224217259e8, 0],
            [0, 0, 0.487218045112782e8]], rtol=1e-14, atol=1e-3)
    np.testing.assert_allclose(
        Db, [[5.817528896869037e6, 1.919784535966782e6, 0], [1.919784535966782e6, 5.817528896869037e6, 0],
             [0, 0, 1.948872180451127e6]], rtol=1e-14, atol=1e-6)
    np.testing.assert_allclose(S, [[1.353383458646616e9, 0], [0, 1.353383458646616e9]], rtol=1e-14)


@pytest.mark.parametrize("off", [0.0, 0.03])
def test_section_thickness_derivatives_fd(oracle, off):
    """The plate element's d/dh of A, B, D_b, S (used by the sensitivity pass)
    against central differences of the section matrices."""
    h = 1e-7
    Ap, Bp, Dp, Sp = oracle.section(E, NU, H + h, off, KAPPA)
    Am, Bm, Dm, Sm = oracle.section(E, NU, H - h, off, KAPPA)
    Dps = oracle.material_D2(E, NU)
    G = E / 2 / (1 + NU)
    np.testing.assert_allclose((Ap - Am) / (2 * h), Dps, rtol=1e-6, atol=1e-6 * Dps.max())
    np.testing.assert_allclose((Bp - Bm) / (2 * h), Dps * off, rtol=1e-6, atol=1e-6 * Dps.max())
    np.testing.assert_allclose((Dp - Dm) / (2 * h), Dps * (H ** 2 / 4 + off ** 2), rtol=1e-6,
                               atol=1e-6 * Dps.max() * H ** 2)
    np.testing.assert_allclose((Sp - Sm) / (2 * h), np.eye(2) * G * KAPPA, rtol=1e-6, atol=1.0)
