"""The oracle's surface radiation restatement
(HeatConductionElementBase::surface_radiation_residual,
src/heat_conduction/heat_conduction_elem_base.cpp:1402-1463), pinned by what
the mathematics fixes: a uniform face temperature gives the total flux
A sigma eps ((T + T0)^4 - (Ta + T0)^4) over the face area A (the shape
functions sum to one), no flux at T = Ta, only the face's four nodes carry
it, and the Jacobian is the derivative of the residual (central
differences)."""
import numpy as np

SB = 5.670374419e-8


def _cube(h=(1.0, 1.0, 1.0)):
    X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                  [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    return X * np.asarray(h)


# libMesh HEX8 sides: 0 z-, 1 y-, 2 x+, 3 y+, 4 x-, 5 z+
SIDE_NODES = [(0, 1, 2, 3), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (0, 3, 7, 4), (4, 5, 6, 7)]


def test_uniform_temperature_total_flux(oracle):
    hx, hy, hz = 0.3, 0.5, 0.7
    X = _cube((hx, hy, hz))
    area = {0: hx * hy, 5: hx * hy, 1: hx * hz, 3: hx * hz, 2: hy * hz, 4: hy * hz}
    for side in range(6):
        for T, Ta, T0, eps in [(350.0, 300.0, 0.0, 0.8), (20.0, -5.0, 273.15, 0.3)]:
            f, K = oracle.hex8_radiation(X, side, SB, eps, Ta, T0, np.full(8, T))
            exact = area[side] * SB * eps * ((T + T0) ** 4 - (Ta + T0) ** 4)
            assert abs(f.sum() - exact) <= 1e-12 * abs(exact)
            off = [a for a in range(8) if a not in SIDE_NODES[side]]
            assert np.all(f[off] == 0.0) and np.all(K[off] == 0.0) and np.all(K[:, off] == 0.0)
            # uniform face: each corner carries a quarter on a rectangle
            assert np.allclose(f[list(SIDE_NODES[side])], exact / 4, rtol=1e-12)
            # d(total flux)/dT = A sigma eps 4 (T + T0)^3 = sum of K
            assert abs(K.sum() - area[side] * SB * eps * 4 * (T + T0) ** 3) <= 1e-12 * K.sum()


def test_no_flux_at_ambient(oracle):
    X = _cube()
    f, _ = oracle.hex8_radiation(X, 2, SB, 0.9, 310.0, 0.0, np.full(8, 310.0))
    # the interpolated face temperature equals Ta to rounding
    assert np.all(np.abs(f) <= 1e-13 * SB * 0.9 * 310.0 ** 4)


def test_jacobian_is_derivative(oracle):
    rng = np.random.default_rng(3)
    X = _cube((0.2, 0.3, 0.4)) + rng.uniform(-0.02, 0.02, (8, 3))
    T = 300.0 + rng.uniform(-40, 40, 8)
    for side in range(6):
        f0, K = oracle.hex8_radiation(X, side, SB, 0.7, 280.0, 0.0, T)
        Kfd = np.zeros((8, 8))
        for b in range(8):
            d = np.zeros(8)
            d[b] = 1e-3
            fp, _ = oracle.hex8_radiation(X, side, SB, 0.7, 280.0, 0.0, T + d)
            fm, _ = oracle.hex8_radiation(X, side, SB, 0.7, 280.0, 0.0, T - d)
            Kfd[:, b] = (fp - fm) / 2e-3
        assert np.max(np.abs(K - Kfd)) <= 1e-8 * np.abs(K).max()
        assert np.allclose(K, K.T, rtol=0, atol=1e-14 * np.abs(K).max())


def test_linear_in_emissivity(oracle):
    X = _cube((0.5, 0.5, 0.5))
    T = np.linspace(290, 330, 8)
    f1, K1 = oracle.hex8_radiation(X, 4, SB, 0.25, 300.0, 0.0, T)
    f2, K2 = oracle.hex8_radiation(X, 4, SB, 0.75, 300.0, 0.0, T)
    assert np.allclose(f2, 3 * f1, rtol=1e-13, atol=0) and np.allclose(K2, 3 * K1, rtol=1e-13, atol=0)
