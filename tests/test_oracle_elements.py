"""Oracle element kernels pinned by closed forms and invariants.

hex8 (StructuralElement3D): symmetric K with exactly the 6 rigid-body zero
eigenvalues; uniform-strain patch tests whose nodal forces are the exact
surface integrals of the constant stress; rigid rotations give no force;
invariance under a rotation of the element; MAST's 1e-20 rotational diagonal.
quad4 Mindlin: six rigid-body modes plus the two zero-energy modes of the 1-point
shear rule; the constant-curvature patch test under
reduced shear integration with the reference's section goldens; pressure load.
heat hex8: constant temperature gives no flux, linear temperature gives the
exact flux; the k(T) Jacobian term (dead code in MAST) against central
differences of the residual; convection totals.
"""
import numpy as np
import pytest

E, NU = 72.0e9, 0.33
UNIT_HEX = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                     [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
DISTORTED = UNIT_HEX * [1.3, 0.8, 1.1] + np.array(
    [[0, 0, 0], [0.1, 0.05, 0], [0.05, -0.1, 0.02], [-0.05, 0.02, 0.1],
     [0.02, 0.03, 0.05], [0, 0.1, -0.05], [0.1, 0.02, 0.03], [0.03, 0, 0.08]])


def rot_matrix(a, b, c):
    Rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    Ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    Rz = np.array([[np.cos(c), -np.sin(c), 0], [np.sin(c), np.cos(c), 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def field_to_u(X, fn, npe=8, nv=6):
    """nodal vector (var-major, 6 vars) of a displacement field fn(x) -> (3,)"""
    u = np.zeros(nv * npe)
    for a in range(npe):
        d = fn(X[a])
        for v in range(3):
            u[v * npe + a] = d[v]
    return u


@pytest.mark.parametrize("X", [UNIT_HEX, DISTORTED])
def test_hex8_symmetry_and_rigid_modes(oracle, X):
    _, K = oracle.hex8_elastic(X, E, NU)
    Kt = K[:24, :24]
    assert np.abs(Kt - Kt.T).max() <= 1e-12 * np.abs(Kt).max()
    ev = np.linalg.eigvalsh(Kt)
    scale = ev.max()
    assert np.sum(np.abs(ev) < 1e-10 * scale) == 6
    assert np.all(ev[6:] > 1e-6 * scale)
    # translations and infinitesimal rotations produce no force
    R = rot_matrix(0.3, -0.2, 0.5)
    for fn in (lambda x: np.array([1.0, -2.0, 0.5]),
               lambda x: np.cross(np.array([0.1, -0.3, 0.2]), x)):
        u = field_to_u(X, fn)
        f, _ = oracle.hex8_elastic(X, E, NU, u, jac=False)
        assert np.abs(f).max() <= 1e-9 * scale * np.abs(u).max()
    # MAST's small rotational diagonal: 1e-20 * max translational diagonal
    md = np.max(np.diag(Kt))
    np.testing.assert_allclose(np.diag(K)[24:], 1e-20 * md, rtol=1e-14)
    assert np.all(K[24:, :24] == 0) and np.all(K[:24, 24:] == 0)


def test_hex8_patch_uniform_strain(oracle):
    """u_x = x on the unit cube: sigma = D[:, 0] (uniform), nodal forces are
    int grad(N_a) . sigma dV = +-1/4 of the stress on the x/y/z faces."""
    D = oracle.material_D3(E, NU)
    u = field_to_u(UNIT_HEX, lambda x: np.array([x[0], 0.0, 0.0]))
    f, _ = oracle.hex8_elastic(UNIT_HEX, E, NU, u, jac=False)
    sig = D[:, 0]
    for a in range(8):
        sx = 0.25 if UNIT_HEX[a, 0] == 1 else -0.25
        sy = 0.25 if UNIT_HEX[a, 1] == 1 else -0.25
        sz = 0.25 if UNIT_HEX[a, 2] == 1 else -0.25
        np.testing.assert_allclose(f[0 * 8 + a], sx * sig[0], rtol=1e-13)
        np.testing.assert_allclose(f[1 * 8 + a], sy * sig[1], rtol=1e-13)
        np.testing.assert_allclose(f[2 * 8 + a], sz * sig[2], rtol=1e-13)


def test_hex8_patch_pure_shear(oracle):
    """u_x = g*y: engineering shear gamma_xy = g, sigma_xy = G g."""
    g = 1e-3
    G = E / 2 / (1 + NU)
    u = field_to_u(UNIT_HEX, lambda x: np.array([g * x[1], 0.0, 0.0]))
    f, _ = oracle.hex8_elastic(UNIT_HEX, E, NU, u, jac=False)
    for a in range(8):
        sx = 0.25 if UNIT_HEX[a, 0] == 1 else -0.25
        sy = 0.25 if UNIT_HEX[a, 1] == 1 else -0.25
        np.testing.assert_allclose(f[0 * 8 + a], sy * G * g, rtol=1e-12)   # d N/dy sigma_xy
        np.testing.assert_allclose(f[1 * 8 + a], sx * G * g, rtol=1e-12)   # d N/dx sigma_xy
        assert abs(f[2 * 8 + a]) <= 1e-12 * G * g


def test_hex8_rotation_invariance(oracle):
    R = rot_matrix(0.4, 0.9, -0.3)
    X2 = DISTORTED @ R.T + np.array([3.0, -1.0, 2.0])
    _, K1 = oracle.hex8_elastic(DISTORTED, E, NU)
    _, K2 = oracle.hex8_elastic(X2, E, NU)
    # node-major reorder of the var-major 24 block, then K2 = T K1 T^T
    P = np.array([v * 8 + a for a in range(8) for v in range(3)])
    k1 = K1[:24, :24][np.ix_(P, P)]
    k2 = K2[:24, :24][np.ix_(P, P)]
    T = np.kron(np.eye(8), R)
    # an isotropic material rotates with the element
    np.testing.assert_allclose(k2, T @ k1 @ T.T, rtol=0, atol=1e-10 * np.abs(k1).max())


def test_hex8_sensitivity_is_dD(oracle):
    rng = np.random.default_rng(0)
    u = rng.standard_normal(48)
    u[24:] = 0
    # d/dE: residual is linear in E
    dD = oracle.material_dD3(E, NU, 1.0, 0.0)
    df = oracle.hex8_elastic_sens(DISTORTED, dD.ravel(), u)
    f1, _ = oracle.hex8_elastic(DISTORTED, 1.0, NU, u, jac=False)
    np.testing.assert_allclose(df, f1, rtol=1e-12, atol=1e-14 * np.abs(f1).max())
    # d/dnu against a central difference of the residual
    dD = oracle.material_dD3(E, NU, 0.0, 1.0)
    df = oracle.hex8_elastic_sens(DISTORTED, dD.ravel(), u)
    h = 1e-6
    fp, _ = oracle.hex8_elastic(DISTORTED, E, NU + h, u, jac=False)
    fm, _ = oracle.hex8_elastic(DISTORTED, E, NU - h, u, jac=False)
    np.testing.assert_allclose(df, (fp - fm) / (2 * h), rtol=1e-6, atol=1e-7 * np.abs(df).max())


# ------------------------------------------------------------------ Mindlin --
UNIT_QUAD = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], float)
SKEW_QUAD = np.array([[0, 0, 0], [1.2, 0.1, 0], [1.3, 0.9, 0], [-0.1, 1.1, 0]], float)
H, OFF, KAPPA = 0.06, 0.0, 5.0 / 6.0


def plate_u(X, u=None, v=None, w=None, tx=None, ty=None):
    out = np.zeros(24)
    for a in range(4):
        for k, fn in enumerate((u, v, w, tx, ty)):
            if fn is not None:
                out[k * 4 + a] = fn(X[a])
    return out


@pytest.mark.parametrize("X", [UNIT_QUAD, SKEW_QUAD])
def test_mindlin_rigid_modes(oracle, X):
    _, K = oracle.quad4_mindlin(X, E, NU, H, OFF, KAPPA)
    assert np.abs(K - K.T).max() <= 1e-12 * np.abs(K).max()
    np.testing.assert_allclose(np.diag(K)[20:], 1e-8)          # theta_z diagonal
    k = K[:20, :20]
    ev = np.linalg.eigvalsh(k)
    # 6 rigid-body modes + the 2 zero-energy modes of the 1-point shear rule
    # (w hourglass; theta_x = x, theta_y = y with zero bending strain)
    assert np.sum(np.abs(ev) < 1e-10 * ev.max()) == 8
    hg = plate_u(X, w=lambda x: 1.0 if (x[0] > 0.5) == (x[1] > 0.5) else -1.0)
    if X is UNIT_QUAD:
        f, _ = oracle.quad4_mindlin(X, E, NU, H, OFF, KAPPA, hg)
        assert np.abs(f).max() <= 1e-9 * ev.max()
    t = 0.01
    modes = [plate_u(X, u=lambda x: 1.0), plate_u(X, v=lambda x: 1.0),
             plate_u(X, u=lambda x: -t * x[1], v=lambda x: t * x[0]),
             plate_u(X, w=lambda x: 1.0),
             plate_u(X, w=lambda x: t * x[1], tx=lambda x: t),     # rotation about x
             plate_u(X, w=lambda x: -t * x[0], ty=lambda x: t)]    # rotation about y
    for m in modes:
        f, _ = oracle.quad4_mindlin(X, E, NU, H, OFF, KAPPA, m)
        assert np.abs(f).max() <= 1e-9 * ev.max() * np.abs(m).max()


def test_mindlin_constant_curvature_patch(oracle):
    """theta_y = k x, w = -k x^2/2 on the unit square: kappa_x = k, no shear at the
    1x1 shear point, so the nodal moments are int dN/dx M dA with M = D_b[:,0] k
    (D_b golden of the reference's section test with off = 0.03)."""
    kk = 1e-3
    Db00, Db01 = 5.817528896869037e6, 1.919784535966782e6
    u = plate_u(UNIT_QUAD, w=lambda x: -kk * x[0] ** 2 / 2, ty=lambda x: kk * x[0])
    f, _ = oracle.quad4_mindlin(UNIT_QUAD, E, NU, 0.06, 0.03, KAPPA, u)
    # with off = 0.03 the coupling B = D h off also loads the membrane dofs:
    Bc00, Bc01 = 1.454382224217259e8, 0.479946133991696e8
    for a in range(4):
        sx = 0.5 if UNIT_QUAD[a, 0] == 1 else -0.5
        sy = 0.5 if UNIT_QUAD[a, 1] == 1 else -0.5
        np.testing.assert_allclose(f[4 * 4 + a], sx * Db00 * kk, rtol=1e-12)        # theta_y: dN/dx Mx
        np.testing.assert_allclose(f[3 * 4 + a], -sy * Db01 * kk, rtol=1e-12)       # theta_x: -dN/dy My
        np.testing.assert_allclose(f[0 * 4 + a], sx * Bc00 * kk, rtol=1e-12)        # u: dN/dx Nx
        np.testing.assert_allclose(f[1 * 4 + a], sy * Bc01 * kk, rtol=1e-12)        # v: dN/dy Ny
        assert abs(f[2 * 4 + a]) <= 1e-12 * Db00 * kk                              # no shear force


def test_mindlin_pressure(oracle):
    P = oracle.Problem(oracle.PLATE, 4, 4, seed=1, mast_exact=True)
    P.set_params(np.full(P.n_elems, 0.01))
    R, _ = P.assemble(want_K=False)
    base, con = P.dofmap()
    # R_w = +p int N on free w dofs; the total over all nodes (free + constrained) is p*A
    w_free = [base[n] + 2 for n in range(P.n_nodes) if not con[base[n] + 2]]
    interior_node_load = 1e3 * (1.0 / 16)
    np.testing.assert_allclose(R[w_free], interior_node_load, rtol=1e-13)


# --------------------------------------------------------------------- heat --
def test_heat_constant_and_linear(oracle):
    k0, alpha = 200.0, 0.0
    f, K = oracle.hex8_heat(DISTORTED, k0, alpha, np.full(8, 7.0))
    assert np.abs(f).max() <= 1e-12 * k0
    assert np.abs(K.sum(axis=1)).max() <= 1e-12 * np.abs(K).max()
    T = UNIT_HEX[:, 2] * 3.0           # dT/dz = 3 on the unit cube -> flux k*3
    f, _ = oracle.hex8_heat(UNIT_HEX, k0, alpha, T)
    np.testing.assert_allclose(f, np.where(UNIT_HEX[:, 2] == 1, 0.25, -0.25) * k0 * 3.0, rtol=1e-13)


def test_heat_kT_jacobian_fd(oracle):
    """The dk/dT Jacobian term (heat_conduction_elem_base.cpp:124-142) makes K the
    exact derivative of the residual with k(T) = k0 (1 + alpha T)."""
    rng = np.random.default_rng(1)
    T = 300 + 100 * rng.standard_normal(8)
    f0, K = oracle.hex8_heat(DISTORTED, 200.0, 1e-3, T)
    h = 1e-4
    Kfd = np.zeros((8, 8))
    for b in range(8):
        Tp = T.copy(); Tp[b] += h
        Tm = T.copy(); Tm[b] -= h
        Kfd[:, b] = (oracle.hex8_heat(DISTORTED, 200.0, 1e-3, Tp)[0] -
                     oracle.hex8_heat(DISTORTED, 200.0, 1e-3, Tm)[0]) / (2 * h)
    np.testing.assert_allclose(K, Kfd, rtol=1e-7, atol=1e-8 * np.abs(K).max())
    assert np.abs(K - K.T).max() > 1e-6 * np.abs(K).max()   # genuinely non-symmetric


def test_heat_convection(oracle):
    hc, Tinf = 50.0, 300.0
    T = np.full(8, 350.0)
    for side in range(6):
        f, K = oracle.hex8_convection(UNIT_HEX, side, hc, Tinf, T)
        np.testing.assert_allclose(f.sum(), hc * (350 - 300) * 1.0, rtol=1e-13)
        np.testing.assert_allclose(K.sum(), hc * 1.0, rtol=1e-13)
        assert np.sum(np.abs(f) > 0) == 4     # only the side's 4 nodes
