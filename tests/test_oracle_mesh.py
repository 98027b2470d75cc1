"""Oracle meshes, DofMap and sparsity pinned by hand enumeration and by the
structured-count formulas (nnz = pairs * nvar^2, pairs = prod(3 n_d + 1)),
which reproduce the counts derived for the benchmark configurations.

Parity against libMesh itself is unpinned: libMesh is not part of the
reference; these fix the restated conventions (build_cube numbering, var-major
first-touch DOF bases, full-coupling sorted rows, boundary ids).
"""
import numpy as np


def test_build_cube_2x1x1_hand(oracle):
    P = oracle.Problem(oracle.CANTILEVER, 2, 1, 1, seed=1, lo=0.3, hi=1.0)
    xyz, conn = P.mesh()
    assert xyz.shape == (12, 3)
    # nodes i-fastest: id = i + 3 j + 6 k, x = i/2, y = j*0.5, z = k*0.5
    np.testing.assert_array_equal(xyz[4], [0.5, 0.5, 0.0])
    np.testing.assert_array_equal(xyz[11], [1.0, 0.5, 0.5])
    np.testing.assert_array_equal(conn, [[0, 1, 4, 3, 6, 7, 10, 9], [1, 2, 5, 4, 7, 8, 11, 10]])
    base, con = P.dofmap()
    # first touch: element 0 nodes 0,1,4,3,6,7,10,9 -> ranks 0..7; element 1 adds 2,5,8,11
    rank = {0: 0, 1: 1, 4: 2, 3: 3, 6: 4, 7: 5, 10: 6, 9: 7, 2: 8, 5: 9, 8: 10, 11: 11}
    np.testing.assert_array_equal(base, [3 * rank[n] for n in range(12)])
    # x = 0 face (boundary id 4) = nodes 0, 3, 6, 9: all 3 variables constrained
    exp = np.zeros(36, np.uint8)
    for n in (0, 3, 6, 9):
        exp[base[n]:base[n] + 3] = 1
    np.testing.assert_array_equal(con, exp)
    e, s, b = P.bsides()
    assert sorted(zip(e, s, b)) == sorted([(0, 0, 0), (0, 5, 5), (0, 1, 1), (0, 3, 3), (0, 4, 4),
                                           (1, 0, 0), (1, 5, 5), (1, 1, 1), (1, 3, 3), (1, 2, 2)])


def test_sparsity_2x1x1_hand(oracle):
    P = oracle.Problem(oracle.CANTILEVER, 2, 1, 1, seed=1, lo=0.3, hi=1.0)
    rp, col = P.csr()
    base, _ = P.dofmap()
    # a node of element 0 only couples with the 8 nodes of element 0; a shared
    # node (x = 0.5) couples with all 12
    r0 = col[rp[base[0]]:rp[base[0] + 1]]
    np.testing.assert_array_equal(r0, np.arange(24))
    r1 = col[rp[base[1]]:rp[base[1] + 1]]
    np.testing.assert_array_equal(r1, np.arange(36))
    assert rp[-1] == (8 * 8 * 2 - 4 * 4) * 9


def structured_pairs(*n):
    p = 1
    for k in n:
        p *= 3 * k + 1
    return p


def test_nnz_formula_configs(oracle):
    # C1 counts (20x10x10): 527,589 active, 2,110,356 with MAST's 6 variables
    P = oracle.Problem(oracle.CANTILEVER, 20, 10, 10, seed=1, lo=0.3, hi=1.0)
    assert len(P.csr()[1]) == 527589 == structured_pairs(20, 10, 10) * 9
    P6 = oracle.Problem(oracle.CANTILEVER, 20, 10, 10, seed=1, lo=0.3, hi=1.0, mast_exact=True)
    assert len(P6.csr()[1]) == 2110356 == structured_pairs(20, 10, 10) * 36
    assert P6.n_dofs == 15246 and P.n_dofs == 7623
    # plate and heat counts on small meshes follow the same rule
    Q = oracle.Problem(oracle.PLATE, 5, 4, seed=1)
    assert len(Q.csr()[1]) == structured_pairs(5, 4) * 25
    Hh = oracle.Problem(oracle.HEAT, 4, seed=1)
    assert len(Hh.csr()[1]) == structured_pairs(4, 4, 4)


def test_csr_sorted_symmetric_and_row_sizes(oracle):
    P = oracle.Problem(oracle.CANTILEVER, 5, 4, 3, seed=1, lo=0.3, hi=1.0)
    rp, col = P.csr()
    n = P.n_dofs
    lens = np.diff(rp)
    assert set(np.unique(lens // 3)) <= {8, 12, 18, 27}
    rows = np.repeat(np.arange(n), lens)
    for r in range(n):
        assert np.all(np.diff(col[rp[r]:rp[r + 1]]) > 0)
    a = set(zip(rows.tolist(), col.tolist()))
    assert all((c, r) in a for r, c in a)
    # sampled rows from the per-node union agree with the full build
    for d in (0, 17, 100, n - 1):
        np.testing.assert_array_equal(P.sparsity_row(d), col[rp[d]:rp[d + 1]])


def test_point_loads_c1(oracle):
    P = oracle.Problem(oracle.CANTILEVER, 20, 10, 10, seed=1, lo=0.3, hi=1.0)
    F = P.point_load()
    base, _ = P.dofmap()
    xyz, _ = P.mesh()
    np.testing.assert_allclose(F.sum(), -1e6 * 0.25, rtol=1e-13)
    on = np.isclose(xyz[:, 0], 1.0)
    fz = F[base + 2]
    # 0.05 x 0.05 faces: corner 625, edge 1250, interior 2500 (N)
    corner = on & np.isin(xyz[:, 1], [0, 0.5]) & np.isin(xyz[:, 2], [0, 0.5])
    edge = on & (np.isin(xyz[:, 1], [0, 0.5]) ^ np.isin(xyz[:, 2], [0, 0.5]))
    inner = on & ~corner & ~edge
    np.testing.assert_allclose(fz[corner], -625.0, rtol=1e-12)
    np.testing.assert_allclose(fz[edge], -1250.0, rtol=1e-12)
    np.testing.assert_allclose(fz[inner], -2500.0, rtol=1e-12)
    assert np.all(fz[~on] == 0) and np.all(F[base] == 0) and np.all(F[base + 1] == 0)


def test_seeded_uniform(oracle):
    a = oracle.uniform(1, 1000, 0.3, 1.0)
    assert a.min() >= 0.3 and a.max() < 1.0
    assert abs(a.mean() - 0.65) < 0.02
    np.testing.assert_array_equal(a, oracle.uniform(1, 1000, 0.3, 1.0))
    P = oracle.Problem(oracle.CANTILEVER, 10, 5, 5, seed=1, lo=0.3, hi=1.0)
    np.testing.assert_array_equal(P.params(), oracle.uniform(1, 250, 0.3, 1.0))
