"""CPU checks of the scipy multigrid restatement the GPU multigrid is pinned
against: trilinear prolongation reproduces linear fields exactly, Galerkin
operators stay symmetric positive definite, and the preconditioned CG reaches
the oracle's compliance in far fewer iterations than Jacobi."""
import numpy as np
import scipy.sparse as sp

import mg_restatement as MR


def test_prolongation_reproduces_linear_fields(oracle):
    P = oracle.Problem(oracle.CANTILEVER, 8, 4, 4, seed=1, lo=0.3, hi=1.0)
    levels, _ = MR.build_levels(P, (8, 4, 4), np.array([1 / 8, 0.5 / 4, 0.5 / 4]), max_coarse=100)
    Pm = levels[0]["P"]
    A, b, con, node, var = MR.active_system(P)
    xyz, _ = P.mesh()
    X = xyz[node]
    ijk_c = np.repeat(np.stack(np.meshgrid(np.arange(5), np.arange(3), np.arange(3), indexing="ij"), -1)
                      .transpose(2, 1, 0, 3).reshape(-1, 3), 3, axis=0)
    Xc = ijk_c * np.array([2 / 8, 2 * 0.5 / 4, 2 * 0.5 / 4])
    var_c = np.tile(np.arange(3), len(Xc) // 3)
    Bm = np.array([[0.3, -1.0, 2.0], [0.5, 0.25, -0.75], [1.5, 0.0, 1.0]])
    uf = (X @ Bm.T)[np.arange(len(var)), var]
    uc = (Xc @ Bm.T)[np.arange(len(var_c)), var_c]
    free = ~con
    # columns of constrained coarse DOFs are removed, so compare where the
    # interpolating coarse DOFs are all free (x > 0 planes away from the clamp)
    away = X[:, 0] > 2 / 8 + 1e-12
    np.testing.assert_allclose((Pm @ uc)[free & away], uf[free & away], atol=1e-13)


def test_galerkin_spd_and_pcg(oracle):
    P = oracle.Problem(oracle.CANTILEVER, 16, 8, 8, seed=2, lo=0.01, hi=1.0)
    levels, b = MR.build_levels(P, (16, 8, 8), np.array([1 / 16, 0.5 / 8, 0.5 / 8]))
    for L in levels[1:]:
        A = L["A"]
        assert abs(A - A.T).max() <= 1e-12 * abs(A).max()
        assert np.linalg.eigvalsh(A.toarray()).min() > 0 if A.shape[0] <= 700 else True
    it, x = MR.pcg_iterations(levels, b)
    ref = P.solve_elastic(1e-10)
    assert abs(b @ x - ref["compliance"]) <= 1e-9 * ref["compliance"]
    assert it * 10 < ref["iters"]
