"""Test-side restatement of the product's multigrid preconditioner, in scipy,
on the oracle's assembled matrix: trilinear transfers with constrained rows /
columns removed, Galerkin coarse operators (constrained coarse DOFs get a unit
diagonal per touching element), Chebyshev smoothing on D^-1 A over
[lam/80, lam] with lam the Gershgorin bound max_i sum_j |a_ij| / a_ii, `k`
steps before and after the coarse correction, exact coarsest solve.  Used to
pin the GPU iteration count, which only matches when every piece matches."""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

CORNER = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]


def lattice_of(xyz, h):
    lo = xyz.min(0)
    return np.rint((xyz - lo) / h).astype(np.int64)


def active_system(P):
    """Oracle K, b restricted to the translational DOFs (node-major order as
    the product's active numbering), with lattice index and var of each DOF."""
    rp, col = P.csr()
    R0, K = P.assemble()
    base, con = P.dofmap()
    n = P.n_dofs
    A = sp.csr_matrix((K, col, rp), shape=(n, n))
    order = np.argsort(base, kind="stable")
    idx = np.concatenate([[base[nd] + v for v in range(3)] for nd in order]).astype(np.int64)
    node = np.repeat(order, 3)
    var = np.tile(np.arange(3), len(order))
    return A[idx][:, idx].tocsr(), -R0[idx], con[idx].astype(bool), node, var


def prolongation(dims_f, ijk_f, var_f, con_f, dims_c, con_c):
    """P[f, c] for fine DOFs (lattice ijk_f, var) and coarse DOFs in lattice
    order (node-major, 3 vars)."""
    cx, cy = dims_c[0] + 1, dims_c[1] + 1
    rows, cols, vals = [], [], []
    for f in range(len(var_f)):
        if con_f[f]:
            continue
        opts = []
        for d in range(3):
            p = ijk_f[f, d]
            opts.append([(p // 2, 1.0)] if p % 2 == 0 else [(p // 2, 0.5), (p // 2 + 1, 0.5)])
        for (i, wi) in opts[0]:
            for (j, wj) in opts[1]:
                for (k, wk) in opts[2]:
                    c = (i + cx * (j + cy * k)) * 3 + var_f[f]
                    if not con_c[c]:
                        rows.append(f); cols.append(c); vals.append(wi * wj * wk)
    return sp.csr_matrix((vals, (rows, cols)), shape=(len(var_f), 3 * (dims_c[0] + 1) * cy * (dims_c[2] + 1)))


def build_levels(P, dims, h, max_coarse=600):
    A, b, con, node, var = active_system(P)
    xyz, _ = P.mesh()
    ijk = lattice_of(xyz, h)[node]
    levels = [dict(A=A, con=con)]
    cur_ijk, cur_var, cur_con, cur_dims = ijk, var, con, tuple(dims)
    while A.shape[0] > max_coarse and all(d % 2 == 0 for d in cur_dims):
        dc = tuple(d // 2 for d in cur_dims)
        g = np.stack(np.meshgrid(*[np.arange(d + 1) for d in dc], indexing="ij"), -1)
        g = g.transpose(2, 1, 0, 3).reshape(-1, 3)          # x fastest
        ijk_c = np.repeat(g, 3, axis=0)
        var_c = np.tile(np.arange(3), len(g))
        # a coarse DOF is constrained when the fine DOF at the same point is
        fine_at = {}
        for f in range(len(cur_var)):
            fine_at[(tuple(cur_ijk[f]), cur_var[f])] = f
        con_c = np.array([cur_con[fine_at[(tuple(2 * ijk_c[c]), var_c[c])]] for c in range(len(var_c))])
        Pm = prolongation(cur_dims, cur_ijk, cur_var, cur_con, dc, con_c)
        Ac = (Pm.T @ levels[-1]["A"] @ Pm).tolil()
        # unit diagonal per coarse element touching a constrained coarse DOF
        for c in np.nonzero(con_c)[0]:
            I, J, K = ijk_c[c]
            cnt = 1
            for dd, nn in zip((I, J, K), dc):
                cnt *= (2 if 0 < dd < nn else 1)
            Ac[c, c] = float(cnt)
        levels[-1]["P"] = Pm
        levels.append(dict(A=Ac.tocsr(), con=con_c))
        A = levels[-1]["A"]
        cur_ijk, cur_var, cur_con, cur_dims = ijk_c, var_c, con_c, dc
    for L in levels:
        d = L["A"].diagonal()
        L["dinv"] = 1.0 / d
        L["lam"] = float(((abs(L["A"]) @ np.ones(L["A"].shape[0])) / d).max())
    levels[-1]["inv"] = np.linalg.inv(levels[-1]["A"].toarray())
    return levels, b


def chebyshev(A, dinv, r, x, k, lam, ratio=80.0):
    a, bnd = lam / ratio, lam
    theta, delta = 0.5 * (bnd + a), 0.5 * (bnd - a)
    sigma = theta / delta
    rho = 1.0 / sigma
    res = r - A @ x
    d = dinv * res / theta
    for i in range(1, k):
        x = x + d
        res = res - A @ d
        rn = 1.0 / (2 * sigma - rho)
        d = rn * rho * d + 2 * rn / delta * (dinv * res)
        rho = rn
    return x + d


def vcycle(levels, l, r, k):
    """k = (steps on the problem level, steps on the coarse levels)."""
    L = levels[l]
    if "inv" in L:
        return L["inv"] @ r
    A, dinv = L["A"], L["dinv"]
    kl = k[0] if l == 0 else k[1]
    x = chebyshev(A, dinv, r, np.zeros_like(r), kl, L["lam"])
    x = x + L["P"] @ vcycle(levels, l + 1, L["P"].T @ (r - A @ x), k)
    return chebyshev(A, dinv, r, x, kl, L["lam"])


def pcg_iterations(levels, b, k=(6, 2), rtol=1e-10, max_it=1000):
    A = levels[0]["A"]
    x = np.zeros_like(b); r = b.copy()
    z = vcycle(levels, 0, r, k); p = z.copy(); rz = r @ z
    bb = b @ b
    for it in range(1, max_it + 1):
        q = A @ p
        a = rz / (p @ q)
        x += a * p; r -= a * q
        if r @ r <= rtol * rtol * bb:
            return it, x
        z = vcycle(levels, 0, r, k)
        rzn = r @ z
        p = z + (rzn / rz) * p
        rz = rzn
    return max_it, x
