"""ctypes view of liboracle.so: the FP64 CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module, as the checker; the product library never loads it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None

_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)
_U8 = C.POINTER(C.c_uint8)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(path)
        sig = {
            "oracle_problem_new": (_P, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64,
                                        C.c_double, C.c_double, C.c_int]),
            "oracle_problem_free": (None, [_P]),
            "oracle_problem_from_arrays": (_P, [C.c_int, C.c_int64, _D, C.c_int64, _I32, C.c_int, _I64,
                                                C.c_int64, _U8, _D, _D]),
            "oracle_set_threads": (None, [_P, C.c_int]),
            "oracle_max_threads": (C.c_int, []),
            "oracle_set_scalar": (C.c_int, [_P, C.c_char_p, C.c_double]),
            "oracle_n_nodes": (C.c_int64, [_P]),
            "oracle_n_elems": (C.c_int64, [_P]),
            "oracle_n_dofs": (C.c_int64, [_P]),
            "oracle_n_vars": (C.c_int, [_P]),
            "oracle_npe": (C.c_int, [_P]),
            "oracle_n_bsides": (C.c_int64, [_P]),
            "oracle_get_mesh": (None, [_P, _D, _I32]),
            "oracle_get_bsides": (None, [_P, _I32, _I32, _I32]),
            "oracle_set_bsides": (None, [_P, C.c_int64, _I32, _I32, _I32]),
            "oracle_get_dofmap": (None, [_P, _I64, _U8]),
            "oracle_get_params": (None, [_P, _D]),
            "oracle_set_params": (None, [_P, _D]),
            "oracle_get_point_load": (None, [_P, _D]),
            "oracle_csr_nnz": (C.c_int64, [_P]),
            "oracle_csr": (None, [_P, _I64, _I32]),
            "oracle_sparsity_row": (C.c_int64, [_P, C.c_int64, _I32]),
            "oracle_assemble": (None, [_P, _D, _D, _D]),
            "oracle_sens_dots": (None, [_P, _D, _D, _D]),
            "oracle_cg": (C.c_int, [_P, _D, _D, C.c_double, C.c_int, _D]),
            "oracle_spmv": (None, [_P, _D, _D]),
            "oracle_solve_elastic": (C.c_int, [_P, C.c_double, C.c_int, _D, _D, _D, _D]),
            "oracle_solve_heat": (C.c_int, [_P, C.c_double, C.c_double, _D, _D, _D, _D, C.c_int]),
            "oracle_material_D3": (None, [C.c_double, C.c_double, _D]),
            "oracle_material_dD3": (None, [C.c_double, C.c_double, C.c_double, C.c_double, _D]),
            "oracle_material_D2": (None, [C.c_double, C.c_double, _D]),
            "oracle_material_dD2": (None, [C.c_double, C.c_double, C.c_double, C.c_double, _D]),
            "oracle_section": (None, [C.c_double] * 5 + [_D] * 4),
            "oracle_hex8_elastic": (None, [_D, C.c_double, C.c_double, _D, _D, _D]),
            "oracle_hex8_elastic_sens": (None, [_D, _D, _D, _D]),
            "oracle_quad4_mindlin": (None, [_D] + [C.c_double] * 5 + [_D, _D, _D]),
            "oracle_hex8_heat": (None, [_D, C.c_double, C.c_double, _D, _D, _D]),
            "oracle_hex8_convection": (None, [_D, C.c_int, C.c_double, C.c_double, _D, _D, _D]),
            "oracle_hex8_pressure": (None, [_D, C.c_int, C.c_double, _D]),
            "oracle_uniform": (None, [C.c_uint64, C.c_int64, C.c_double, C.c_double, _D]),
            "oracle_hex8_radiation": (None, [_D, C.c_int] + [C.c_double] * 4 + [_D, _D, _D]),
            "oracle_hex8_stress": (None, [_D, C.c_double, C.c_double, _D, _D, _D]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a, t=_D):
    if a is None:
        return None
    return a.ctypes.data_as(t)


CANTILEVER, PLATE, HEAT = 1, 2, 3


class Problem:
    """A workload built by the oracle (mesh, DofMap, loads, seeded parameters)."""

    def __init__(self, kind, nx=0, ny=0, nz=0, seed=1, lo=0.0, hi=1.0, mast_exact=False,
                 threads=0, arrays=None):
        """A structured workload, or (arrays = dict(xyz, conn, n_vars, node_base,
        constrained, params[, point_load][, sides = (elem, side, id)])) a
        problem over any mesh and DofMap."""
        self.kind = kind
        if arrays is not None:
            a = {k: v for k, v in arrays.items()}
            xyz = np.ascontiguousarray(a["xyz"], np.float64)
            conn = np.ascontiguousarray(a["conn"], np.int32)
            base = np.ascontiguousarray(a["node_base"], np.int64)
            con = np.ascontiguousarray(a["constrained"], np.uint8)
            par = np.ascontiguousarray(a["params"], np.float64)
            pl = a.get("point_load")
            pl = None if pl is None else np.ascontiguousarray(pl, np.float64)
            self.h = lib().oracle_problem_from_arrays(kind, xyz.shape[0], ptr(xyz), conn.shape[0], ptr(conn, _I32),
                                                      int(a["n_vars"]), ptr(base, _I64), con.size, ptr(con, _U8),
                                                      ptr(par), ptr(pl))
            if a.get("sides") is not None:
                se, si, sb = (np.ascontiguousarray(v, np.int32) for v in a["sides"])
                lib().oracle_set_bsides(self.h, se.size, ptr(se, _I32), ptr(si, _I32), ptr(sb, _I32))
        else:
            self.h = lib().oracle_problem_new(kind, nx, ny, nz, seed, lo, hi, int(mast_exact))
        if threads:
            lib().oracle_set_threads(self.h, threads)
        L = lib()
        self.n_nodes = L.oracle_n_nodes(self.h)
        self.n_elems = L.oracle_n_elems(self.h)
        self.n_dofs = L.oracle_n_dofs(self.h)
        self.n_vars = L.oracle_n_vars(self.h)
        self.npe = L.oracle_npe(self.h)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.oracle_problem_free(self.h)
            self.h = None

    def set(self, **kw):
        for k, v in kw.items():
            if lib().oracle_set_scalar(self.h, k.encode(), float(v)):
                raise KeyError(k)
        return self

    def mesh(self):
        xyz = np.zeros((self.n_nodes, 3))
        conn = np.zeros((self.n_elems, self.npe), np.int32)
        lib().oracle_get_mesh(self.h, ptr(xyz), ptr(conn, _I32))
        return xyz, conn

    def bsides(self):
        n = lib().oracle_n_bsides(self.h)
        e = np.zeros(n, np.int32); s = np.zeros(n, np.int32); b = np.zeros(n, np.int32)
        lib().oracle_get_bsides(self.h, ptr(e, _I32), ptr(s, _I32), ptr(b, _I32))
        return e, s, b

    def dofmap(self):
        base = np.zeros(self.n_nodes, np.int64)
        con = np.zeros(self.n_dofs, np.uint8)
        lib().oracle_get_dofmap(self.h, ptr(base, _I64), ptr(con, _U8))
        return base, con

    def params(self):
        p = np.zeros(self.n_elems)
        lib().oracle_get_params(self.h, ptr(p))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float64)
        lib().oracle_set_params(self.h, ptr(p))

    def point_load(self):
        f = np.zeros(self.n_dofs)
        lib().oracle_get_point_load(self.h, ptr(f))
        return f

    def csr(self):
        nnz = lib().oracle_csr_nnz(self.h)
        rp = np.zeros(self.n_dofs + 1, np.int64)
        col = np.zeros(nnz, np.int32)
        lib().oracle_csr(self.h, ptr(rp, _I64), ptr(col, _I32))
        return rp, col

    def sparsity_row(self, dof):
        n = lib().oracle_sparsity_row(self.h, dof, None)
        out = np.zeros(n, np.int32)
        lib().oracle_sparsity_row(self.h, dof, ptr(out, _I32))
        return out

    def assemble(self, X=None, want_K=True):
        R = np.zeros(self.n_dofs)
        K = np.zeros(lib().oracle_csr_nnz(self.h)) if want_K else None
        Xa = None if X is None else np.ascontiguousarray(X, np.float64)
        lib().oracle_assemble(self.h, ptr(Xa), ptr(R), ptr(K))
        return R, K

    def sens_dots(self, X, lam):
        g = np.zeros(self.n_elems)
        X = np.ascontiguousarray(X, np.float64)
        lam = np.ascontiguousarray(lam, np.float64)
        lib().oracle_sens_dots(self.h, ptr(X), ptr(lam), ptr(g))
        return g

    def cg(self, b, rtol=1e-10, max_it=100000):
        x = np.zeros(self.n_dofs)
        rr = C.c_double()
        b = np.ascontiguousarray(b, np.float64)
        it = lib().oracle_cg(self.h, ptr(b), ptr(x), rtol, max_it, C.byref(rr))
        return x, it, rr.value

    def spmv(self, x):
        y = np.zeros(self.n_dofs)
        x = np.ascontiguousarray(x, np.float64)
        lib().oracle_spmv(self.h, ptr(x), ptr(y))
        return y

    def solve_elastic(self, rtol=1e-10, max_it=200000):
        u = np.zeros(self.n_dofs)
        c = C.c_double()
        dc = np.zeros(self.n_elems)
        rr = C.c_double()
        it = lib().oracle_solve_elastic(self.h, rtol, max_it, ptr(u), C.byref(c), ptr(dc),
                                        C.byref(rr))
        return dict(u=u, compliance=c.value, dc=dc, iters=it, rel_res=rr.value)

    def solve_heat(self, newton_rtol=1e-10, lin_rtol=1e-12):
        T = np.zeros(self.n_dofs)
        m = C.c_double()
        dJ = np.zeros(self.n_elems)
        hist = np.zeros(64)
        it = lib().oracle_solve_heat(self.h, newton_rtol, lin_rtol, ptr(T), C.byref(m), ptr(dJ),
                                     ptr(hist), 64)
        return dict(T=T, mean_T=m.value, dJ=dJ, newton_iters=it, res_hist=hist[: it + 1])


def uniform(seed, n, lo, hi):
    out = np.zeros(n)
    lib().oracle_uniform(seed, n, lo, hi, ptr(out))
    return out


def material_D3(E, nu):
    D = np.zeros(36)
    lib().oracle_material_D3(E, nu, ptr(D))
    return D.reshape(6, 6)


def material_dD3(E, nu, dE, dnu):
    D = np.zeros(36)
    lib().oracle_material_dD3(E, nu, dE, dnu, ptr(D))
    return D.reshape(6, 6)


def material_D2(E, nu):
    D = np.zeros(9)
    lib().oracle_material_D2(E, nu, ptr(D))
    return D.reshape(3, 3)


def material_dD2(E, nu, dE, dnu):
    D = np.zeros(9)
    lib().oracle_material_dD2(E, nu, dE, dnu, ptr(D))
    return D.reshape(3, 3)


def section(E, nu, h, off, kappa):
    A = np.zeros(9); B = np.zeros(9); Db = np.zeros(9); S = np.zeros(4)
    lib().oracle_section(E, nu, h, off, kappa, ptr(A), ptr(B), ptr(Db), ptr(S))
    return A.reshape(3, 3), B.reshape(3, 3), Db.reshape(3, 3), S.reshape(2, 2)


def hex8_elastic(X, E, nu, u=None, jac=True):
    X = np.ascontiguousarray(X, np.float64).reshape(8, 3)
    u = np.zeros(48) if u is None else np.ascontiguousarray(u, np.float64)
    f = np.zeros(48)
    K = np.zeros((48, 48)) if jac else None
    lib().oracle_hex8_elastic(ptr(X), E, nu, ptr(u), ptr(f), ptr(K))
    return f, K


def hex8_elastic_sens(X, dD, u):
    X = np.ascontiguousarray(X, np.float64).reshape(8, 3)
    dD = np.ascontiguousarray(dD, np.float64)
    u = np.ascontiguousarray(u, np.float64)
    df = np.zeros(48)
    lib().oracle_hex8_elastic_sens(ptr(X), ptr(dD), ptr(u), ptr(df))
    return df


def quad4_mindlin(X, E, nu, h, off, kappa, u=None):
    X = np.ascontiguousarray(X, np.float64).reshape(4, 3)
    u = np.zeros(24) if u is None else np.ascontiguousarray(u, np.float64)
    f = np.zeros(24)
    K = np.zeros((24, 24))
    lib().oracle_quad4_mindlin(ptr(X), E, nu, h, off, kappa, ptr(u), ptr(f), ptr(K))
    return f, K


def hex8_heat(X, k0, alpha, T):
    X = np.ascontiguousarray(X, np.float64).reshape(8, 3)
    T = np.ascontiguousarray(T, np.float64)
    f = np.zeros(8)
    K = np.zeros((8, 8))
    lib().oracle_hex8_heat(ptr(X), k0, alpha, ptr(T), ptr(f), ptr(K))
    return f, K


def hex8_pressure(X, side, p):
    """Element vector of a uniform pressure p on side `side` (variable-major
    u0..u7, v0..v7, w0..w7): f = -int N p n dA."""
    X = np.ascontiguousarray(X, np.float64).reshape(8, 3)
    f = np.zeros(24)
    lib().oracle_hex8_pressure(ptr(X), side, p, ptr(f))
    return f


def hex8_convection(X, side, h, Tinf, T):
    X = np.ascontiguousarray(X, np.float64).reshape(8, 3)
    T = np.ascontiguousarray(T, np.float64)
    f = np.zeros(8)
    K = np.zeros((8, 8))
    lib().oracle_hex8_convection(ptr(X), side, h, Tinf, ptr(T), ptr(f), ptr(K))
    return f, K


def hex8_radiation(X, side, sigma, emiss, Tamb, T0, T):
    X = np.ascontiguousarray(X, np.float64).reshape(8, 3)
    T = np.ascontiguousarray(T, np.float64)
    f = np.zeros(8)
    K = np.zeros((8, 8))
    lib().oracle_hex8_radiation(ptr(X), side, sigma, emiss, Tamb, T0, ptr(T), ptr(f), ptr(K))
    return f, K


def hex8_stress(X, E, nu, u):
    """Stresses [8, 6] (xx yy zz xy yz zx) and JxW [8] at the quadrature points."""
    X = np.ascontiguousarray(X, np.float64).reshape(8, 3)
    u = np.ascontiguousarray(u, np.float64)
    sig = np.zeros((8, 6))
    jxw = np.zeros(8)
    lib().oracle_hex8_stress(ptr(X), E, nu, ptr(u), ptr(sig), ptr(jxw))
    return sig, jxw
