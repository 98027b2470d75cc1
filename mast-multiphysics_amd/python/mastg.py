"""Python view of the mastg C ABI (include/mastg/mastg.h) through ctypes.

Loads the in-tree libmastg.so; there is no fallback: if the library is not
built, or its calls fail, the error is raised.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "libmastg.so"))

HEX8_ELASTIC, QUAD4_MINDLIN, HEX8_HEAT = 1, 2, 3
HOST, DEVICE = 0, 1
ELEM_FP64, ELEM_FP32 = 0, 1
SOLVER_CG_JACOBI, SOLVER_MIXED, SOLVER_BICGSTAB, SOLVER_MATFREE, SOLVER_MG = 0, 1, 2, 3, 4
OUTPUT_COMPLIANCE, OUTPUT_MEAN_T, OUTPUT_KS_STRESS, OUTPUT_PNORM_STRESS, OUTPUT_RAMP_STRESS = 1, 2, 3, 4, 5
PARAM_E, PARAM_NU, PARAM_PRESSURE, PARAM_K0, PARAM_CONV_H, PARAM_EMISSIVITY = 1, 2, 3, 4, 5, 6
OPT_GRAPHS = 1
OPT_MG_STEPS_FINE = 2
OPT_MG_STEPS_COARSE = 3
OPT_MG_CHEB_RATIO = 4
OPT_ROW_ASSEMBLY = 5
ERRORS = {0: "ok", -1: "invalid", -2: "out of memory", -3: "not converged", -4: "nan",
          -5: "device error", -6: "admission", -7: "busy"}


class ProblemDesc(C.Structure):
    _fields_ = [
        ("physics", C.c_int32), ("n_nodes", C.c_int64), ("n_elems", C.c_int64),
        ("xyz", C.POINTER(C.c_double)), ("conn", C.POINTER(C.c_int32)),
        ("n_vars", C.c_int32), ("node_dof_base", C.POINTER(C.c_int64)),
        ("n_constrained", C.c_int64), ("constrained_dofs", C.POINTER(C.c_int64)),
        ("n_sides", C.c_int64), ("side_elem", C.POINTER(C.c_int32)),
        ("side_index", C.POINTER(C.c_int32)), ("side_id", C.POINTER(C.c_int32)),
        ("point_loads", C.POINTER(C.c_double)),
        ("E", C.c_double), ("nu", C.c_double), ("penalty", C.c_double),
        ("kappa", C.c_double), ("offset", C.c_double), ("pressure", C.c_double),
        ("k0", C.c_double), ("alpha", C.c_double), ("conv_h", C.c_double),
        ("conv_tinf", C.c_double), ("conv_bid_mask", C.c_uint32), ("flags", C.c_int32),
        ("rad_emissivity", C.c_double), ("rad_tamb", C.c_double), ("rad_tzero", C.c_double),
        ("rad_sigma", C.c_double), ("rad_bid_mask", C.c_uint32), ("pres_bid_mask", C.c_uint32),
    ]


class Sizes(C.Structure):
    _fields_ = [("n_dofs", C.c_int64), ("n_active", C.c_int64), ("nnz", C.c_int64),
                ("n_blocks", C.c_int64), ("n_active_vars", C.c_int32),
                ("nodes_per_elem", C.c_int32), ("device_bytes", C.c_int64)]


class SolveOpts(C.Structure):
    _fields_ = [("solver", C.c_int32), ("max_iters", C.c_int32), ("rtol", C.c_double),
                ("reliable_delta", C.c_double), ("transpose", C.c_int32),
                ("check_every", C.c_int32)]


class SolveStats(C.Structure):
    _fields_ = [("iters", C.c_int32), ("replacements", C.c_int32), ("rel_res", C.c_double),
                ("seconds", C.c_double)]


class ServerOpts(C.Structure):
    _fields_ = [("n_slots", C.c_int32), ("solver", C.c_int32), ("hbm_budget", C.c_int64),
                ("queue_capacity", C.c_int32), ("rtol", C.c_double), ("rtol_plate", C.c_double)]


class Request(C.Structure):
    _fields_ = [("kind", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("params", C.POINTER(C.c_double)), ("grad_out", C.POINTER(C.c_double))]


class RequestResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("iters", C.c_int32), ("objective", C.c_double),
                ("t_queue", C.c_double), ("t_run", C.c_double)]


class DesignResult(C.Structure):
    _fields_ = [("objective", C.c_double), ("iters", C.c_int32), ("newton_iters", C.c_int32),
                ("rel_res", C.c_double), ("t_assemble", C.c_double), ("t_solve", C.c_double),
                ("t_sens", C.c_double)]


_lib = None
_P = C.c_void_p
_D = C.POINTER(C.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libmastg.so is not built: run __graft_entry__.build() "
                               "or make -C mast-multiphysics_amd")
        L = C.CDLL(LIB_PATH)
        sig = {
            "mastg_ctx_create": (C.c_int, [C.POINTER(ProblemDesc), C.POINTER(_P)]),
            "mastg_ctx_destroy": (C.c_int, [_P]),
            "mastg_last_error": (C.c_char_p, [_P]),
            "mastg_get_sizes": (C.c_int, [_P, C.POINTER(Sizes)]),
            "mastg_get_csr_pattern": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
            "mastg_get_active_map": (C.c_int, [_P, C.POINTER(C.c_int64)]),
            "mastg_set_params": (C.c_int, [_P, _D, C.c_int]),
            "mastg_assemble": (C.c_int, [_P, _D, _D, _D, C.c_int, C.c_int]),
            "mastg_solve": (C.c_int, [_P, _D, _D, C.c_int, C.POINTER(SolveOpts),
                                      C.POINTER(SolveStats)]),
            "mastg_sensitivity": (C.c_int, [_P, _D, _D, _D, C.c_int]),
            "mastg_jacobian_product": (C.c_int, [_P, _D, _D, C.c_int, C.c_int]),
            "mastg_output": (C.c_int, [_P, C.c_int, _D, C.c_int, _D, _D]),
            "mastg_output_noalloc": (C.c_int, [_P, C.c_int, C.c_void_p, C.c_void_p]),
            "mastg_csr_full_size": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
            "mastg_export_csr_full": (C.c_int, [_P, C.c_void_p, C.c_void_p, C.c_void_p]),
            "mastg_set_ks_stress": (C.c_int, [_P, C.c_double, C.c_double]),
            "mastg_set_stress_aggregation": (C.c_int, [_P, C.c_double, C.c_double, C.c_double, C.c_double]),
            "mastg_output_sensitivity": (C.c_int, [_P, C.c_int, _D, C.c_int, _D]),
            "mastg_device_alloc": (C.c_int, [_P, C.c_size_t, C.POINTER(C.c_void_p)]),
            "mastg_device_free": (C.c_int, [_P, C.c_void_p]),
            "mastg_copy": (C.c_int, [_P, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
            "mastg_sensitivity_rhs": (C.c_int, [_P, C.c_int, _D, _D, C.c_int]),
            "mastg_sensitivity_global": (C.c_int, [_P, C.c_int, _D, _D, C.c_int, _D]),
            "mastg_sensitivity_solve": (C.c_int, [_P, C.c_int, _D, _D, C.c_int, C.POINTER(SolveOpts),
                                                  C.POINTER(SolveStats)]),
            "mastg_design_eval": (C.c_int, [_P, _D, _D, C.c_int, C.POINTER(SolveOpts),
                                            C.POINTER(DesignResult)]),
            "mastg_enable_kernel_timing": (C.c_int, [_P, C.c_int]),
            "mastg_kernel_time": (C.c_int, [_P, C.c_int, _D, C.POINTER(C.c_int64)]),
            "mastg_synchronize": (C.c_int, [_P]),
            "mastg_get_stream": (C.c_int, [_P, C.POINTER(C.c_void_p)]),
            "mastg_set_stream": (C.c_int, [_P, C.c_void_p]),
            "mastg_bench_kernel": (C.c_int, [_P, C.c_int, C.c_int, _D]),
            "mastg_get_tuning": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "mastg_set_option": (C.c_int, [_P, C.c_int, C.c_int]),
            "mastg_set_tuning": (C.c_int, [_P, C.c_int, C.c_int]),
            "mastg_batch_create": (C.c_int, [C.POINTER(ProblemDesc), C.c_int, C.POINTER(_P)]),
            "mastg_batch_destroy": (C.c_int, [_P]),
            "mastg_batch_lanes": (C.c_int, [_P]),
            "mastg_batch_lane": (C.c_int, [_P, C.c_int, C.POINTER(_P)]),
            "mastg_batch_set_option": (C.c_int, [_P, C.c_int, C.c_int]),
            "mastg_batch_eval": (C.c_int, [_P, C.c_int, _D, _D, C.c_int, C.POINTER(SolveOpts),
                                           C.POINTER(DesignResult), _D]),
            "mastg_batch_last_error": (C.c_char_p, [_P]),
            "mastg_server_create": (C.c_int, [C.POINTER(ServerOpts), C.POINTER(_P)]),
            "mastg_server_submit": (C.c_int, [_P, C.POINTER(Request), C.POINTER(C.c_int64)]),
            "mastg_server_poll": (C.c_int, [_P, C.c_int64, C.POINTER(RequestResult)]),
            "mastg_server_wait": (C.c_int, [_P, C.c_int64, C.POINTER(RequestResult)]),
            "mastg_server_stats": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                              C.POINTER(C.c_int32)]),
            "mastg_server_destroy": (C.c_int, [_P]),
            "mastg_server_last_error": (C.c_char_p, [_P]),
            "mastg_request_footprint": (C.c_int, [C.POINTER(Request), C.POINTER(C.c_int64)]),
            "mastg_workload_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64,
                                                C.c_double, C.c_double, C.c_int, C.POINTER(_P)]),
            "mastg_workload_desc": (C.POINTER(ProblemDesc), [_P]),
            "mastg_workload_params": (_D, [_P]),
            "mastg_workload_destroy": (C.c_int, [_P]),
            "mastg_nastran_read": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
            "mastg_nastran_last_error": (C.c_char_p, []),
            "mastg_nastran_desc": (C.POINTER(ProblemDesc), [_P]),
            "mastg_nastran_params": (_D, [_P]),
            "mastg_nastran_ids": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                            C.POINTER(C.c_int32)]),
            "mastg_nastran_spc_sets": (C.c_int, [_P]),
            "mastg_nastran_destroy": (C.c_int, [_P]),
            "mastg_design_create": (C.c_int, [C.POINTER(ProblemDesc), C.c_double, C.POINTER(_P)]),
            "mastg_design_destroy": (C.c_int, [_P]),
            "mastg_design_last_error": (C.c_char_p, [_P]),
            "mastg_design_info": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), _D]),
            "mastg_design_filter": (C.c_int, [_P, _D, _D, C.c_int, C.c_int]),
            "mastg_design_volume": (C.c_int, [_P, _D, C.c_int, _D, _D]),
            "mastg_design_oc_update": (C.c_int, [_P, _D, _D, _D, C.c_double, C.c_double, C.c_double,
                                                 _D, C.c_int, _D]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class MastgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mastg error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


def _p(a, t=C.c_double):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def desc_from_arrays(physics, xyz, conn, n_vars, node_dof_base=None, constrained_dofs=None,
                     sides=None, point_loads=None, **mat):
    """Build a ProblemDesc from numpy arrays; the returned object keeps them alive."""
    keep = {}

    def arr(name, a, dt):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dt)
        keep[name] = a
        return a

    d = ProblemDesc()
    d.physics = physics
    xyz = arr("xyz", xyz, np.float64)
    conn = arr("conn", conn, np.int32)
    d.n_nodes = xyz.shape[0]
    d.n_elems = conn.shape[0]
    d.xyz = _p(xyz)
    d.conn = _p(conn, C.c_int32)
    d.n_vars = n_vars
    nb = arr("nb", node_dof_base, np.int64)
    d.node_dof_base = _p(nb, C.c_int64)
    cd = arr("cd", constrained_dofs, np.int64)
    d.n_constrained = 0 if cd is None else cd.size
    d.constrained_dofs = _p(cd, C.c_int64)
    if sides is not None:
        se = arr("se", sides[0], np.int32); si = arr("si", sides[1], np.int32)
        sb = arr("sb", sides[2], np.int32)
        d.n_sides = se.size
        d.side_elem = _p(se, C.c_int32); d.side_index = _p(si, C.c_int32)
        d.side_id = _p(sb, C.c_int32)
    pl = arr("pl", point_loads, np.float64)
    d.point_loads = _p(pl)
    for k, v in mat.items():
        setattr(d, k, v)
    d._keep = keep
    return d


class Workload:
    """One of the structured benchmark problems, built by the library."""

    def __init__(self, kind, nx, ny=0, nz=0, seed=1, lo=0.0, hi=1.0, mast_exact=False):
        h = _P()
        rc = lib().mastg_workload_create(kind, nx, ny, nz, seed, lo, hi, int(mast_exact), C.byref(h))
        if rc:
            raise MastgError(rc, "workload_create")
        self.h = h
        self.nx, self.ny, self.nz = nx, ny, nz
        self.desc = lib().mastg_workload_desc(h)
        self.n_elems = self.desc.contents.n_elems
        self.params = np.ctypeslib.as_array(lib().mastg_workload_params(h), (self.n_elems,)).copy()

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.mastg_workload_destroy(self.h)
            self.h = None


class NastranModel:
    """A Nastran bulk-data model read by the library (mesh, DofMap, SPCs,
    FORCE loads, MAT1, PSHELL thicknesses)."""

    def __init__(self, path):
        h = _P()
        rc = lib().mastg_nastran_read(os.fsencode(path), C.byref(h))
        if rc:
            raise MastgError(rc, lib().mastg_nastran_last_error().decode())
        self.h = h
        self.desc = lib().mastg_nastran_desc(h)
        d = self.desc.contents
        self.n_nodes, self.n_elems = d.n_nodes, d.n_elems
        self.npe = 4 if d.physics == QUAD4_MINDLIN else 8
        self.params = np.ctypeslib.as_array(lib().mastg_nastran_params(h), (self.n_elems,)).copy()
        self.xyz = np.ctypeslib.as_array(d.xyz, (self.n_nodes * 3,)).reshape(-1, 3).copy()
        self.conn = np.ctypeslib.as_array(d.conn, (self.n_elems * self.npe,)).reshape(-1, self.npe).copy()
        self.node_base = np.ctypeslib.as_array(d.node_dof_base, (self.n_nodes,)).copy()
        self.n_dofs = int(self.node_base.max()) + d.n_vars if self.n_nodes else 0
        self.constrained = (np.ctypeslib.as_array(d.constrained_dofs, (d.n_constrained,)).copy()
                            if d.n_constrained else np.zeros(0, np.int64))
        self.point_loads = np.ctypeslib.as_array(d.point_loads, (self.n_dofs,)).copy()
        nid = np.zeros(self.n_nodes, np.int64)
        eid = np.zeros(self.n_elems, np.int64)
        pid = np.zeros(self.n_elems, np.int32)
        lib().mastg_nastran_ids(h, nid.ctypes.data_as(C.POINTER(C.c_int64)),
                                eid.ctypes.data_as(C.POINTER(C.c_int64)), pid.ctypes.data_as(C.POINTER(C.c_int32)))
        self.node_ids, self.elem_ids, self.elem_pids = nid, eid, pid
        self.spc_sets = lib().mastg_nastran_spc_sets(h)

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.mastg_nastran_destroy(self.h)
            self.h = None

    __del__ = close


class Context:
    def __init__(self, desc):
        h = _P()
        dp = desc if isinstance(desc, C.POINTER(ProblemDesc)) else C.byref(desc)
        rc = lib().mastg_ctx_create(dp, C.byref(h))
        if rc:
            raise MastgError(rc, lib().mastg_last_error(None).decode())
        self.h = h
        s = Sizes()
        self._chk(lib().mastg_get_sizes(h, C.byref(s)))
        self.sizes = s
        self.n_dofs, self.n_active, self.nnz = s.n_dofs, s.n_active, s.nnz
        self.n_elems = desc.contents.n_elems if isinstance(desc, C.POINTER(ProblemDesc)) else desc.n_elems

    @classmethod
    def _borrow(cls, h, n_elems):
        """A view of a context another object owns (a batch lane): closing it
        does not destroy the context."""
        ctx = cls.__new__(cls)
        ctx.h = h
        ctx._borrowed = True
        s = Sizes()
        ctx._chk(lib().mastg_get_sizes(h, C.byref(s)))
        ctx.sizes = s
        ctx.n_dofs, ctx.n_active, ctx.nnz = s.n_dofs, s.n_active, s.nnz
        ctx.n_elems = n_elems
        return ctx

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            if not getattr(self, "_borrowed", False):
                _lib.mastg_ctx_destroy(self.h)
            self.h = None

    __del__ = close

    def _chk(self, rc):
        if rc:
            raise MastgError(rc, lib().mastg_last_error(self.h).decode())
        return rc

    def csr_pattern(self):
        rp = np.zeros(self.n_active + 1, np.int64)
        col = np.zeros(self.nnz, np.int32)
        self._chk(lib().mastg_get_csr_pattern(self.h, _p(rp, C.c_int64), _p(col, C.c_int32)))
        return rp, col

    def active_map(self):
        a = np.zeros(self.n_dofs, np.int64)
        self._chk(lib().mastg_get_active_map(self.h, _p(a, C.c_int64)))
        return a

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float64)
        self._chk(lib().mastg_set_params(self.h, _p(p), HOST))

    def assemble(self, X=None, want_R=True, want_J=True, elem_precision=ELEM_FP64):
        R = np.zeros(self.n_dofs) if want_R else None
        J = np.zeros(self.nnz) if want_J else None
        Xa = None if X is None else np.ascontiguousarray(X, np.float64)
        self._chk(lib().mastg_assemble(self.h, _p(Xa), _p(R), _p(J), HOST, elem_precision))
        return R, J

    def solve(self, rhs, solver=SOLVER_CG_JACOBI, rtol=1e-10, max_iters=0, transpose=False,
              reliable_delta=0.0, check_every=0, raise_on_error=True):
        x = np.zeros(self.n_dofs)
        rhs = np.ascontiguousarray(rhs, np.float64)
        o = SolveOpts(solver, max_iters, rtol, reliable_delta, int(transpose), check_every)
        st = SolveStats()
        rc = lib().mastg_solve(self.h, _p(rhs), _p(x), HOST, C.byref(o), C.byref(st))
        if raise_on_error:
            self._chk(rc)
        return x, dict(iters=st.iters, replacements=st.replacements, rel_res=st.rel_res,
                       seconds=st.seconds, rc=rc)

    def jacobian_product(self, dX, op=0):
        y = np.zeros(self.n_dofs)
        dX = np.ascontiguousarray(dX, np.float64)
        self._chk(lib().mastg_jacobian_product(self.h, _p(dX), _p(y), HOST, op))
        return y

    def sensitivity(self, X, lam):
        g = np.zeros(self.n_elems)
        X = np.ascontiguousarray(X, np.float64)
        lam = np.ascontiguousarray(lam, np.float64)
        self._chk(lib().mastg_sensitivity(self.h, _p(X), _p(lam), _p(g), HOST))
        return g

    def sensitivity_rhs(self, param, X):
        """dR/dp at X for a global parameter (MASTG_PARAM_*)."""
        out = np.zeros(self.n_dofs)
        X = np.ascontiguousarray(X, np.float64)
        self._chk(lib().mastg_sensitivity_rhs(self.h, param, _p(X), _p(out), HOST))
        return out

    def sensitivity_global(self, param, X, lam):
        """lambda . dR/dp for a global parameter."""
        v = C.c_double()
        X = np.ascontiguousarray(X, np.float64)
        lam = np.ascontiguousarray(lam, np.float64)
        self._chk(lib().mastg_sensitivity_global(self.h, param, _p(X), _p(lam), HOST, C.byref(v)))
        return v.value

    def sensitivity_solve(self, param, X, solver=SOLVER_CG_JACOBI, rtol=1e-10, max_iters=0):
        """dX/dp from J dX/dp = -dR/dp (the Jacobian of the last assembly)."""
        out = np.zeros(self.n_dofs)
        X = np.ascontiguousarray(X, np.float64)
        o = SolveOpts(solver, max_iters, rtol, 0.0, 0, 0)
        st = SolveStats()
        self._chk(lib().mastg_sensitivity_solve(self.h, param, _p(X), _p(out), HOST, C.byref(o), C.byref(st)))
        return out, dict(iters=st.iters, rel_res=st.rel_res)

    def set_ks_stress(self, k, sigma_ref):
        self._chk(lib().mastg_set_ks_stress(self.h, k, sigma_ref))

    def set_stress_aggregation(self, p, p_weight, rho, sigma0):
        """StressStrainOutputBase::set_aggregation_coefficients(p1, p2, rho, sigma0)."""
        self._chk(lib().mastg_set_stress_aggregation(self.h, p, p_weight, rho, sigma0))

    def output_sensitivity(self, kind, X):
        """Direct dJ/dp_e of an output at fixed state X."""
        X = np.ascontiguousarray(X, np.float64)
        g = np.zeros(self.n_elems)
        self._chk(lib().mastg_output_sensitivity(self.h, kind, _p(X), HOST, _p(g)))
        return g

    def output(self, kind, X, want_grad=False):
        v = C.c_double()
        X = np.ascontiguousarray(X, np.float64)
        dv = np.zeros(self.n_dofs) if want_grad else None
        self._chk(lib().mastg_output(self.h, kind, _p(X), HOST, C.byref(v), _p(dv)))
        return (v.value, dv) if want_grad else v.value

    def design_eval(self, params, solver=SOLVER_CG_JACOBI, rtol=1e-10, max_iters=0,
                    reliable_delta=0.0, check_every=0, want_grad=True, raise_on_error=True):
        params = np.ascontiguousarray(params, np.float64)
        g = np.zeros(self.n_elems) if want_grad else None
        o = SolveOpts(solver, max_iters, rtol, reliable_delta, 0, check_every)
        r = DesignResult()
        rc = lib().mastg_design_eval(self.h, _p(params), _p(g), HOST, C.byref(o), C.byref(r))
        if raise_on_error:
            self._chk(rc)
        return dict(objective=r.objective, iters=r.iters, newton_iters=r.newton_iters,
                    rel_res=r.rel_res, t_assemble=r.t_assemble, t_solve=r.t_solve,
                    t_sens=r.t_sens, grad=g, rc=rc)

    def enable_timing(self, on=True):
        self._chk(lib().mastg_enable_kernel_timing(self.h, int(on)))

    def bench_kernel(self, kind, n_launch=20):
        ms = C.c_double()
        self._chk(lib().mastg_bench_kernel(self.h, kind, n_launch, C.byref(ms)))
        return ms.value

    def export_csr_full(self):
        """The MAST-exact 6-variable matrix of the last FP64 assembly:
        (row_ptr, col, val) in the caller's numbering."""
        n, nnz = C.c_int64(), C.c_int64()
        self._chk(lib().mastg_csr_full_size(self.h, C.byref(n), C.byref(nnz)))
        rp = np.zeros(n.value + 1, np.int64)
        col = np.zeros(nnz.value, np.int32)
        val = np.zeros(nnz.value, np.float64)
        self._chk(lib().mastg_export_csr_full(self.h, rp.ctypes.data, col.ctypes.data, val.ctypes.data))
        return rp, col, val

    def set_option(self, option, value):
        self._chk(lib().mastg_set_option(self.h, option, value))

    def stream(self):
        """The HIP stream (an integer handle) the context enqueues on."""
        h = C.c_void_p()
        self._chk(lib().mastg_get_stream(self.h, C.byref(h)))
        return h.value or 0

    def set_stream(self, handle):
        """Enqueue on the caller's HIP stream `handle` (0: the context's own)."""
        self._chk(lib().mastg_set_stream(self.h, C.c_void_p(handle or None)))

    def tuning(self):
        a, b = C.c_int(), C.c_int()
        self._chk(lib().mastg_get_tuning(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_tuning(self, lanes, grid):
        self._chk(lib().mastg_set_tuning(self.h, lanes, grid))

    def kernel_time(self, kind):
        ms = C.c_double()
        n = C.c_int64()
        self._chk(lib().mastg_kernel_time(self.h, kind, C.byref(ms), C.byref(n)))
        return ms.value, n.value


class Design:
    """Element-wise design space: density filter, volume, optimality-criteria step."""

    def __init__(self, desc, radius):
        h = _P()
        dp = desc if isinstance(desc, C.POINTER(ProblemDesc)) else C.byref(desc)
        rc = lib().mastg_design_create(dp, radius, C.byref(h))
        if rc:
            raise MastgError(rc, "design_create")
        self.h = h
        n, nnz, vt = C.c_int64(), C.c_int64(), C.c_double()
        lib().mastg_design_info(h, C.byref(n), C.byref(nnz), C.byref(vt))
        self.n, self.nnz, self.v_total = n.value, nnz.value, vt.value

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.mastg_design_destroy(self.h)
            self.h = None

    __del__ = close

    def _chk(self, rc):
        if rc:
            raise MastgError(rc, lib().mastg_design_last_error(self.h).decode())

    def filter(self, x, transpose=False):
        x = np.ascontiguousarray(x, np.float64)
        out = np.zeros(self.n)
        self._chk(lib().mastg_design_filter(self.h, _p(x), _p(out), int(transpose), HOST))
        return out

    def volume(self, rho, want_grad=False):
        rho = np.ascontiguousarray(rho, np.float64)
        v = C.c_double()
        dv = np.zeros(self.n) if want_grad else None
        self._chk(lib().mastg_design_volume(self.h, _p(rho), HOST, C.byref(v), _p(dv)))
        return (v.value, dv) if want_grad else v.value

    def oc_update(self, rho, dc, dv, volfrac, move=0.2, rho_min=1e-3):
        rho = np.ascontiguousarray(rho, np.float64)
        dc = np.ascontiguousarray(dc, np.float64)
        dv = np.ascontiguousarray(dv, np.float64)
        out = np.zeros(self.n)
        lam = C.c_double()
        self._chk(lib().mastg_design_oc_update(self.h, _p(rho), _p(dc), _p(dv), volfrac, move, rho_min,
                                               _p(out), HOST, C.byref(lam)))
        return out, lam.value


class Batch:
    """Batched design evaluations (mastg_batch_*): n_lanes contexts over one
    description, each on its own HIP stream and host thread; the designs of
    one eval() call are taken in turn by whichever lane is free."""

    def __init__(self, desc, n_lanes=2):
        h = _P()
        dp = desc if isinstance(desc, C.POINTER(ProblemDesc)) else C.byref(desc)
        rc = lib().mastg_batch_create(dp, n_lanes, C.byref(h))
        if rc:
            raise MastgError(rc, lib().mastg_last_error(None).decode())
        self.h = h
        self.n_lanes = n_lanes
        self.n_elems = desc.contents.n_elems if isinstance(desc, C.POINTER(ProblemDesc)) else desc.n_elems

    def _chk(self, rc):
        if rc:
            raise MastgError(rc, lib().mastg_batch_last_error(self.h).decode())
        return rc

    def lane(self, i):
        c = _P()
        self._chk(lib().mastg_batch_lane(self.h, i, C.byref(c)))
        return Context._borrow(c, self.n_elems)

    def set_option(self, option, value):
        self._chk(lib().mastg_batch_set_option(self.h, option, value))

    def eval(self, designs, solver=SOLVER_MG, rtol=1e-10, max_iters=0, reliable_delta=0.0, check_every=0,
             want_grad=True, raise_on_error=True):
        """designs: [n][n_elems]; a list of per-design result dicts (as
        Context.design_eval's, plus 'seconds', the design's wall time)."""
        P = np.ascontiguousarray(np.asarray(designs, np.float64).reshape(-1, self.n_elems))
        n = P.shape[0]
        G = np.zeros_like(P) if want_grad else None
        o = SolveOpts(solver, max_iters, rtol, reliable_delta, 0, check_every)
        out = (DesignResult * max(n, 1))()
        secs = np.zeros(max(n, 1))
        rc = lib().mastg_batch_eval(self.h, n, _p(P), _p(G), HOST, C.byref(o), out, _p(secs))
        if raise_on_error:
            self._chk(rc)
        return [dict(objective=r.objective, iters=r.iters, newton_iters=r.newton_iters, rel_res=r.rel_res,
                     t_assemble=r.t_assemble, t_solve=r.t_solve, t_sens=r.t_sens, seconds=float(secs[i]),
                     grad=None if G is None else G[i], rc=rc) for i, r in enumerate(out[:n])]

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.mastg_batch_destroy(self.h)
            self.h = None

    __del__ = close


class Server:
    """Request server (continuous batching of design evaluations)."""

    def __init__(self, n_slots=4, solver=SOLVER_MG, hbm_budget=0, queue_capacity=0, rtol=1e-10, rtol_plate=1e-7):
        h = _P()
        o = ServerOpts(n_slots, solver, hbm_budget, queue_capacity, rtol, rtol_plate)
        rc = lib().mastg_server_create(C.byref(o), C.byref(h))
        if rc:
            raise MastgError(rc, "server_create")
        self.h = h
        self._keep = {}

    def submit(self, kind, nx, ny, nz, params):
        params = np.ascontiguousarray(params, np.float64)
        grad = np.zeros(params.size)
        r = Request(kind, nx, ny, nz, _p(params), _p(grad))
        rid = C.c_int64()
        rc = lib().mastg_server_submit(self.h, C.byref(r), C.byref(rid))
        if rc:
            raise MastgError(rc, "server_submit")
        self._keep[rid.value] = (params, grad)
        return rid.value

    def _result(self, rid, res):
        params, grad = self._keep.pop(rid)
        return dict(status=res.status, objective=res.objective, iters=res.iters,
                    t_queue=res.t_queue, t_run=res.t_run, grad=grad)

    def poll(self, rid):
        res = RequestResult()
        rc = lib().mastg_server_poll(self.h, rid, C.byref(res))
        if rc:
            raise MastgError(rc, "server_poll")
        if res.status == 2 or res.status < 0:
            return self._result(rid, res)
        return None

    def wait(self, rid):
        res = RequestResult()
        rc = lib().mastg_server_wait(self.h, rid, C.byref(res))
        if rc:
            raise MastgError(rc, "server_wait")
        return self._result(rid, res)

    def last_error(self):
        return lib().mastg_server_last_error(self.h).decode()

    def stats(self):
        b, n, q = C.c_int64(), C.c_int32(), C.c_int32()
        lib().mastg_server_stats(self.h, C.byref(b), C.byref(n), C.byref(q))
        return dict(bytes_in_use=b.value, contexts=n.value, queued=q.value)

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.mastg_server_destroy(self.h)
            self.h = None

    __del__ = close


def request_footprint(kind, nx, ny, nz):
    r = Request(kind, nx, ny, nz, None, None)
    b = C.c_int64()
    rc = lib().mastg_request_footprint(C.byref(r), C.byref(b))
    if rc:
        raise MastgError(rc, "request_footprint")
    return b.value
