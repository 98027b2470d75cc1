"""The context on a caller's HIP stream (mastg_get_stream / mastg_set_stream):
the results are bitwise those on the context's own stream, its work lands on
the caller's stream in order with the caller's own, and a null stream
restores the context's own.  The caller's stream is made with the HIP
runtime the library itself uses."""
import ctypes as C

import numpy as np
import pytest

from helpers import ctx_from_oracle

pytestmark = pytest.mark.gpu


def _hip():
    # the runtime libmastg.so is linked against (already loaded with it)
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipStreamCreateWithFlags.argtypes = [C.POINTER(C.c_void_p), C.c_uint]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    hip.hipMemsetAsync.argtypes = [C.c_void_p, C.c_int, C.c_size_t, C.c_void_p]
    return hip


def test_context_on_caller_stream(oracle, mastg):
    print("[stream] hex8 8x4x4 on a caller's stream", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, 8, 4, 4, seed=3, lo=0.3, hi=1.0)
    ctx = ctx_from_oracle(mastg, P)
    own = ctx.stream()
    assert own != 0
    r0 = ctx.design_eval(P.params(), solver=mastg.SOLVER_MG)
    hip = _hip()
    s = C.c_void_p()
    assert hip.hipStreamCreateWithFlags(C.byref(s), 1) == 0   # non-blocking, default priority
    L = mastg.lib()
    dx, dv = C.c_void_p(), C.c_void_p()
    try:
        ctx.set_stream(s.value)
        assert ctx.stream() == s.value
        r1 = ctx.design_eval(P.params(), solver=mastg.SOLVER_MG)
        assert r1["objective"] == r0["objective"] and np.array_equal(r1["grad"], r0["grad"])
        # the caller's work and the context's on one stream, in order: the
        # caller clears the value, the allocation-free output then writes it
        X = np.random.default_rng(1).standard_normal(P.n_dofs) * 1e-4
        ctx._chk(L.mastg_device_alloc(ctx.h, X.nbytes, C.byref(dx)))
        ctx._chk(L.mastg_device_alloc(ctx.h, 8, C.byref(dv)))
        ctx._chk(L.mastg_copy(ctx.h, dx, X.ctypes.data, X.nbytes, 0))
        assert hip.hipMemsetAsync(dv, 0xff, 8, s) == 0
        ctx._chk(L.mastg_output_noalloc(ctx.h, mastg.OUTPUT_COMPLIANCE, dx, dv))
        assert hip.hipStreamSynchronize(s) == 0
        v = np.zeros(1)
        ctx._chk(L.mastg_copy(ctx.h, v.ctypes.data, dv, 8, 1))
        ref = ctx.output(mastg.OUTPUT_COMPLIANCE, X)
        assert abs(v[0] - ref) <= 1e-12 * abs(ref)
        ctx.set_stream(0)
        assert ctx.stream() == own
        r2 = ctx.design_eval(P.params(), solver=mastg.SOLVER_MG)
        assert r2["objective"] == r0["objective"]
    finally:
        if dx.value:
            L.mastg_device_free(ctx.h, dx)
        if dv.value:
            L.mastg_device_free(ctx.h, dv)
        ctx.set_stream(0)
        ctx.close()
        hip.hipStreamDestroy(s)
