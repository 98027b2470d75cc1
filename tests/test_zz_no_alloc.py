"""The allocation-free output entry point (mastg_output_noalloc) against
mastg_output, through the C ABI: device input and output pointers, the
reduction's partial sums in the library's file-scope scratch array.  Two
calls in a row, the second on a smaller problem (fewer partial sums than the
first left in the shared scratch), then a mean-temperature call; each value
within 1e-12 (relative) of mastg_output's, the FP64 bar of the outputs."""
import ctypes as C

import numpy as np
import pytest

from helpers import ctx_from_oracle

pytestmark = pytest.mark.gpu


def _device_output(mastg, ctx, kind, X):
    L = mastg.lib()
    X = np.ascontiguousarray(X, np.float64)
    dX, dv = C.c_void_p(), C.c_void_p()
    ctx._chk(L.mastg_device_alloc(ctx.h, X.nbytes, C.byref(dX)))
    ctx._chk(L.mastg_device_alloc(ctx.h, 8, C.byref(dv)))
    try:
        ctx._chk(L.mastg_copy(ctx.h, dX, X.ctypes.data, X.nbytes, 0))
        ctx._chk(L.mastg_output_noalloc(ctx.h, kind, dX, dv))
        ctx._chk(L.mastg_synchronize(ctx.h))
        out = np.zeros(1)
        ctx._chk(L.mastg_copy(ctx.h, out.ctypes.data, dv, 8, 1))
        return float(out[0])
    finally:
        L.mastg_device_free(ctx.h, dX)
        L.mastg_device_free(ctx.h, dv)


def test_output_noalloc_matches_output(oracle, mastg):
    cases = [(oracle.CANTILEVER, (24, 12, 12)), (oracle.CANTILEVER, (3, 2, 2))]
    for kind, mesh in cases:
        print(f"[no-alloc output] compliance, hex8 {mesh}", flush=True)
        P = oracle.Problem(kind, *mesh, seed=3, lo=0.3, hi=1.0)
        ctx = ctx_from_oracle(mastg, P)
        X = np.random.default_rng(len(mesh) + mesh[0]).standard_normal(P.n_dofs)
        ref = ctx.output(mastg.OUTPUT_COMPLIANCE, X)
        got = _device_output(mastg, ctx, mastg.OUTPUT_COMPLIANCE, X)
        print(f"  {got!r} vs {ref!r}", flush=True)
        assert abs(got - ref) <= 1e-12 * abs(ref)
        ctx.close()
    print("[no-alloc output] mean temperature, heat 5^3", flush=True)
    P = oracle.Problem(oracle.HEAT, 5, seed=2)
    ctx = ctx_from_oracle(mastg, P)
    T = 300 + np.random.default_rng(9).standard_normal(P.n_dofs)
    ref = ctx.output(mastg.OUTPUT_MEAN_T, T)
    got = _device_output(mastg, ctx, mastg.OUTPUT_MEAN_T, T)
    assert abs(got - ref) <= 1e-12 * abs(ref)
    ctx.close()
