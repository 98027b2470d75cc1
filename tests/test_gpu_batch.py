"""Batched design evaluations (mastg_batch_*) through the C ABI.

Several designs of one mesh on n lanes (contexts with their own streams and
host threads) overlap on the device.  Bars: every design's compliance and
gradient bitwise equal to a lone context's with the same launch shapes,
whatever lane takes it and whatever runs beside it (more designs than lanes,
two calls in a row); the first design against the oracle (compliance 1e-9,
dC/drho 1e-7); the plate with CG + Jacobi (launch shapes tuned on lane 0 and
shared); an invalid design fails alone, with its index in the message, while
the others complete.
"""
import numpy as np
import pytest

from helpers import rel_err_floor

pytestmark = pytest.mark.gpu


def _desc(oracle, mastg, P):
    xyz, conn = P.mesh()
    base, con = P.dofmap()
    if P.kind == oracle.CANTILEVER:
        return mastg.desc_from_arrays(mastg.HEX8_ELASTIC, xyz, conn, P.n_vars, node_dof_base=base,
                                      constrained_dofs=np.nonzero(con)[0], point_loads=P.point_load(),
                                      E=72e9, nu=0.33, penalty=3.0)
    return mastg.desc_from_arrays(mastg.QUAD4_MINDLIN, xyz, conn, P.n_vars, node_dof_base=base,
                                  constrained_dofs=np.nonzero(con)[0], E=70e9, nu=0.3, kappa=5.0 / 6.0,
                                  offset=0.0, pressure=1.0e3)


def _designs(P, n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    return [P.params()] + [lo + (hi - lo) * rng.random(P.n_elems) for _ in range(n - 1)]


def test_batch_hex8_multigrid(oracle, mastg):
    print("[batch] hex8 16x8x8, 3 lanes, 7 designs, multigrid", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, 16, 8, 8, seed=4, lo=0.01, hi=1.0)
    desc = _desc(oracle, mastg, P)
    designs = _designs(P, 7, 0.01, 1.0, 11)
    B = mastg.Batch(desc, 3)
    first = B.eval(designs[:2], solver=mastg.SOLVER_MG)
    res = B.eval(designs, solver=mastg.SOLVER_MG)
    assert all(r["rc"] == 0 and r["seconds"] > 0 for r in res)
    lanes = B.lane(0).tuning()
    ctx = mastg.Context(desc)
    ctx.set_tuning(*lanes)
    for i, d in enumerate(designs):
        r = ctx.design_eval(d, solver=mastg.SOLVER_MG)
        assert r["objective"] == res[i]["objective"] and np.array_equal(r["grad"], res[i]["grad"]), i
        if i < 2:
            assert first[i]["objective"] == r["objective"] and np.array_equal(first[i]["grad"], r["grad"])
    ref = P.solve_elastic(1e-12)
    ec = abs(res[0]["objective"] - ref["compliance"]) / abs(ref["compliance"])
    eg = rel_err_floor(res[0]["grad"], ref["dc"], floor=1e-6)
    print(f"  design 0 against the oracle: C {ec:.1e}, grad {eg:.1e}", flush=True)
    assert ec <= 1e-9 and eg <= 1e-7
    ctx.close()
    B.close()


def test_batch_plate_cg(oracle, mastg):
    print("[batch] plate 12x12, 2 lanes, 5 designs, CG + Jacobi", flush=True)
    P = oracle.Problem(oracle.PLATE, 12, 12, seed=2)
    desc = _desc(oracle, mastg, P)
    designs = _designs(P, 5, 0.002, 0.01, 5)
    B = mastg.Batch(desc, 2)
    res = B.eval(designs, solver=mastg.SOLVER_CG_JACOBI)
    ctx = mastg.Context(desc)
    ctx.set_tuning(*B.lane(1).tuning())
    for i, d in enumerate(designs):
        r = ctx.design_eval(d, solver=mastg.SOLVER_CG_JACOBI)
        assert r["objective"] == res[i]["objective"] and np.array_equal(r["grad"], res[i]["grad"]), i
    ref = P.solve_elastic(1e-10)
    assert abs(res[0]["objective"] - ref["compliance"]) <= 1e-9 * abs(ref["compliance"])
    ctx.close()
    B.close()


def test_batch_invalid_design(oracle, mastg):
    print("[batch] one invalid design among four", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, 6, 3, 3, seed=1, lo=0.3, hi=1.0)
    desc = _desc(oracle, mastg, P)
    designs = _designs(P, 4, 0.3, 1.0, 3)
    designs[2] = designs[2].copy()
    designs[2][5] = np.nan
    B = mastg.Batch(desc, 2)
    with pytest.raises(mastg.MastgError) as e:
        B.eval(designs, solver=mastg.SOLVER_CG_JACOBI)
    assert "design 2" in str(e.value)
    res = B.eval(designs, solver=mastg.SOLVER_CG_JACOBI, raise_on_error=False)
    ctx = mastg.Context(desc)
    ctx.set_tuning(*B.lane(0).tuning())
    for i in (0, 1, 3):
        r = ctx.design_eval(designs[i], solver=mastg.SOLVER_CG_JACOBI)
        assert r["objective"] == res[i]["objective"], i
    ctx.close()
    B.close()


def test_batch_refusals(mastg):
    w = mastg.Workload(1, 4, 2, 2, seed=1)
    for lanes in (0, 17):
        with pytest.raises(mastg.MastgError):
            mastg.Batch(w.desc, lanes)


def test_batch_device_pointers(oracle, mastg):
    """params and gradients in device memory (where = MASTG_DEVICE) give the
    host-pointer results bitwise."""
    import ctypes as C
    print("[batch] device pointers, hex8 8x4x4, 2 lanes, 3 designs", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, 8, 4, 4, seed=2, lo=0.1, hi=1.0)
    desc = _desc(oracle, mastg, P)
    designs = np.ascontiguousarray(np.stack(_designs(P, 3, 0.1, 1.0, 8)))
    B = mastg.Batch(desc, 2)
    host = B.eval(designs, solver=mastg.SOLVER_MG)
    lane = B.lane(0)
    L = mastg.lib()
    nbytes = designs.nbytes
    dp, dg = C.c_void_p(), C.c_void_p()
    lane._chk(L.mastg_device_alloc(lane.h, nbytes, C.byref(dp)))
    lane._chk(L.mastg_device_alloc(lane.h, nbytes, C.byref(dg)))
    try:
        lane._chk(L.mastg_copy(lane.h, dp, designs.ctypes.data, nbytes, 0))
        o = mastg.SolveOpts(mastg.SOLVER_MG, 0, 1e-10, 0.0, 0, 0)
        out = (mastg.DesignResult * 3)()
        rc = L.mastg_batch_eval(B.h, 3, C.cast(dp, C.POINTER(C.c_double)), C.cast(dg, C.POINTER(C.c_double)),
                                mastg.DEVICE, C.byref(o), out, None)
        B._chk(rc)
        g = np.zeros_like(designs)
        lane._chk(L.mastg_copy(lane.h, g.ctypes.data, dg, nbytes, 1))
    finally:
        L.mastg_device_free(lane.h, dp)
        L.mastg_device_free(lane.h, dg)
    for i in range(3):
        assert out[i].objective == host[i]["objective"] and np.array_equal(g[i], host[i]["grad"]), i
    B.close()
