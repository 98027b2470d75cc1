"""Request server (continuous batching) on the GPU.

* each request's result is the same alone and in a concurrent mix of meshes and
  physics (deterministic kernels with fixed launch shapes: bitwise equal);
* a served request equals a direct design evaluation on a context of the same
  mesh (same operator and launch shapes);
* admission by device footprint: a request larger than the budget is refused,
  and a budget that holds one mesh at a time still completes a mixed stream by
  evicting idle contexts;
* invalid design parameters are rejected at submit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def params_for(oracle, kind, nx, ny, nz, seed):
    n = nx * ny * nz if kind == 1 else nx * ny
    lo, hi = (0.01, 1.0) if kind == 1 else (0.002, 0.01)
    return oracle.uniform(seed, n, lo, hi)


MIX = [(1, 16, 8, 8), (1, 24, 12, 12), (2, 32, 32, 0), (1, 32, 16, 16), (1, 16, 8, 8),
       (2, 24, 24, 0), (1, 24, 12, 12), (1, 48, 24, 24), (1, 16, 8, 8), (2, 32, 32, 0)]


def test_alone_equals_in_mix(oracle, mastg):
    print("[server] alone vs mix", flush=True)
    target = (1, 24, 12, 12)
    p = params_for(oracle, *target, seed=99)
    s = mastg.Server(n_slots=1)
    alone = s.wait(s.submit(*target, p))
    s.close()
    s = mastg.Server(n_slots=4)
    ids = []
    for i, m in enumerate(MIX):
        ids.append(s.submit(*m, params_for(oracle, *m, seed=i)))
        if i == 4:
            tid = s.submit(*target, p)
    outs = [s.wait(i) for i in ids]
    mixed = s.wait(tid)
    err = s.last_error()
    s.close()
    assert alone["status"] == 2 and mixed["status"] == 2, err
    assert all(o["status"] == 2 for o in outs)
    assert mixed["objective"] == alone["objective"]
    assert np.array_equal(mixed["grad"], alone["grad"])


def test_served_equals_direct_context(oracle, mastg):
    print("[server] served vs direct", flush=True)
    w = mastg.Workload(1, 16, 8, 8, seed=4, lo=0.01, hi=1.0)
    ctx = mastg.Context(w.desc)
    ctx.set_tuning(8, 2048)
    r = ctx.design_eval(w.params, solver=mastg.SOLVER_MG)
    ctx.close()
    s = mastg.Server(n_slots=2)
    out = s.wait(s.submit(1, 16, 8, 8, w.params))
    s.close()
    assert out["objective"] == r["objective"]
    assert np.array_equal(out["grad"], r["grad"])


def test_admission_and_eviction(oracle, mastg):
    print("[server] admission", flush=True)
    big = mastg.request_footprint(1, 48, 24, 24)
    small = mastg.request_footprint(1, 16, 8, 8)
    s = mastg.Server(n_slots=2, hbm_budget=small + small // 2)
    with pytest.raises(mastg.MastgError) as e:
        s.submit(1, 48, 24, 24, params_for(oracle, 1, 48, 24, 24, 1))
    assert e.value.code == -6
    s.close()
    # room for one of the two meshes at a time: requests wait, idle contexts are evicted
    s = mastg.Server(n_slots=2, hbm_budget=big + small // 2)
    ids = [s.submit(*m, params_for(oracle, *m, seed=i))
           for i, m in enumerate([(1, 48, 24, 24), (1, 16, 8, 8), (1, 48, 24, 24), (1, 16, 8, 8)])]
    outs = [s.wait(i) for i in ids]
    st = s.stats()
    err = s.last_error()
    s.close()
    assert all(o["status"] == 2 for o in outs), ([o["status"] for o in outs], err)
    assert st["bytes_in_use"] <= big + small // 2


def test_invalid_params_rejected(mastg):
    s = mastg.Server(n_slots=1)
    p = np.full(16 * 8 * 8, 0.5)
    p[3] = np.nan
    with pytest.raises(mastg.MastgError) as e:
        s.submit(1, 16, 8, 8, p)
    assert e.value.code == -4
    s.close()
