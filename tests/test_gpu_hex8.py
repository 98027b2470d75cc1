"""GPU parity of the hex8 elasticity path against the FP64 oracle, through the C ABI.

Bars: CSR row_ptr / col_idx bit-exact; K entries within 1e-12 relative to
max(|entry|, 1e-2 x row max) (entries that nearly cancel keep rounding of a few
ulp of the row scale whatever the FP64 summation order); R within 1e-12 of max|R|; sensitivity dots within 1e-12; solve-derived
compliance within 1e-9 and dC/drho within 1e-7 relative (both solves stop at
||r||/||b|| <= 1e-10, so solutions agree to the conditioning of K times that).
"""
import numpy as np
import pytest

from helpers import ctx_from_oracle, rel_err_floor, restrict_csr

pytestmark = pytest.mark.gpu

MESHES = [(1, 1, 1), (2, 1, 1), (3, 2, 1), (20, 10, 10)]


@pytest.mark.parametrize("mesh", MESHES)
@pytest.mark.parametrize("exact", [False, True])
def test_csr_pattern_and_K(oracle, mastg, mesh, exact):
    print(f"[hex8 K] mesh={mesh} mast_exact={exact}", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, *mesh, seed=11, lo=0.3, hi=1.0, mast_exact=exact)
    ctx = ctx_from_oracle(mastg, P)
    rp, col = ctx.csr_pattern()
    orp, ocol = P.csr()
    R0, K0 = P.assemble()
    if exact:
        base, _ = P.dofmap()
        keep = np.sort(np.concatenate([base + v for v in range(3)]))
        orp, ocol, K0 = restrict_csr(orp, ocol, K0, keep)
        amap = ctx.active_map()
        assert np.array_equal(np.nonzero(amap >= 0)[0], keep)
    assert np.array_equal(rp, orp), "row_ptr differs"
    assert np.array_equal(col, ocol), "col_idx differs"
    R, J = ctx.assemble()
    assert rel_err_floor(J, K0, orp, floor=1e-2) <= 1e-12
    np.testing.assert_allclose(R, R0, rtol=1e-12, atol=1e-12 * np.abs(R0).max())
    ctx.close()


@pytest.mark.parametrize("mesh", [(2, 1, 1), (6, 3, 3)])
def test_residual_at_state(oracle, mastg, mesh):
    print(f"[hex8 R(X)] mesh={mesh}", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, *mesh, seed=5, lo=0.01, hi=1.0, mast_exact=True)
    ctx = ctx_from_oracle(mastg, P)
    X = np.random.default_rng(3).standard_normal(P.n_dofs) * 1e-4
    _, con = P.dofmap()
    X[con == 1] = 0.0
    R0, _ = P.assemble(X)
    R, _ = ctx.assemble(X)
    base, _ = P.dofmap()
    act = np.sort(np.concatenate([base + v for v in range(3)]))
    assert np.max(np.abs(R - R0)[act]) <= 1e-12 * np.abs(R0).max()
    rot = np.setdiff1d(np.arange(P.n_dofs), act)
    assert np.all(R[rot] == 0.0)          # rotational DOFs carry no residual
    ctx.close()


def test_sensitivity_dots(oracle, mastg):
    print("[hex8 sens dots] 6x3x3", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, 6, 3, 3, seed=9, lo=0.01, hi=1.0)
    ctx = ctx_from_oracle(mastg, P)
    rng = np.random.default_rng(4)
    X = rng.standard_normal(P.n_dofs)
    lam = rng.standard_normal(P.n_dofs)
    g0 = P.sens_dots(X, lam)
    g = ctx.sensitivity(X, lam)
    assert rel_err_floor(g, g0, floor=1e-4) <= 1e-12
    ctx.close()


@pytest.mark.parametrize("mesh,lo", [((20, 10, 10), 0.3), ((16, 8, 8), 0.01)])
def test_design_eval_plain(oracle, mastg, mesh, lo):
    print(f"[hex8 design eval] mesh={mesh} lo={lo}", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, *mesh, seed=1, lo=lo, hi=1.0)
    ref = P.solve_elastic(1e-10)
    ctx = ctx_from_oracle(mastg, P)
    r = ctx.design_eval(P.params())
    assert r["rel_res"] <= 1.5e-10
    assert abs(r["objective"] - ref["compliance"]) <= 1e-9 * abs(ref["compliance"])
    assert rel_err_floor(r["grad"], ref["dc"], floor=1e-6) <= 1e-7
    # iteration counts of the two CG implementations agree closely
    assert abs(r["iters"] - ref["iters"]) <= max(3, 0.02 * ref["iters"])
    ctx.close()


def test_solve_and_output_api(oracle, mastg):
    print("[hex8 solve api] 8x4x4", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, 8, 4, 4, seed=2, lo=0.3, hi=1.0, mast_exact=True)
    ctx = ctx_from_oracle(mastg, P)
    R, J = ctx.assemble()
    x, st = ctx.solve(-R)
    ref = P.solve_elastic(1e-10)
    assert st["rel_res"] <= 1.5e-10
    # the reported residual is the true one: recompute it on the host
    import scipy.sparse as sp
    rp, col = ctx.csr_pattern()
    act = ctx.active_map()
    idx = np.nonzero(act >= 0)[0]
    A = sp.csr_matrix((J, col, rp), shape=(ctx.n_active, ctx.n_active))
    b = -R[idx]
    r_host = np.linalg.norm(b - A @ x[idx]) / np.linalg.norm(b)
    assert abs(r_host - st["rel_res"]) <= 1e-3 * st["rel_res"] + 1e-14
    assert st["iters"] > 10
    np.testing.assert_allclose(x, ref["u"], rtol=0, atol=1e-8 * np.abs(ref["u"]).max())
    C = ctx.output(mastg.OUTPUT_COMPLIANCE, x)
    assert abs(C - ref["compliance"]) <= 1e-9 * ref["compliance"]
    ctx.close()


def test_deterministic_assembly(oracle, mastg):
    print("[hex8 determinism] 10x5x5", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, 10, 5, 5, seed=3, lo=0.01, hi=1.0)
    ctx = ctx_from_oracle(mastg, P)
    _, J1 = ctx.assemble()
    _, J2 = ctx.assemble()
    assert np.array_equal(J1, J2)
    r1 = ctx.design_eval(P.params())
    r2 = ctx.design_eval(P.params())
    assert r1["objective"] == r2["objective"]
    assert np.array_equal(r1["grad"], r2["grad"])
    ctx.close()


def test_error_paths(oracle, mastg):
    print("[hex8 errors]", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, 4, 2, 2, seed=3, lo=0.3, hi=1.0)
    ctx = ctx_from_oracle(mastg, P)
    bad = P.params().copy()
    bad[3] = np.nan
    with pytest.raises(mastg.MastgError) as e:
        ctx.set_params(bad)
    assert e.value.code == -4
    ctx.set_params(P.params())
    r = ctx.design_eval(P.params(), max_iters=3, raise_on_error=False)
    assert r["rc"] == -3
    ctx.close()


@pytest.mark.parametrize("mesh", [(2, 1, 1), (20, 10, 10), (7, 5, 3)])
def test_matfree_operator_equals_csr(oracle, mastg, mesh):
    print(f"[hex8 matrix-free J dX] mesh={mesh}", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, *mesh, seed=4, lo=0.01, hi=1.0, mast_exact=True)
    ctx = ctx_from_oracle(mastg, P)
    R, J = ctx.assemble()
    dX = np.random.default_rng(7).standard_normal(P.n_dofs)
    y0 = ctx.jacobian_product(dX, op=0)
    y1 = ctx.jacobian_product(dX, op=1)
    # the oracle's own matrix-vector product (it keeps the rotational DOFs)
    P.assemble()
    yo = P.spmv(np.where(np.isin(np.arange(P.n_dofs), np.nonzero(ctx.active_map() >= 0)[0]), dX, 0.0))
    act = ctx.active_map() >= 0
    scale = np.abs(y0).max()
    assert np.max(np.abs(y0 - yo)[act]) <= 1e-12 * scale
    assert np.max(np.abs(y1 - y0)) <= 1e-12 * scale
    ctx.close()


@pytest.mark.parametrize("mesh,lo", [((20, 10, 10), 0.3), ((16, 8, 8), 0.01)])
def test_design_eval_matfree(oracle, mastg, mesh, lo):
    print(f"[hex8 design eval matrix-free] mesh={mesh} lo={lo}", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, *mesh, seed=1, lo=lo, hi=1.0)
    ref = P.solve_elastic(1e-10)
    ctx = ctx_from_oracle(mastg, P)
    r = ctx.design_eval(P.params(), solver=mastg.SOLVER_MATFREE)
    assert r["rel_res"] <= 1.5e-10
    assert abs(r["objective"] - ref["compliance"]) <= 1e-9 * abs(ref["compliance"])
    assert rel_err_floor(r["grad"], ref["dc"], floor=1e-6) <= 1e-7
    assert abs(r["iters"] - ref["iters"]) <= max(3, 0.02 * ref["iters"])
    ctx.close()


def test_autotuner_choice_is_valid(oracle, mastg):
    print("[autotune] 12x6x6", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, 12, 6, 6, seed=2, lo=0.3, hi=1.0)
    ctx = ctx_from_oracle(mastg, P)
    r0 = ctx.design_eval(P.params())
    lanes, grid = ctx.tuning()
    assert lanes in (4, 8, 16, 32, 64) and 1024 <= grid <= 8192
    # every launch shape gives the same answer to rounding
    for lanes, grid in [(4, 1024), (64, 8192)]:
        ctx.set_tuning(lanes, grid)
        r = ctx.design_eval(P.params())
        assert abs(r["objective"] - r0["objective"]) <= 1e-11 * r0["objective"]
    ctx.close()


def test_direct_launch_path_matches_graphs(oracle, mastg):
    print("[no graphs] 10x5x5", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, 10, 5, 5, seed=6, lo=0.01, hi=1.0)
    ctx = ctx_from_oracle(mastg, P)
    for sv in (mastg.SOLVER_CG_JACOBI, mastg.SOLVER_MATFREE):
        ctx.set_option(mastg.OPT_GRAPHS, 1)
        r1 = ctx.design_eval(P.params(), solver=sv)
        ctx.set_option(mastg.OPT_GRAPHS, 0)
        r2 = ctx.design_eval(P.params(), solver=sv)
        assert r1["iters"] == r2["iters"]
        assert r1["objective"] == r2["objective"]
        assert np.array_equal(r1["grad"], r2["grad"])
    ctx.close()


def distorted_cantilever(oracle, mesh, seed):
    """Oracle cantilever mesh with interior nodes moved by up to 15% of the
    element size (general trilinear geometry), 3 variables per node numbered
    node-major, no constraints; and the dense FP64 oracle matrix over it."""
    P = oracle.Problem(oracle.CANTILEVER, *mesh, seed=seed, lo=0.01, hi=1.0)
    xyz, conn = P.mesh()
    xyz = xyz.copy()
    h = np.array([1.0 / mesh[0], 0.5 / mesh[1], 0.5 / mesh[2]])
    lo, hi = xyz.min(0), xyz.max(0)
    inner = np.all((xyz > lo + 1e-9) & (xyz < hi - 1e-9), axis=1)
    rng = np.random.default_rng(seed)
    xyz[inner] += rng.uniform(-0.15, 0.15, (inner.sum(), 3)) * h
    rho = P.params()
    n = xyz.shape[0]
    K = np.zeros((3 * n, 3 * n))
    for e in range(conn.shape[0]):
        _, Ke = oracle.hex8_elastic(xyz[conn[e]], 72e9 * rho[e] ** 3, 0.33)
        dofs = np.array([3 * conn[e][a] + v for v in range(3) for a in range(8)])
        K[np.ix_(dofs, dofs)] += Ke[:24, :24]
    return xyz, conn, rho, K


@pytest.mark.parametrize("mesh", [(1, 1, 1), (3, 2, 2), (8, 4, 4)])
def test_fp32_matrix_core_elements(oracle, mastg, mesh):
    """FP32 element matrices on the matrix cores (v_mfma_f32_16x16x4f32),
    summed in FP64, against the FP64 oracle on distorted elements.  Bars: FP64
    path 1e-12 as above; FP32 path every entry within 1e-5 of its row's largest
    entry (operands rounded to FP32, 48-term FP32 dot products) and a median
    error below 1e-6 of it."""
    print(f"[hex8 FP32 MFMA K] mesh={mesh}", flush=True)
    xyz, conn, rho, K0 = distorted_cantilever(oracle, mesh, seed=7)
    n = xyz.shape[0]
    desc = mastg.desc_from_arrays(mastg.HEX8_ELASTIC, xyz, conn, 3, node_dof_base=3 * np.arange(n),
                                  E=72e9, nu=0.33, penalty=3.0)
    ctx = mastg.Context(desc)
    ctx.set_params(rho)
    rp, col = ctx.csr_pattern()
    rows = np.repeat(np.arange(3 * n), np.diff(rp))
    ref = K0[rows, col]
    _, J64 = ctx.assemble(want_R=False)
    assert rel_err_floor(J64, ref, rp, floor=1e-2) <= 1e-12
    _, J32 = ctx.assemble(want_R=False, elem_precision=mastg.ELEM_FP32)
    rowmax = np.maximum.reduceat(np.abs(ref), rp[:-1])[rows]
    err = np.abs(J32 - ref) / rowmax
    assert err.max() <= 1e-5, err.max()
    assert np.median(err) <= 1e-6
    ctx.close()


def test_fp32_jacobian_solve_and_limits(oracle, mastg):
    """A solve on the FP32-element matrix: compliance moves by the rounding of
    the element matrices (documented 1e-4..1e-3 on SIMP designs, bar 1e-2); the
    FP32 path is refused for physics without it."""
    print("[hex8 FP32 solve] 12x6x6", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, 12, 6, 6, seed=2, lo=0.01, hi=1.0)
    ref = P.solve_elastic(1e-10)
    ctx = ctx_from_oracle(mastg, P)
    R, _ = ctx.assemble(elem_precision=mastg.ELEM_FP32)
    x, st = ctx.solve(-R, rtol=1e-10)
    C32 = ctx.output(mastg.OUTPUT_COMPLIANCE, x)
    assert abs(C32 - ref["compliance"]) <= 1e-2 * ref["compliance"]
    ctx.close()
    Pp = oracle.Problem(oracle.PLATE, 8, 8, seed=1, lo=0.002, hi=0.01)
    cp = ctx_from_oracle(mastg, Pp)
    with pytest.raises(mastg.MastgError) as e:
        cp.assemble(elem_precision=mastg.ELEM_FP32)
    assert e.value.code == -1
    cp.close()


@pytest.mark.parametrize("mesh", [(2, 1, 1), (7, 5, 3)])
def test_fp32_node_block_operator(oracle, mastg, mesh):
    """The mixed solver's operator (FP32-rounded CSR values, read per node
    block, FP64 products): within FP32 rounding of the FP64 CSR product,
    |y32 - y| <= 1e-6 * (|A| |x|) entry-wise."""
    print(f"[hex8 FP32 node-block SpMV] mesh={mesh}", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, *mesh, seed=8, lo=0.01, hi=1.0)
    ctx = ctx_from_oracle(mastg, P)
    _, J = ctx.assemble()
    rp, col = ctx.csr_pattern()
    x = np.random.default_rng(2).standard_normal(P.n_dofs)
    amap = ctx.active_map()
    act = np.nonzero(amap >= 0)[0]
    y64 = ctx.jacobian_product(x, op=0)
    y32 = ctx.jacobian_product(x, op=2)
    xa = x[act]
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    absAx = np.zeros(rp.size - 1)
    np.add.at(absAx, rows, np.abs(J) * np.abs(xa[col]))
    assert np.all(np.abs(y32[act] - y64[act]) <= 1e-6 * absAx + 1e-300)
    assert np.max(np.abs(y32[act] - y64[act])) > 0.0     # it is the FP32 copy that is read
    ctx.close()


@pytest.mark.parametrize("mesh,lo", [((20, 10, 10), 0.3), ((16, 8, 8), 0.01)])
def test_design_eval_mixed(oracle, mastg, mesh, lo):
    """FP32 node-block SpMV inside CG, residual replaced by the FP64 true
    residual (reliable updates): the answer is fixed by the FP64 matrix, so the
    bars are the FP64 paths' (true ||b - A x||/||b|| <= 1e-10, compliance 1e-9,
    dC/drho 1e-7); the FP32 operator may cost extra iterations."""
    print(f"[hex8 design eval mixed] mesh={mesh} lo={lo}", flush=True)
    P = oracle.Problem(oracle.CANTILEVER, *mesh, seed=1, lo=lo, hi=1.0)
    ref = P.solve_elastic(1e-10)
    ctx = ctx_from_oracle(mastg, P)
    r = ctx.design_eval(P.params(), solver=mastg.SOLVER_MIXED)
    assert r["rel_res"] <= 1.5e-10
    assert abs(r["objective"] - ref["compliance"]) <= 1e-9 * abs(ref["compliance"])
    assert rel_err_floor(r["grad"], ref["dc"], floor=1e-6) <= 1e-7
    assert r["iters"] <= 3 * ref["iters"]
    R, _ = ctx.assemble()
    x, st = ctx.solve(-R, solver=mastg.SOLVER_MIXED)
    assert st["replacements"] >= 1 and st["rel_res"] <= 1.5e-10
    ctx.close()


@pytest.mark.parametrize("mesh,exact,distort", [((3, 2, 2), True, 0.0), ((12, 6, 6), False, 0.2),
                                                ((20, 10, 10), False, 0.0), ((9, 4, 3), False, -1.0)])
def test_row_assembly_equals_element_scatter(oracle, mastg, mesh, exact, distort):
    """The fused row-owner assembly (default: over the row plan where the
    elements fall into congruence classes) against the same kernel without the
    plan (MASTG_OPT_ROW_ASSEMBLY 2, bitwise), the element kernel + segmented
    scatter (MASTG_OPT_ROW_ASSEMBLY 0) and the oracle: J, R and the design
    evaluation.  distort < 0: cells of alternating widths along x and y (four
    congruence classes)."""
    print(f"[hex8 row assembly] mesh={mesh} mast_exact={exact} distortion={distort}", flush=True)
    P0 = oracle.Problem(oracle.CANTILEVER, *mesh, seed=21, lo=0.05, hi=1.0, mast_exact=exact)
    xyz, conn = P0.mesh()
    if distort > 0:
        lo, hi = xyz.min(0), xyz.max(0)
        inner = np.all((xyz > lo + 1e-12) & (xyz < hi - 1e-12), axis=1)
        xyz = xyz.copy()
        xyz[inner] += distort / mesh[0] * np.random.default_rng(2).uniform(-1, 1, (int(inner.sum()), 3))
    elif distort < 0:
        xyz = xyz.copy()
        for d in (0, 1):
            u = np.unique(xyz[:, d])
            i = np.searchsorted(u, xyz[:, d])
            h = u[1] - u[0]
            xyz[:, d] = i * h + 0.3 * h * (i % 2)      # widths 1.3 h and 0.7 h in turn
    base, con = P0.dofmap()
    arrays = dict(xyz=xyz, conn=conn, n_vars=P0.n_vars, node_base=base, constrained=con, params=P0.params(),
                  point_load=P0.point_load())
    P = oracle.Problem(oracle.CANTILEVER, arrays=arrays)
    desc = mastg.desc_from_arrays(mastg.HEX8_ELASTIC, xyz, conn, P0.n_vars, node_dof_base=base,
                                  constrained_dofs=np.nonzero(con)[0], point_loads=P0.point_load(),
                                  E=72e9, nu=0.33, penalty=3.0)
    ctx = mastg.Context(desc)
    ctx.set_params(P0.params())
    X = np.random.default_rng(4).standard_normal(P.n_dofs) * 1e-4
    X[con == 1] = 0.0
    R1, J1 = ctx.assemble(X)
    ctx.set_option(mastg.OPT_ROW_ASSEMBLY, 2)
    R2, J2 = ctx.assemble(X)
    assert np.array_equal(J1, J2)
    assert np.max(np.abs(R1 - R2)) <= 1e-14 * np.abs(R2).max()
    ctx.set_option(mastg.OPT_ROW_ASSEMBLY, 0)
    R0, J0 = ctx.assemble(X)
    rp, _ = ctx.csr_pattern()
    assert rel_err_floor(J1, J0, rp, floor=1e-2) <= 1e-12
    assert np.max(np.abs(R1 - R0)) <= 1e-12 * np.abs(R0).max()
    # and against the oracle
    orp, ocol = P.csr()
    Ro, Ko = P.assemble(X)
    if exact:
        keep = np.sort(np.concatenate([base + v for v in range(3)]))
        orp, ocol, Ko = restrict_csr(orp, ocol, Ko, keep)
        act = keep
    else:
        act = np.arange(P.n_dofs)
    assert rel_err_floor(J1, Ko, orp, floor=1e-2) <= 1e-12
    assert np.max(np.abs(R1 - Ro)[act]) <= 1e-12 * np.abs(Ro[act]).max()
    ctx.set_option(mastg.OPT_ROW_ASSEMBLY, 1)
    ref = P.solve_elastic(1e-10)
    r = ctx.design_eval(P0.params())
    assert abs(r["objective"] - ref["compliance"]) <= 1e-9 * ref["compliance"]
    assert rel_err_floor(r["grad"], ref["dc"], floor=1e-6) <= 1e-7
    ctx.close()
