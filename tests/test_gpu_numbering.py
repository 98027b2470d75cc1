"""GPU parity of the global assembly under an arbitrary caller numbering.

The expected matrix is built here by brute force, independently of the
oracle's sparsity and scatter code: a dense global matrix over every caller
DOF, filled element by element from the oracle's element matrices (pinned by
the reference's goldens and closed forms), with libMesh's homogeneous
Dirichlet rule applied per element before the add
(DofMap::constrain_element_matrix_and_vector, libMesh 1.5.1
src/base/dof_map_constraints.C: K <- C^T K C zeroes the constrained row and
column, then K(i,i) = 1 and f(i) = 0), and the sparsity pattern taken as the
set of (i, j) pairs that share an element (libMesh SparsityPattern::Build
with full variable coupling).  The node DOF bases are a random permutation,
not libMesh's first-touch order, so the test pins that the library keeps the
caller's numbering whatever it is
(src/base/nonlinear_implicit_assembly.cpp:111-183 adds element matrices at
the dof_indices the DofMap hands out).

Bars: row_ptr / col_idx bit-exact; values (constrained rows included) within
1e-12 of the row's largest entry; residuals within 1e-12 of max |R|.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

E, NU = 72e9, 0.33
PE, PNU, KAPPA = 70e9, 0.3, 5.0 / 6.0


def _element_K(oracle, kind, X, p, T=None, sides=()):
    """Element matrix over the caller's variables (var-major), its residual
    function, for one element."""
    if kind == oracle.CANTILEVER:
        Ee = p ** 3 * E
        _, K = oracle.hex8_elastic(X, Ee, NU)
        return K
    if kind == oracle.PLATE:
        _, K = oracle.quad4_mindlin(X, PE, PNU, p, 0.0, KAPPA)
        return K
    _, K = oracle.hex8_heat(X, 200.0, 1e-3, T)
    for s in sides:
        _, Kc = oracle.hex8_convection(X, s, 50.0, 300.0, T)
        K = K + Kc
    return K


def _element_R(oracle, kind, X, p, u, sides=()):
    if kind == oracle.CANTILEVER:
        f, _ = oracle.hex8_elastic(X, p ** 3 * E, NU, u)
        return f
    if kind == oracle.PLATE:
        f, _ = oracle.quad4_mindlin(X, PE, PNU, p, 0.0, KAPPA, u)
        return f
    f, _ = oracle.hex8_heat(X, 200.0, 1e-3, u)
    for s in sides:
        fc, _ = oracle.hex8_convection(X, s, 50.0, 300.0, u)
        f = f + fc
    return f


def _brute_force(oracle, P, base, con, params, X, nav):
    """Dense global K (at state X for heat), R at X, pattern; caller numbering."""
    xyz, conn = P.mesh()
    npe, nv = P.npe, P.n_vars
    n = int(P.n_dofs)
    K = np.zeros((n, n))
    R = np.zeros(n)
    pat = np.zeros((n, n), bool)
    be, bs, bid = P.bsides()
    for e in range(P.n_elems):
        Xe = xyz[conn[e]]
        dofs = np.array([base[conn[e, a]] + v for v in range(nv) for a in range(npe)])
        ue = X[dofs]
        sides = [int(s) for s, b, ee in zip(bs, bid, be) if ee == e and ((0x3E >> b) & 1)] \
            if P.kind == oracle.HEAT else []
        Ke = _element_K(oracle, P.kind, Xe, params[e], ue, sides).copy()
        fe = _element_R(oracle, P.kind, Xe, params[e], ue, sides).copy()
        for i, d in enumerate(dofs):
            if con[d]:
                Ke[i, :] = 0.0
                Ke[:, i] = 0.0
                Ke[i, i] = 1.0
                fe[i] = 0.0
        K[np.ix_(dofs, dofs)] += Ke
        R[dofs] += fe
        pat[np.ix_(dofs, dofs)] = True
    # the active rows / columns: the first nav variables of every node, in
    # caller order
    act = np.sort(np.concatenate([base + v for v in range(nav)]))
    Ka = K[np.ix_(act, act)]
    pa = pat[np.ix_(act, act)]
    rp = np.zeros(act.size + 1, np.int64)
    cols, vals = [], []
    for i in range(act.size):
        c = np.nonzero(pa[i])[0]
        cols.append(c)
        vals.append(Ka[i, c])
        rp[i + 1] = rp[i] + c.size
    return rp, np.concatenate(cols).astype(np.int32), np.concatenate(vals), R, act


CASES = [("hex8", (2, 1, 1)), ("hex8", (2, 2, 1)), ("hex8", (3, 2, 2)),
         ("plate", (2, 1)), ("plate", (3, 2)), ("plate", (3, 3)),
         ("heat", (2,)), ("heat", (3,))]


@pytest.mark.parametrize("kind,mesh", CASES)
def test_permuted_numbering_bruteforce(oracle, mastg, kind, mesh):
    print(f"[numbering] {kind} mesh={mesh}", flush=True)
    K = {"hex8": oracle.CANTILEVER, "plate": oracle.PLATE, "heat": oracle.HEAT}[kind]
    P = oracle.Problem(K, *mesh, seed=17, lo=0.3, hi=1.0, mast_exact=(kind != "heat"))
    xyz, conn = P.mesh()
    base0, con0 = P.dofmap()
    nv = P.n_vars
    nav = {"hex8": 3, "plate": 5, "heat": 1}[kind]
    rng = np.random.default_rng(hash(mesh) % 1000 + nav)
    perm = rng.permutation(P.n_nodes)
    base = (perm * nv).astype(np.int64)
    # the same constrained (node, variable) pairs in the new numbering
    con = np.zeros(P.n_dofs, np.uint8)
    for nd in range(P.n_nodes):
        for v in range(nv):
            con[base[nd] + v] = con0[base0[nd] + v]
    params = P.params()
    if kind == "heat":
        params = np.zeros(P.n_elems)       # no source: R is internal + convection
    X = rng.standard_normal(P.n_dofs) * (1e-4 if kind != "heat" else 30.0)
    X[con == 1] = 0.0
    erp, ecol, eval_, eR, act = _brute_force(oracle, P, base, con, params, X, nav)

    if kind == "hex8":
        mat = dict(E=E, nu=NU, penalty=3.0)
        phys = mastg.HEX8_ELASTIC
    elif kind == "plate":
        mat = dict(E=PE, nu=PNU, kappa=KAPPA, offset=0.0, pressure=0.0)
        phys = mastg.QUAD4_MINDLIN
    else:
        mat = dict(k0=200.0, alpha=1e-3, conv_h=50.0, conv_tinf=300.0, conv_bid_mask=0x3E)
        phys = mastg.HEX8_HEAT
    desc = mastg.desc_from_arrays(phys, xyz, conn, nv, node_dof_base=base,
                                  constrained_dofs=np.nonzero(con)[0], sides=P.bsides(), **mat)
    ctx = mastg.Context(desc)
    ctx.set_params(params)
    amap = ctx.active_map()
    assert np.array_equal(np.nonzero(amap >= 0)[0], act)
    assert np.array_equal(amap[act], np.arange(act.size))
    rp, col = ctx.csr_pattern()
    assert np.array_equal(rp, erp), "row_ptr differs from the brute-force pattern"
    assert np.array_equal(col, ecol), "col_idx differs from the brute-force pattern"
    R, J = ctx.assemble(X)
    rows = np.repeat(np.arange(act.size), np.diff(erp))
    rowmax = np.maximum.reduceat(np.abs(eval_), erp[:-1])
    err = np.abs(J - eval_) / rowmax[rows]
    assert err.max() <= 1e-12, f"K differs: {err.max():.3e}"
    # constrained rows: exactly the unit diagonal per element
    for i in np.nonzero(con[act])[0]:
        seg = slice(erp[i], erp[i + 1])
        d = np.nonzero(ecol[seg] == i)[0][0]
        assert J[seg][d] == eval_[seg][d] and float(J[seg][d]).is_integer()
        assert np.all(np.delete(J[seg], d) == 0.0)
    # residual (caller numbering; inactive variables carry none)
    assert np.max(np.abs(R[act] - eR[act])) <= 1e-12 * np.abs(eR[act]).max()
    ctx.close()
