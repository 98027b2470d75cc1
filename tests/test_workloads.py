"""The library's benchmark workloads (host code, no GPU needed) reproduce the
oracle's problems exactly: coordinates, connectivity, DOF bases, Dirichlet
set, nodal loads and the seeded design parameters.  Also checks that the C ABI
library loads and exports every entry point declared in include/mastg/mastg.h.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def arrays(M, w, npe):
    d = w.desc.contents
    nn, ne = d.n_nodes, d.n_elems
    xyz = np.ctypeslib.as_array(d.xyz, (nn * 3,)).reshape(nn, 3).copy()
    conn = np.ctypeslib.as_array(d.conn, (ne * npe,)).reshape(ne, npe).copy()
    base = np.ctypeslib.as_array(d.node_dof_base, (nn,)).copy()
    con = np.ctypeslib.as_array(d.constrained_dofs, (d.n_constrained,)).copy()
    ndofs = int(base.max()) + d.n_vars
    loads = (np.ctypeslib.as_array(d.point_loads, (ndofs,)).copy()
             if d.point_loads else np.zeros(ndofs))
    return xyz, conn, base, np.sort(con), loads, w.params


@pytest.mark.parametrize("kind,dims,npe", [(1, (20, 10, 10), 8), (1, (7, 3, 2), 8),
                                           (2, (12, 9, 0), 4), (3, (5, 5, 5), 8)])
@pytest.mark.parametrize("exact", [False, True])
def test_workload_matches_oracle(oracle, mastg, kind, dims, npe, exact):
    if kind == 3 and exact:
        pytest.skip("heat has one variable")
    nx, ny, nz = dims
    lo, hi = (0.3, 1.0) if kind == 1 else (0.0, 1.0)
    w = mastg.Workload(kind, nx, ny, nz, seed=7, lo=lo, hi=hi, mast_exact=exact)
    P = oracle.Problem(kind, nx, ny, nz, seed=7, lo=lo, hi=hi, mast_exact=exact)
    xyz, conn, base, con, loads, params = arrays(mastg, w, npe)
    oxyz, oconn = P.mesh()
    obase, ocon = P.dofmap()
    assert np.array_equal(xyz, oxyz)
    assert np.array_equal(conn, oconn)
    assert np.array_equal(base, obase)
    assert np.array_equal(con, np.nonzero(ocon)[0])
    if kind == 1:
        np.testing.assert_allclose(loads, P.point_load(), rtol=1e-14, atol=1e-9)
    assert np.array_equal(params, P.params())


def test_abi_exports_every_declared_symbol(mastg):
    hdr = open(os.path.join(ROOT, "include", "mastg", "mastg.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = set(re.findall(r"\b(mastg_[a-z_0-9]+)\s*\(", hdr))
    assert len(names) >= 20
    L = C.CDLL(mastg.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
