"""Invalid problem descriptions are refused at context creation with
MASTG_ERR_INVALID and a message, before any device work (host code: runs
without a GPU).  The reference stops on the same conditions with libMesh
errors or asserts (DofMap and MeshBase consistency checks); here each is a
status and mastg_last_error's text, as the C ABI promises."""
import numpy as np
import pytest


def _cube(nx=2, ny=1, nz=1):
    xs = [(i, j, k) for k in range(nz + 1) for j in range(ny + 1) for i in range(nx + 1)]
    xyz = np.array(xs, float)
    nid = lambda i, j, k: i + (nx + 1) * (j + (ny + 1) * k)
    conn = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                conn.append([nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                             nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1)])
    return xyz, np.array(conn, np.int32)


def _square(n=2):
    xyz = np.array([(i, j, 0.0) for j in range(n + 1) for i in range(n + 1)], float)
    conn = np.array([[i + (n + 1) * j, i + 1 + (n + 1) * j, i + 1 + (n + 1) * (j + 1), i + (n + 1) * (j + 1)]
                     for j in range(n) for i in range(n)], np.int32)
    return xyz, conn


def _refused(M, desc, text):
    with pytest.raises(M.MastgError) as e:
        M.Context(desc)
    assert e.value.code == -1, e.value
    assert text in str(e.value), e.value


def test_invalid_descriptions_are_refused(mastg):
    M = mastg
    xyz, conn = _cube()
    _refused(M, M.desc_from_arrays(7, xyz, conn, 3), "unknown physics")
    _refused(M, M.desc_from_arrays(M.HEX8_ELASTIC, xyz, conn[:0], 3), "empty mesh")
    _refused(M, M.desc_from_arrays(M.HEX8_ELASTIC, xyz, conn, 4), "3 or 6 variables")
    _refused(M, M.desc_from_arrays(M.HEX8_HEAT, xyz, conn, 3), "1 variable")
    bad = conn.copy(); bad[1, 6] = len(xyz)
    _refused(M, M.desc_from_arrays(M.HEX8_ELASTIC, xyz, bad, 3), "connectivity out of range")
    bad = conn.copy(); bad[0, 0] = -1
    _refused(M, M.desc_from_arrays(M.HEX8_ELASTIC, xyz, bad, 3), "connectivity out of range")
    bad = conn.copy(); bad[1, 5] = bad[1, 2]
    _refused(M, M.desc_from_arrays(M.HEX8_ELASTIC, xyz, bad, 3), "repeated node")
    base = np.arange(len(xyz), dtype=np.int64) * 3
    nb = base.copy(); nb[4] = -3
    _refused(M, M.desc_from_arrays(M.HEX8_ELASTIC, xyz, conn, 3, node_dof_base=nb), "negative node DOF base")
    nb = base.copy(); nb[4] = nb[3] + 1
    _refused(M, M.desc_from_arrays(M.HEX8_ELASTIC, xyz, conn, 3, node_dof_base=nb), "overlap")
    n_dofs = 3 * len(xyz)
    for c in ([n_dofs], [-1]):
        _refused(M, M.desc_from_arrays(M.HEX8_ELASTIC, xyz, conn, 3, node_dof_base=base, constrained_dofs=c),
                 "constrained DOF out of range")
    sq, sc = _square()
    _refused(M, M.desc_from_arrays(M.QUAD4_MINDLIN, sq, sc, 3), "5 or 6 variables")
    tilted = sq.copy(); tilted[:, 2] = 0.1 * tilted[:, 0]
    _refused(M, M.desc_from_arrays(M.QUAD4_MINDLIN, tilted, sc, 5), "need 6 variables")
