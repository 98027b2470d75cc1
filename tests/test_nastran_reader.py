"""The library's Nastran bulk-data reader (host code, no GPU needed).

MAST reads Nastran models through pyNastran (src/mesh/nastran_io.cpp:217-255,
src/mesh/pynastran_io.pyx), which is not part of the reference and is not
available here, so the reader is pinned by hand-written cards whose meaning
the Nastran format fixes (small, large and free field; continuations; the
implicit-exponent reals; SPC1 THRU) and by the unstructured fixtures of
tests/golden/ checked against the arrays they were written from
(tests/golden/make_nastran_fixtures.py).  Order conventions follow the
reference: nodes by Nastran id, elements by EID (their libMesh ids), DOF
bases first-touch over the elements with 6 variables per node.
"""
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")

TINY = """\
$ two hex8 elements sharing a face, three field formats
SOL 101
CEND
BEGIN BULK
GRID          10              0.      0.      0.
GRID*                 12                             1.0             0.0*
*                    0.0
GRID,14,,1.,1.,0.
GRID          16              0.     1.0      0.
GRID          20              0.      0.     1.0
GRID,22,0,1.,0.,1.
GRID          24              1.      1.      1.
GRID,26,,0.,1.,1.
GRID          30              0.      0.     2.0
GRID          32              1.      0.     2.0
GRID          34              1.      1.     2.0
GRID          36             0.0      1. 2.0+0
CHEXA        200       5      20      22      24      26      30      32+
+             34      36
CHEXA,100,5,10,12,14,16,20,22
,24,26
SPC1           1     123      10    THRU      16
SPC,2,30,456,0.,34,3,0.
FORCE          7      36       0   2.5+3      0.      0.     -1.
FORCE,7,34,,500.,1.,0.,0.
MAT1           1  7.2+10  2.7+10
PSOLID         1       1
ENDDATA
GRID          99              5.      5.      5.
"""


def test_tiny_model_all_formats(mastg, tmp_path):
    p = tmp_path / "tiny.bdf"
    p.write_text(TINY)
    m = mastg.NastranModel(str(p))
    d = m.desc.contents
    assert d.physics == mastg.HEX8_ELASTIC and d.n_vars == 6
    assert list(m.node_ids) == [10, 12, 14, 16, 20, 22, 24, 26, 30, 32, 34, 36]   # ENDDATA ends the bulk
    assert list(m.elem_ids) == [100, 200]
    np.testing.assert_array_equal(m.xyz[:4], [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]])
    np.testing.assert_array_equal(m.xyz[8:], [[0, 0, 2], [1, 0, 2], [1, 1, 2], [0, 1, 2]])
    idx = {int(n): k for k, n in enumerate(m.node_ids)}
    np.testing.assert_array_equal(m.conn[0], [idx[n] for n in (10, 12, 14, 16, 20, 22, 24, 26)])
    np.testing.assert_array_equal(m.conn[1], [idx[n] for n in (20, 22, 24, 26, 30, 32, 34, 36)])
    # first touch over EID order: element 100 numbers grids 10..26, then 30..36
    np.testing.assert_array_equal(m.node_base, 6 * np.arange(12))
    # SPC1 123 on grids 10 THRU 16: the grids of a THRU range need not exist
    # (11, 13, 15 do not); SPC 456 on 30 and 3 on 34
    con = set(m.constrained.tolist())
    assert m.spc_sets == 2
    expect = {6 * idx[n] + c for n in (10, 12, 14, 16) for c in range(3)} | \
             {6 * idx[30] + c for c in (3, 4, 5)} | {6 * idx[34] + 2}
    assert con == expect
    F = m.point_loads
    assert F[6 * idx[36] + 2] == -2500.0 and F[6 * idx[34] + 0] == 500.0
    assert np.count_nonzero(F) == 2
    assert d.E == 7.2e10 and abs(d.nu - (7.2e10 / (2 * 2.7e10) - 1)) <= 1e-15
    np.testing.assert_array_equal(m.params, [1.0, 1.0])
    m.close()


def test_errors_are_reported(mastg, tmp_path):
    cases = {
        "GRID,1,3,0.,0.,0.\nCHEXA,1,1,1,1,1,1,1,1\n,1,1\n": "basic coordinate",
        "GRID,1,,0.,0.,0.\nCHEXA,1,1,1,2,3,4,5,6\n,7,8\n": "missing grid",
        "GRID,1,,0.,0.,0.\nSPC,1,1,3,0.5\n": "homogeneous",
        "GRID,1,,0.,0.,0.\n": "CHEXA or CQUAD4",
    }
    for text, msg in cases.items():
        p = tmp_path / "bad.bdf"
        p.write_text(text)
        with pytest.raises(mastg.MastgError) as ei:
            mastg.NastranModel(str(p))
        assert msg in str(ei.value), (text, str(ei.value))


@pytest.mark.parametrize("name", ["butterfly_hex", "butterfly_quad"])
def test_unstructured_fixture(mastg, name):
    g = np.load(os.path.join(GOLD, name + ".npz"))
    m = mastg.NastranModel(os.path.join(GOLD, name + ".bdf"))
    order = np.argsort(g["node_ids"])
    assert np.array_equal(m.node_ids, g["node_ids"][order])
    # small-field grids carry 8-character reals (about 5 digits)
    assert np.max(np.abs(m.xyz - g["xyz"][order])) <= 5e-5
    eorder = np.argsort(g["elem_ids"])
    assert np.array_equal(m.elem_ids, g["elem_ids"][eorder])
    assert np.array_equal(m.node_ids[m.conn], g["conn_ids"][eorder])
    # first-touch bases over EID order
    seen = {}
    for e in range(m.n_elems):
        for v in m.conn[e]:
            if v not in seen:
                seen[v] = 6 * len(seen)
    assert np.array_equal(m.node_base, np.array([seen[k] for k in range(m.n_nodes)]))
    idx = {int(n): k for k, n in enumerate(m.node_ids)}
    spc = {int(m.node_base[idx[int(n)]]) + c for n in g["spc_ids"]
           for c in (range(3) if name == "butterfly_quad" else range(6))}
    assert set(m.constrained.tolist()) == spc
    assert m.spc_sets == 1
    if name == "butterfly_quad":
        assert m.desc.contents.physics == mastg.QUAD4_MINDLIN
        assert np.array_equal(m.params, g["thickness"][eorder])
        Fz = m.point_loads[[m.node_base[idx[int(n)]] + 2 for n in g["force_ids"]]]
        np.testing.assert_array_equal(Fz, g["force_z"])
        assert np.count_nonzero(m.point_loads) == len(g["force_ids"])
    else:
        F = m.point_loads.reshape(-1, 6)
        rows = [m.node_base[idx[int(n)]] // 6 for n in g["force_ids"]]
        np.testing.assert_allclose(F[rows, :3], np.tile([6000.0, 8000.0, 0.0], (len(rows), 1)), rtol=1e-15)
        assert np.count_nonzero(F) == 2 * len(rows)
    d = m.desc.contents
    assert d.E == 7.2e10 and d.nu == 0.33
    m.close()
