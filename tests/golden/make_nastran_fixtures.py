"""Writes the Nastran bulk-data fixtures of the reader's tests (run from the
repository root: python tests/golden/make_nastran_fixtures.py).

butterfly_hex.bdf / butterfly_quad.bdf: an unstructured mesh of a disk (a
square core and four blocks to the rim: the nodes at the core's corners are
shared by three cells, not four), as CQUAD4 plates or extruded to CHEXA;
Nastran node and element ids shuffled and sparse; the cards written in a
shuffled order and in the three field formats (small, large GRID*, free).
The plate is simply supported on its rim (SPC1 123) and loaded by FORCE
cards normal to it on the interior grids, with two PSHELL thicknesses; the
solid is clamped on its bottom face (SPC1 123456) and pushed sideways on its
top face.  Next to each .bdf a .npz holds the mesh the file was written from
(coordinates and connectivity in Nastran id order, the loads, constrained
grid components and thicknesses) for the reader's checks.
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def butterfly(n=4, m=3, a=0.4, R=1.0):
    """2D quads of a disk of radius R: an n x n core square of half-width a
    and four n x m blocks from its sides to the rim."""
    pts, key = [], {}

    def node(x, y):
        k = (round(x, 12), round(y, 12))
        if k not in key:
            key[k] = len(pts)
            pts.append((x, y))
        return key[k]

    quads = []
    g = np.linspace(-a, a, n + 1)
    for j in range(n):
        for i in range(n):
            quads.append([node(g[i], g[j]), node(g[i + 1], g[j]), node(g[i + 1], g[j + 1]), node(g[i], g[j + 1])])
    for side in range(4):
        ang0 = np.pi / 4 + side * np.pi / 2          # corners of the core at 45, 135, ...
        rot = np.array([[np.cos(side * np.pi / 2), -np.sin(side * np.pi / 2)],
                        [np.sin(side * np.pi / 2), np.cos(side * np.pi / 2)]])

        def P(s, t):
            # s in [0, 1] along the core side (x = a face rotated), t in [0, 1] to the rim
            inner = rot @ np.array([a, -a + 2 * a * s])
            th = -np.pi / 4 + s * np.pi / 2 + side * np.pi / 2
            outer = R * np.array([np.cos(th), np.sin(th)])
            return (1 - t) * inner + t * outer
        for j in range(m):
            for i in range(n):
                c = [P(i / n, j / m), P((i + 1) / n, j / m), P((i + 1) / n, (j + 1) / m), P(i / n, (j + 1) / m)]
                q = [node(*p) for p in c]
                quads.append(q)
        del ang0
    xy = np.array(pts)
    quads = np.array(quads)
    # counter-clockwise
    for e, q in enumerate(quads):
        x = xy[q]
        area = 0.5 * np.sum(x[:, 0] * np.roll(x[:, 1], -1) - np.roll(x[:, 0], -1) * x[:, 1])
        if area < 0:
            quads[e] = q[::-1]
    return xy, quads


def fmt_small(name, fields):
    out, line = [], name.ljust(8)
    for i, f in enumerate(fields):
        if i and i % 8 == 0:
            out.append(line + "+C" + str(len(out)).ljust(6))
            line = ("+C" + str(len(out) - 1)).ljust(8)
        line += str(f)[:8].rjust(8) if not isinstance(f, str) else f[:8].ljust(8)
    out.append(line)
    return out


def real8(v):
    # Nastran 8-character real: implicit exponent for small magnitudes
    if v == 0.0:
        return "0."
    s = "%.6e" % v
    mant, ex = s.split("e")
    ex = int(ex)
    if -4 <= ex <= 4:
        r = ("%.7g" % v)
        if "." not in r and "e" not in r:
            r += "."
        return r[:8]
    tail = ("-" if ex < 0 else "+") + str(abs(ex))
    for dec in (4, 3, 2, 1, 0):
        m = ("%.*f" % (dec, float(mant))).rstrip("0")
        if len(m) + len(tail) <= 8:
            return m + tail
    raise ValueError(v)


def write(path, grids, elems, etype, spc, forces, pshell, rng):
    lines = ["$ written by tests/golden/make_nastran_fixtures.py", "SOL 101", "CEND", "BEGIN BULK"]
    cards = []
    for k, (nid, x) in enumerate(grids):
        style = k % 3
        if style == 0:
            cards.append(["GRID".ljust(8) + str(nid).rjust(8) + "".rjust(8) + "".join(real8(c).rjust(8) for c in x)])
        elif style == 1:
            l1 = "GRID*".ljust(8) + str(nid).rjust(16) + "".rjust(16) + ("%.9E" % x[0]).rjust(16) + \
                ("%.9E" % x[1]).rjust(16) + "*G%d" % nid
            l2 = ("*G%d" % nid).ljust(8) + ("%.9E" % x[2]).rjust(16)
            cards.append([l1, l2])
        else:
            cards.append(["GRID,%d,,%.15g,%.15g,%.15g" % (nid, x[0], x[1], x[2])])
    for k, (eid, pid, g) in enumerate(elems):
        if k % 2 == 0:
            f = [eid, pid] + list(g)
            cards.append(fmt_small(etype, f))
        else:
            if len(g) == 8:
                cards.append(["%s,%d,%d,%s" % (etype, eid, pid, ",".join(map(str, g[:6]))),
                              "+,%s" % ",".join(map(str, g[6:]))])
            else:
                cards.append(["%s,%d,%d,%s" % (etype, eid, pid, ",".join(map(str, g)))])
    for sid, comp, nodes in spc:
        nodes = sorted(nodes)
        # a THRU range where the ids run consecutively, else a list
        cards.append(fmt_small("SPC1", [sid, comp] + nodes))
    for sid, nid, F, n in forces:
        cards.append(["FORCE,%d,%d,0,%.15g,%.15g,%.15g,%.15g" % (sid, nid, F, n[0], n[1], n[2])])
    cards.append(["MAT1,1,7.2+10,,0.33"])
    for pid, t in pshell:
        cards.append(fmt_small("PSHELL", [pid, 1, real8(t)]))
    cards.append(["PSOLID,1,1"])
    order = rng.permutation(len(cards))
    for i in order:
        lines += cards[i]
    lines.append("ENDDATA")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def main():
    rng = np.random.default_rng(2024)
    xy, quads = butterfly()
    nn2 = len(xy)
    # --- plate
    nid = np.sort(rng.choice(np.arange(1, 20 * nn2), nn2, replace=False))
    nid = rng.permutation(nid)                      # Nastran id of 2D node k
    eid = rng.permutation(np.sort(rng.choice(np.arange(100, 100 + 30 * len(quads)), len(quads), replace=False)))
    pid = np.where(np.arange(len(quads)) < 16, 11, 12)
    thick = {11: 0.008, 12: 0.005}
    xyz = np.c_[xy, np.zeros(nn2)]
    rim = np.isclose(np.hypot(xy[:, 0], xy[:, 1]), 1.0)
    F = [(7, int(nid[k]), 250.0, (0.0, 0.0, -1.0)) for k in range(nn2) if not rim[k]]
    write(os.path.join(HERE, "butterfly_quad.bdf"), [(int(nid[k]), xyz[k]) for k in range(nn2)],
          [(int(eid[e]), int(pid[e]), [int(nid[v]) for v in quads[e]]) for e in range(len(quads))],
          "CQUAD4", [(3, "123", [int(nid[k]) for k in range(nn2) if rim[k]])], F, list(thick.items()), rng)
    np.savez(os.path.join(HERE, "butterfly_quad.npz"), node_ids=nid, xyz=xyz, elem_ids=eid,
             conn_ids=nid[quads], pid=pid, thickness=np.array([thick[p] for p in pid]),
             force_ids=np.array([f[1] for f in F]), force_z=np.array([-250.0] * len(F)),
             spc_ids=nid[rim])
    # --- solid: 3 layers of height 0.2
    nz, H = 3, 0.2
    nn3 = nn2 * (nz + 1)
    xyz3 = np.concatenate([np.c_[xy, np.full(nn2, H * l)] for l in range(nz + 1)])
    hexes = np.array([np.r_[q + l * nn2, q + (l + 1) * nn2] for l in range(nz) for q in quads])
    nid3 = rng.permutation(np.sort(rng.choice(np.arange(1, 10 * nn3), nn3, replace=False)))
    eid3 = rng.permutation(np.arange(5000, 5000 + len(hexes)) * 3)
    bottom = np.arange(nn2)
    top = np.arange(nz * nn2, nn3)
    F3 = [(9, int(nid3[k]), 1.0e4, (0.6, 0.8, 0.0)) for k in top]
    write(os.path.join(HERE, "butterfly_hex.bdf"), [(int(nid3[k]), xyz3[k]) for k in range(nn3)],
          [(int(eid3[e]), 1, [int(nid3[v]) for v in hexes[e]]) for e in range(len(hexes))],
          "CHEXA", [(1, "123456", [int(nid3[k]) for k in bottom])], F3, [], rng)
    np.savez(os.path.join(HERE, "butterfly_hex.npz"), node_ids=nid3, xyz=xyz3, elem_ids=eid3,
             conn_ids=nid3[hexes], force_ids=np.array([f[1] for f in F3]), spc_ids=nid3[bottom])


if __name__ == "__main__":
    main()
