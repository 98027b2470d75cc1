"""Shared helpers: build product contexts from oracle problems, error norms."""
import numpy as np


def ctx_from_oracle(M, P, physics=None, **extra):
    """A mastg Context over exactly the oracle problem's mesh, DofMap, BCs and
    loads (`extra`: further description fields, e.g. radiation, that the
    oracle problem was given too)."""
    xyz, conn = P.mesh()
    base, con = P.dofmap()
    sides = P.bsides()
    phys = physics or {1: M.HEX8_ELASTIC, 2: M.QUAD4_MINDLIN, 3: M.HEX8_HEAT}[P.kind]
    mat = {}
    if phys == M.HEX8_ELASTIC:
        mat = dict(E=72e9, nu=0.33, penalty=3.0)
    elif phys == M.QUAD4_MINDLIN:
        mat = dict(E=70e9, nu=0.3, kappa=5.0 / 6.0, offset=0.0, pressure=1.0e3)
    else:
        mat = dict(k0=200.0, alpha=1e-3, conv_h=50.0, conv_tinf=300.0, conv_bid_mask=0x3E)
    mat.update(extra)
    pl = P.point_load() if phys == M.HEX8_ELASTIC else None
    desc = M.desc_from_arrays(phys, xyz, conn, P.n_vars, node_dof_base=base,
                              constrained_dofs=np.nonzero(con)[0], sides=sides,
                              point_loads=pl, **mat)
    ctx = M.Context(desc)
    ctx.set_params(P.params())
    return ctx


def rel_err_floor(a, b, rp=None, floor=1e-4):
    """max |a-b| / max(|b|, floor * max|row of b|): the entry-wise relative error
    with a floor at `floor` times the largest entry of the row (for CSR values;
    without rp the floor is relative to the largest entry overall).  Entries that
    cancel to ~0 in exact arithmetic carry rounding of order 1e-16 of the row
    scale in any FP64 summation order, hence the floor."""
    a = np.asarray(a); b = np.asarray(b)
    if rp is None:
        den = np.maximum(np.abs(b), floor * np.abs(b).max())
    else:
        rowmax = np.maximum.reduceat(np.abs(b), rp[:-1]) if len(b) else np.zeros(0)
        rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
        den = np.maximum(np.abs(b), floor * rowmax[rows])
        den = np.where(den == 0, 1.0, den)
    return float(np.max(np.abs(a - b) / den)) if len(b) else 0.0


def active_restriction(P, base, n_active_vars):
    """Indices (in the oracle's caller numbering) of the active variables."""
    nv = P.n_vars
    rows = []
    order = np.argsort(base, kind="stable")
    for n in order:
        for v in range(n_active_vars):
            rows.append(base[n] + v)
    return np.asarray(rows, np.int64)


def restrict_csr(rp, col, val, keep):
    """Restrict a CSR matrix to rows/cols in `keep` (sorted caller indices),
    renumbering columns by position in `keep`."""
    pos = -np.ones(rp.size - 1, np.int64)
    pos[keep] = np.arange(keep.size)
    out_rp = [0]
    out_col = []
    out_val = []
    for r in keep:
        c = col[rp[r]:rp[r + 1]]
        m = pos[c] >= 0
        out_col.append(pos[c][m])
        if val is not None:
            out_val.append(val[rp[r]:rp[r + 1]][m])
        out_rp.append(out_rp[-1] + int(m.sum()))
    out_col = np.concatenate(out_col).astype(np.int32)
    out_val = np.concatenate(out_val) if val is not None else None
    return np.asarray(out_rp, np.int64), out_col, out_val
