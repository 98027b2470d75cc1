"""GPU parity under the DOF numbering of a mesh partitioned over P processors.

With P MPI ranks libMesh numbers the DOFs processor by processor: a node
belongs to the lowest-numbered partition among its elements
(Partitioner::set_node_processor_ids), and each processor numbers the nodes it
owns first-touch over its own elements, its DOF range following the previous
processor's (DofMap::distribute_dofs).  MAST hands these dof_indices to the
assembly unchanged (src/base/nonlinear_implicit_assembly.cpp:111-183).  Here
the elements are split into P slabs along x; the numbering is built as
described and handed to both the oracle and the library.

Bars: row_ptr / col_idx bit-exact against the oracle's pattern in that
numbering; K within 1e-12 of max(|K_ij|, 1e-2 x row max); compliance 1e-9 and
dC/drho 1e-7 after solves to 1e-10 (CG + Jacobi and multigrid).
"""
import numpy as np
import pytest

from helpers import rel_err_floor

pytestmark = pytest.mark.gpu


def partitioned_numbering(conn, n_nodes, elem_part, n_parts, n_vars):
    """Node DOF bases of the processor-by-processor numbering."""
    npe = conn.shape[1]
    owner = np.full(n_nodes, n_parts, np.int64)
    for e in range(conn.shape[0]):
        owner[conn[e]] = np.minimum(owner[conn[e]], elem_part[e])
    base = np.full(n_nodes, -1, np.int64)
    nxt = 0
    for p in range(n_parts):
        for e in np.nonzero(elem_part == p)[0]:
            for a in range(npe):
                n = conn[e, a]
                if owner[n] == p and base[n] < 0:
                    base[n] = nxt
                    nxt += n_vars
    assert (base >= 0).all()
    return base


@pytest.mark.parametrize("parts", [2, 5])
def test_hex8_partitioned_numbering(oracle, mastg, parts):
    print(f"[partitioned numbering] hex8 8x4x4 over {parts} partitions", flush=True)
    P0 = oracle.Problem(oracle.CANTILEVER, 8, 4, 4, seed=6, lo=0.2, hi=1.0, mast_exact=True)
    xyz, conn = P0.mesh()
    base0, con0 = P0.dofmap()
    nn, ne = len(xyz), len(conn)
    ix = np.arange(ne) % 8                      # element x index (x fastest)
    part = (ix * parts) // 8
    base = partitioned_numbering(conn, nn, part, parts, P0.n_vars)
    # move the constraints and loads to the new numbering
    con = np.zeros(P0.n_dofs, np.uint8)
    F0 = P0.point_load()
    F = np.zeros(P0.n_dofs)
    for n in range(nn):
        for v in range(P0.n_vars):
            con[base[n] + v] = con0[base0[n] + v]
            F[base[n] + v] = F0[base0[n] + v]
    P = oracle.Problem(oracle.CANTILEVER, arrays=dict(xyz=xyz, conn=conn, n_vars=P0.n_vars, node_base=base,
                                                      constrained=con, params=P0.params(), point_load=F))
    desc = mastg.desc_from_arrays(mastg.HEX8_ELASTIC, xyz, conn, P0.n_vars, node_dof_base=base,
                                  constrained_dofs=np.nonzero(con)[0], point_loads=F, E=72e9, nu=0.33, penalty=3.0)
    ctx = mastg.Context(desc)
    ctx.set_params(P0.params())
    ctx.assemble()
    rp, col, val = ctx.export_csr_full()
    orp, ocol = P.csr()
    _, K = P.assemble()
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    eK = rel_err_floor(val, K, orp, floor=1e-2)
    ref = P.solve_elastic(1e-12)
    ref0 = P0.solve_elastic(1e-12)
    # the numbering does not change the answer
    assert abs(ref["compliance"] - ref0["compliance"]) <= 1e-10 * abs(ref0["compliance"])
    errs = [eK]
    for sv in (mastg.SOLVER_CG_JACOBI, mastg.SOLVER_MG):
        r = ctx.design_eval(P0.params(), solver=sv)
        ec = abs(r["objective"] - ref["compliance"]) / abs(ref["compliance"])
        eg = rel_err_floor(r["grad"], ref["dc"], floor=1e-6)
        errs += [ec, eg]
        assert ec <= 1e-9 and eg <= 1e-7
    print(f"  K {errs[0]:.1e}; CG C {errs[1]:.1e} grad {errs[2]:.1e}; MG C {errs[3]:.1e} grad {errs[4]:.1e}",
          flush=True)
    assert eK <= 1e-12
    ctx.close()
