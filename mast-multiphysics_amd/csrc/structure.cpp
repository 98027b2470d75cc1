// Host-side mesh structure (see structure.h).
//
// The DOF numbering restates libMesh's DofMap::distribute_local_dofs_var_major
// for one variable group (MAST declares its 6 structural variables with one FE
// type, src/elasticity/structural_system_initialization.cpp:37-55): node DOF
// bases in first-touch order of the element traversal, a node's variables
// contiguous.  The sparsity is libMesh's full-coupling pattern: row i holds the
// sorted union of the DOFs of every element containing DOF i.
#include "structure.h"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace mastg {

static int physics_npe(int physics) { return physics == 2 ? 4 : 8; }

std::string build_structure(int physics, int64_t n_nodes, int64_t n_elems, const double* xyz,
                            const int32_t* conn, int n_vars, const int64_t* node_dof_base,
                            int64_t n_constrained, const int64_t* constrained_dofs,
                            Structure& S) {
    if (physics < 1 || physics > 3) return "unknown physics";
    if (n_nodes <= 0 || n_elems <= 0 || !xyz || !conn) return "empty mesh";
    const int npe = physics_npe(physics);
    S.npe = npe;
    S.n_vars = n_vars;
    S.n_nodes = n_nodes;
    S.n_elems = n_elems;
    if (physics == 1) {
        if (n_vars != 3 && n_vars != 6) return "hex8 elasticity needs 3 or 6 variables per node";
        S.nav = 3;
        for (int v = 0; v < 3; v++) S.active_var[v] = v;
    } else if (physics == 2) {
        if (n_vars != 5 && n_vars != 6) return "Mindlin plate needs 5 or 6 variables per node";
        // an element whose nodes do not share one z is computed in its own
        // frame, which couples theta_z: then all six variables are active
        bool framed = false;
        for (int64_t e = 0; e < n_elems && !framed; e++) {
            const int32_t* en = conn + e * 4;
            bool in_range = true;
            for (int a = 0; a < 4; a++) in_range = in_range && en[a] >= 0 && en[a] < n_nodes;
            if (!in_range) return "connectivity out of range";
            const double z0 = xyz[int64_t(en[0]) * 3 + 2];
            for (int a = 1; a < 4; a++) framed = framed || xyz[int64_t(en[a]) * 3 + 2] != z0;
        }
        if (framed && n_vars != 6)
            return "plate elements outside a z = const plane need 6 variables per node (theta_z couples)";
        S.nav = framed ? 6 : 5;
        for (int v = 0; v < S.nav; v++) S.active_var[v] = v;
    } else {
        if (n_vars != 1) return "heat conduction needs 1 variable per node";
        S.nav = 1;
        S.active_var[0] = 0;
    }
    if (n_elems * int64_t(npe) * npe >= (int64_t(1) << 31)) return "mesh too large for 32-bit maps";
    for (int64_t i = 0; i < n_elems * npe; i++)
        if (conn[i] < 0 || conn[i] >= n_nodes) return "connectivity out of range";
    // a node twice in one element: a degenerate element (zero Jacobian), and
    // the assembly places each of an element's nodes once in a row
    for (int64_t e = 0; e < n_elems; e++)
        for (int a = 1; a < npe; a++)
            for (int b = 0; b < a; b++)
                if (conn[e * npe + a] == conn[e * npe + b]) return "element with a repeated node";

    // ---- caller DOF bases (libMesh first-touch when not given)
    std::vector<int64_t> base(n_nodes, -1);
    if (node_dof_base) {
        for (int64_t n = 0; n < n_nodes; n++) base[n] = node_dof_base[n];
    } else {
        int64_t next = 0;
        for (int64_t i = 0; i < n_elems * npe; i++) {
            const int32_t n = conn[i];
            if (base[n] < 0) { base[n] = next; next += n_vars; }
        }
        for (int64_t n = 0; n < n_nodes; n++)
            if (base[n] < 0) { base[n] = next; next += n_vars; }
    }
    S.n_dofs = 0;
    for (int64_t n = 0; n < n_nodes; n++) {
        if (base[n] < 0) return "negative node DOF base";
        S.n_dofs = std::max(S.n_dofs, base[n] + n_vars);
    }
    // ---- rank nodes by DOF base: active numbering is monotone in caller order
    S.node_of_rank.resize(n_nodes);
    std::iota(S.node_of_rank.begin(), S.node_of_rank.end(), 0);
    std::sort(S.node_of_rank.begin(), S.node_of_rank.end(),
              [&](int32_t a, int32_t b) { return base[a] < base[b]; });
    S.rank_of_node.resize(n_nodes);
    for (int64_t r = 0; r < n_nodes; r++) {
        if (r > 0 && base[S.node_of_rank[r]] < base[S.node_of_rank[r - 1]] + n_vars)
            return "node DOF ranges overlap";
        S.rank_of_node[S.node_of_rank[r]] = int32_t(r);
    }
    const int nav = S.nav;
    S.n_active = n_nodes * nav;
    S.caller_dof_of_active.resize(S.n_active);
    for (int64_t r = 0; r < n_nodes; r++)
        for (int v = 0; v < nav; v++)
            S.caller_dof_of_active[r * nav + v] = base[S.node_of_rank[r]] + S.active_var[v];
    S.xyz_r.resize(n_nodes * 3);
    for (int64_t r = 0; r < n_nodes; r++)
        for (int i = 0; i < 3; i++) S.xyz_r[r * 3 + i] = xyz[int64_t(S.node_of_rank[r]) * 3 + i];
    S.econn.resize(n_elems * npe);
    for (int64_t i = 0; i < n_elems * npe; i++) S.econn[i] = S.rank_of_node[conn[i]];

    // ---- node -> (element, local node) incidence, in element order
    S.ninc_ptr.assign(n_nodes + 1, 0);
    for (int64_t i = 0; i < n_elems * npe; i++) S.ninc_ptr[S.econn[i] + 1]++;
    for (int64_t r = 0; r < n_nodes; r++) S.ninc_ptr[r + 1] += S.ninc_ptr[r];
    S.ninc.resize(n_elems * npe);
    {
        // filled local-index-major so each node's incidences are sorted by la:
        // on structured meshes the lanes of a wave then walk la in lock step
        std::vector<int32_t> next(S.ninc_ptr.begin(), S.ninc_ptr.end() - 1);
        for (int a = 0; a < npe; a++)
            for (int64_t e = 0; e < n_elems; e++) S.ninc[next[S.econn[e * npe + a]]++] = int32_t(e * npe + a);
    }
    // ---- congruence classes
    {
        S.elem_class.assign(n_elems, 0);
        S.n_classes = 0;
        std::vector<std::vector<double>> rep;
        bool ok = true;
        std::vector<double> rel(npe * 3);
        for (int64_t e = 0; e < n_elems && ok; e++) {
            double size = 0;
            for (int a = 0; a < npe; a++)
                for (int i = 0; i < 3; i++) {
                    rel[a * 3 + i] = S.xyz_r[int64_t(S.econn[e * npe + a]) * 3 + i] -
                                     S.xyz_r[int64_t(S.econn[e * npe]) * 3 + i];
                    size = std::max(size, std::abs(rel[a * 3 + i]));
                }
            int found = -1;
            for (int c = 0; c < int(rep.size()) && found < 0; c++) {
                bool same = true;
                for (int k = 0; k < npe * 3 && same; k++)
                    same = std::abs(rel[k] - rep[c][k]) <= 1e-12 * size;
                if (same) found = c;
            }
            if (found < 0) {
                if (int(rep.size()) == Structure::kMaxClasses) { ok = false; break; }
                found = int(rep.size());
                rep.push_back(rel);
                S.class_rep.push_back(int32_t(e));
            }
            S.elem_class[e] = uint8_t(found);
        }
        S.n_classes = ok ? int(rep.size()) : 0;
        if (!ok) { S.class_rep.clear(); S.elem_class.clear(); }
    }
    // ---- node-block pattern: sorted union of element node ranks around each node
    S.blk_ptr.assign(n_nodes + 1, 0);
    std::vector<std::vector<int32_t>> nbr(n_nodes);
    for (int64_t r = 0; r < n_nodes; r++) {
        auto& v = nbr[r];
        for (int32_t k = S.ninc_ptr[r]; k < S.ninc_ptr[r + 1]; k++) {
            const int64_t e = S.ninc[k] / npe;
            for (int b = 0; b < npe; b++) v.push_back(S.econn[e * npe + b]);
        }
        if (v.empty()) v.push_back(int32_t(r));   // unattached node keeps its diagonal
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        S.blk_ptr[r + 1] = S.blk_ptr[r] + int64_t(v.size());
    }
    S.n_blocks = S.blk_ptr[n_nodes];
    for (int64_t r = 0; r < n_nodes; r++)
        S.max_blocks_per_row = std::max(S.max_blocks_per_row, S.blk_ptr[r + 1] - S.blk_ptr[r]);
    if (S.n_blocks >= (int64_t(1) << 31)) return "too many blocks for 32-bit maps";
    S.blk_col.resize(S.n_blocks);
    for (int64_t r = 0; r < n_nodes; r++) {
        std::copy(nbr[r].begin(), nbr[r].end(), S.blk_col.begin() + S.blk_ptr[r]);
        std::vector<int32_t>().swap(nbr[r]);
    }
    // ---- scalar CSR: rows (r, v), columns (nbr, w); rows of a node are contiguous
    S.nnz = S.n_blocks * nav * nav;
    S.row_ptr.resize(S.n_active + 1);
    S.col.resize(S.nnz);
    S.diag_pos.resize(S.n_active);
    for (int64_t r = 0; r < n_nodes; r++) {
        const int64_t deg = S.blk_ptr[r + 1] - S.blk_ptr[r];
        const int64_t b0 = S.blk_ptr[r] * nav * nav;
        for (int v = 0; v < nav; v++) {
            const int64_t rp = b0 + v * deg * nav;
            S.row_ptr[r * nav + v] = rp;
            for (int64_t k = 0; k < deg; k++) {
                const int32_t c = S.blk_col[S.blk_ptr[r] + k];
                for (int w = 0; w < nav; w++) S.col[rp + k * nav + w] = c * nav + w;
                if (c == r) S.diag_pos[r * nav + v] = rp + k * nav + v;
            }
        }
    }
    S.row_ptr[S.n_active] = S.nnz;
    // ---- element-to-block positions and per-block contribution segments
    const int64_t n_contrib = n_elems * npe * npe;
    std::vector<int32_t> pos(n_contrib);
    S.cptr.assign(S.n_blocks + 1, 0);
    for (int64_t e = 0; e < n_elems; e++)
        for (int a = 0; a < npe; a++) {
            const int32_t ra = S.econn[e * npe + a];
            const int32_t* b0 = S.blk_col.data() + S.blk_ptr[ra];
            const int32_t* b1 = S.blk_col.data() + S.blk_ptr[ra + 1];
            for (int b = 0; b < npe; b++) {
                const int32_t* p = std::lower_bound(b0, b1, S.econn[e * npe + b]);
                const int32_t k = int32_t(p - S.blk_col.data());
                pos[(e * npe + a) * npe + b] = k;
                S.cptr[k + 1]++;
            }
        }
    for (int64_t k = 0; k < S.n_blocks; k++) S.cptr[k + 1] += S.cptr[k];
    S.cidx.resize(n_contrib);
    {
        std::vector<int32_t> next(S.cptr.begin(), S.cptr.end() - 1);
        for (int64_t c = 0; c < n_contrib; c++) S.cidx[next[pos[c]]++] = int32_t(c);
    }
    S.blk_transpose.resize(S.n_blocks);
    for (int64_t r = 0; r < n_nodes; r++)
        for (int64_t k = S.blk_ptr[r]; k < S.blk_ptr[r + 1]; k++) {
            const int32_t c = S.blk_col[k];
            const int32_t* b0 = S.blk_col.data() + S.blk_ptr[c];
            const int32_t* b1 = S.blk_col.data() + S.blk_ptr[c + 1];
            S.blk_transpose[k] = int32_t(std::lower_bound(b0, b1, int32_t(r)) - S.blk_col.data());
        }
    // ---- Dirichlet mask over active DOFs (and over every caller variable)
    S.constrained.assign(S.n_active, 0);
    S.node_base_r.resize(n_nodes);
    for (int64_t r = 0; r < n_nodes; r++) S.node_base_r[r] = base[S.node_of_rank[r]];
    S.constrained_var.assign(size_t(n_nodes) * n_vars, 0);
    if (n_constrained > 0 && constrained_dofs) {
        // caller dof -> active index via the node ranks
        const std::vector<int64_t>& sorted_base = S.node_base_r;
        for (int64_t i = 0; i < n_constrained; i++) {
            const int64_t d = constrained_dofs[i];
            auto it = std::upper_bound(sorted_base.begin(), sorted_base.end(), d);
            if (it == sorted_base.begin()) return "constrained DOF out of range";
            const int64_t r = int64_t(it - sorted_base.begin()) - 1;
            const int64_t var = d - sorted_base[r];
            if (var < 0 || var >= n_vars) return "constrained DOF out of range";
            S.constrained_var[size_t(r) * n_vars + var] = 1;
            for (int v = 0; v < nav; v++)
                if (S.active_var[v] == var) S.constrained[r * nav + v] = 1;
        }
    }
    return "";
}

}  // namespace mastg
