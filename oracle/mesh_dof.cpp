// Oracle (test infrastructure only): meshes, DofMap, sparsity, Dirichlet sets.
// Restates the libMesh behaviour MAST builds on; see oracle.h for citations.
#include "oracle.h"
#include <algorithm>
#include <stdexcept>

namespace oracle {

// libMesh Hex8::side_nodes_map / Quad4::side_nodes_map
const int HEX8_SIDE_NODES[6][4] = {
    {0, 3, 2, 1}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}, {4, 5, 6, 7}};
const int QUAD4_SIDE_NODES[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};

Mesh build_cube(int nx, int ny, int nz, double xmin, double xmax, double ymin,
                double ymax, double zmin, double zmax) {
    Mesh m;
    m.dim = 3; m.npe = 8; m.nx = nx; m.ny = ny; m.nz = nz;
    m.n_nodes = int64_t(nx + 1) * (ny + 1) * (nz + 1);
    m.n_elems = int64_t(nx) * ny * nz;
    m.xyz.resize(m.n_nodes * 3);
    int64_t id = 0;
    // unit-cube points i/nx, j/ny, k/nz, then scaled by (max-min) and shifted
    for (int k = 0; k <= nz; k++)
        for (int j = 0; j <= ny; j++)
            for (int i = 0; i <= nx; i++, id++) {
                double px = double(i) / double(nx), py = double(j) / double(ny),
                       pz = double(k) / double(nz);
                m.xyz[id * 3 + 0] = px * (xmax - xmin) + xmin;
                m.xyz[id * 3 + 1] = py * (ymax - ymin) + ymin;
                m.xyz[id * 3 + 2] = pz * (zmax - zmin) + zmin;
            }
    auto idx = [&](int i, int j, int k) {
        return int32_t(i + (nx + 1) * (j + (ny + 1) * k));
    };
    m.conn.resize(m.n_elems * 8);
    int64_t e = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++, e++) {
                int32_t* c = &m.conn[e * 8];
                c[0] = idx(i, j, k);         c[1] = idx(i + 1, j, k);
                c[2] = idx(i + 1, j + 1, k); c[3] = idx(i, j + 1, k);
                c[4] = idx(i, j, k + 1);     c[5] = idx(i + 1, j, k + 1);
                c[6] = idx(i + 1, j + 1, k + 1); c[7] = idx(i, j + 1, k + 1);
                auto add = [&](int s, int b) {
                    m.bs_elem.push_back(int32_t(e)); m.bs_side.push_back(s); m.bs_id.push_back(b);
                };
                if (k == 0) add(0, 0);
                if (k == nz - 1) add(5, 5);
                if (j == 0) add(1, 1);
                if (j == ny - 1) add(3, 3);
                if (i == 0) add(4, 4);
                if (i == nx - 1) add(2, 2);
            }
    return m;
}

Mesh build_square(int nx, int ny, double xmin, double xmax, double ymin, double ymax) {
    Mesh m;
    m.dim = 2; m.npe = 4; m.nx = nx; m.ny = ny; m.nz = 0;
    m.n_nodes = int64_t(nx + 1) * (ny + 1);
    m.n_elems = int64_t(nx) * ny;
    m.xyz.resize(m.n_nodes * 3);
    int64_t id = 0;
    for (int j = 0; j <= ny; j++)
        for (int i = 0; i <= nx; i++, id++) {
            double px = double(i) / double(nx), py = double(j) / double(ny);
            m.xyz[id * 3 + 0] = px * (xmax - xmin) + xmin;
            m.xyz[id * 3 + 1] = py * (ymax - ymin) + ymin;
            m.xyz[id * 3 + 2] = 0.0;
        }
    auto idx = [&](int i, int j) { return int32_t(i + (nx + 1) * j); };
    m.conn.resize(m.n_elems * 4);
    int64_t e = 0;
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++, e++) {
            int32_t* c = &m.conn[e * 4];
            c[0] = idx(i, j); c[1] = idx(i + 1, j); c[2] = idx(i + 1, j + 1); c[3] = idx(i, j + 1);
            auto add = [&](int s, int b) {
                m.bs_elem.push_back(int32_t(e)); m.bs_side.push_back(s); m.bs_id.push_back(b);
            };
            if (j == 0) add(0, 0);
            if (j == ny - 1) add(2, 2);
            if (i == 0) add(3, 3);
            if (i == nx - 1) add(1, 1);
        }
    return m;
}

void DofMap::elem_dofs(const Mesh& m, int64_t e, int64_t* out) const {
    const int32_t* c = &m.conn[e * m.npe];
    for (int v = 0; v < n_vars; v++)
        for (int a = 0; a < m.npe; a++) out[v * m.npe + a] = node_base[c[a]] + v;
}

DofMap build_dofmap(const Mesh& m, int n_vars) {
    DofMap d;
    d.n_vars = n_vars;
    d.node_base.assign(m.n_nodes, -1);
    int64_t next = 0;
    for (int64_t e = 0; e < m.n_elems; e++)
        for (int a = 0; a < m.npe; a++) {
            int32_t n = m.conn[e * m.npe + a];
            if (d.node_base[n] < 0) { d.node_base[n] = next; next += n_vars; }
        }
    // nodes not attached to any element come last, in id order
    for (int64_t n = 0; n < m.n_nodes; n++)
        if (d.node_base[n] < 0) { d.node_base[n] = next; next += n_vars; }
    d.n_dofs = next;
    d.constrained.assign(d.n_dofs, 0);
    return d;
}

void add_dirichlet(const Mesh& m, DofMap& d, uint32_t bid_mask, uint32_t var_mask) {
    for (size_t s = 0; s < m.bs_elem.size(); s++) {
        if (!((bid_mask >> m.bs_id[s]) & 1u)) continue;
        int64_t e = m.bs_elem[s];
        int nsn = m.dim == 3 ? 4 : 2;
        for (int k = 0; k < nsn; k++) {
            int ln = m.dim == 3 ? HEX8_SIDE_NODES[m.bs_side[s]][k] : QUAD4_SIDE_NODES[m.bs_side[s]][k];
            int32_t n = m.conn[e * m.npe + ln];
            for (int v = 0; v < d.n_vars; v++)
                if ((var_mask >> v) & 1u) d.constrained[d.node_base[n] + v] = 1;
        }
    }
}

int64_t CSR::find(int64_t r, int64_t c) const {
    const int32_t* b = col.data() + row_ptr[r];
    const int32_t* en = col.data() + row_ptr[r + 1];
    const int32_t* p = std::lower_bound(b, en, int32_t(c));
    if (p == en || *p != c) return -1;
    return p - col.data();
}

std::vector<std::vector<int32_t>> node_to_elems(const Mesh& m) {
    std::vector<std::vector<int32_t>> ne(m.n_nodes);
    for (int64_t e = 0; e < m.n_elems; e++)
        for (int a = 0; a < m.npe; a++) ne[m.conn[e * m.npe + a]].push_back(int32_t(e));
    return ne;
}

CSR build_sparsity(const Mesh& m, const DofMap& d) {
    // per-row union of element DOF sets (libMesh SparsityPattern::Build)
    std::vector<std::vector<int32_t>> rows(d.n_dofs);
    const int nd = m.npe * d.n_vars;
    std::vector<int64_t> dofs(nd);
    for (int64_t e = 0; e < m.n_elems; e++) {
        d.elem_dofs(m, e, dofs.data());
        for (int i = 0; i < nd; i++)
            for (int j = 0; j < nd; j++) rows[dofs[i]].push_back(int32_t(dofs[j]));
    }
    CSR A;
    A.n = d.n_dofs;
    A.row_ptr.assign(d.n_dofs + 1, 0);
    for (int64_t r = 0; r < d.n_dofs; r++) {
        auto& v = rows[r];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        if (v.empty()) v.push_back(int32_t(r));  // unattached DOF keeps its diagonal
        A.row_ptr[r + 1] = A.row_ptr[r] + int64_t(v.size());
    }
    A.col.resize(A.row_ptr.back());
    for (int64_t r = 0; r < d.n_dofs; r++) {
        std::copy(rows[r].begin(), rows[r].end(), A.col.begin() + A.row_ptr[r]);
        std::vector<int32_t>().swap(rows[r]);
    }
    A.val.assign(A.col.size(), 0.0);
    return A;
}

std::vector<int32_t> sparsity_row(const Mesh& m, const DofMap& d,
                                  const std::vector<std::vector<int32_t>>& node_elems,
                                  int64_t dof) {
    // the node owning `dof`
    int64_t node = -1;
    for (int64_t n = 0; n < m.n_nodes; n++)
        if (dof >= d.node_base[n] && dof < d.node_base[n] + d.n_vars) { node = n; break; }
    if (node < 0) throw std::runtime_error("dof not found");
    std::vector<int32_t> row;
    std::vector<int64_t> dofs(m.npe * d.n_vars);
    for (int32_t e : node_elems[node]) {
        d.elem_dofs(m, e, dofs.data());
        for (int64_t x : dofs) row.push_back(int32_t(x));
    }
    std::sort(row.begin(), row.end());
    row.erase(std::unique(row.begin(), row.end()), row.end());
    return row;
}

void uniform_fill(uint64_t seed, int64_t n, double lo, double hi, double* out) {
    uint64_t s = seed;
    for (int64_t i = 0; i < n; i++) {
        s += 0x9E3779B97F4A7C15ULL;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z = z ^ (z >> 31);
        double u = double(z >> 11) * (1.0 / 9007199254740992.0);
        out[i] = lo + (hi - lo) * u;
    }
}

}  // namespace oracle
