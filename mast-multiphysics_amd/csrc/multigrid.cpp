// Multigrid hierarchy construction (see multigrid.h).
#include "multigrid.h"

#include <algorithm>
#include <climits>
#include <cmath>

namespace mastg {

namespace {

// libMesh hex8 / quad4 local node -> cell corner offset (quad4: the first four)
constexpr int kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                               {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

bool detect_lattice(const Structure& S, int dim, MGLevelHost& L, std::string& why) {
    const int npe = dim == 3 ? 8 : 4;
    const double* X = S.xyz_r.data();
    const int32_t* en = S.econn.data();
    // element 0 fixes the cell size (corner 0 to the opposite corner); it must
    // be an axis-aligned box in libMesh order
    const int opp = dim == 3 ? 6 : 2;
    double h[3];
    for (int d = 0; d < 3; d++) h[d] = X[int64_t(en[opp]) * 3 + d] - X[int64_t(en[0]) * 3 + d];
    if (!(h[0] > 0 && h[1] > 0 && (dim == 2 || h[2] > 0))) { why = "elements are not lattice cells"; return false; }
    double lo[3] = {1e300, 1e300, 1e300};
    for (int64_t r = 0; r < S.n_nodes; r++)
        for (int d = 0; d < 3; d++) lo[d] = std::min(lo[d], X[r * 3 + d]);
    std::vector<int32_t> ijk(size_t(S.n_nodes) * 3, 0);
    int n[3] = {0, 0, 0};
    for (int64_t r = 0; r < S.n_nodes; r++)
        for (int d = 0; d < dim; d++) {
            const double t = (X[r * 3 + d] - lo[d]) / h[d];
            const double ti = std::nearbyint(t);
            if (std::fabs(t - ti) > 1e-9 || ti > 2e9) { why = "nodes are not on a lattice"; return false; }
            ijk[r * 3 + d] = int32_t(ti);
            n[d] = std::max(n[d], int32_t(ti));
        }
    L.dim = dim;
    L.npe = npe;
    L.nav = S.nav;
    L.nx = n[0]; L.ny = n[1]; L.nz = n[2];
    const int64_t nlat = int64_t(n[0] + 1) * (n[1] + 1) * (n[2] + 1);
    const int64_t ncell = int64_t(n[0]) * n[1] * (dim == 3 ? n[2] : 1);
    if (nlat != S.n_nodes || ncell != S.n_elems) { why = "the mesh does not fill its lattice"; return false; }
    L.n_nodes = S.n_nodes;
    L.n_elems = S.n_elems;
    L.rank_of_lat.assign(size_t(nlat), -1);
    for (int64_t r = 0; r < S.n_nodes; r++) {
        const int64_t p = ijk[r * 3] + int64_t(n[0] + 1) * (ijk[r * 3 + 1] + int64_t(n[1] + 1) * ijk[r * 3 + 2]);
        if (L.rank_of_lat[p] >= 0) { why = "coincident nodes"; return false; }
        L.rank_of_lat[p] = int32_t(r);
    }
    L.elem_of_cell.assign(size_t(S.n_elems), -1);
    for (int64_t e = 0; e < S.n_elems; e++) {
        const int32_t* c = en + e * npe;
        const int32_t i0 = ijk[int64_t(c[0]) * 3], j0 = ijk[int64_t(c[0]) * 3 + 1], k0 = ijk[int64_t(c[0]) * 3 + 2];
        if (i0 >= n[0] || j0 >= n[1] || (dim == 3 && k0 >= n[2])) { why = "elements are not lattice cells"; return false; }
        for (int a = 0; a < npe; a++) {
            const int64_t p = (i0 + kCorner[a][0]) +
                              int64_t(n[0] + 1) * ((j0 + kCorner[a][1]) + int64_t(n[1] + 1) * (k0 + kCorner[a][2]));
            if (L.rank_of_lat[p] != c[a]) { why = "elements are not lattice cells in libMesh order"; return false; }
        }
        const int64_t cell = i0 + int64_t(n[0]) * (j0 + int64_t(n[1]) * k0);
        if (L.elem_of_cell[cell] >= 0) { why = "overlapping elements"; return false; }
        L.elem_of_cell[cell] = int32_t(e);
    }
    return true;
}

// The same lattice found from the connectivity alone (node coordinates may be
// displaced, e.g. a distorted mesh of a block): element 0's corners fix the
// lattice indices of its nodes, and every element that shares an indexed node
// places its corners at its base point + the libMesh corner offsets; the
// indices must agree everywhere and fill a box.
bool detect_lattice_topology(const Structure& S, int dim, MGLevelHost& L, std::string& why) {
    const int npe = dim == 3 ? 8 : 4;
    const int32_t* en = S.econn.data();
    const int64_t nn = S.n_nodes, ne = S.n_elems;
    std::vector<int64_t> ijk(size_t(nn) * 3, INT64_MIN);
    std::vector<std::vector<int32_t>> elems_of(static_cast<size_t>(nn));
    for (int64_t e = 0; e < ne; e++)
        for (int a = 0; a < npe; a++) elems_of[size_t(en[e * npe + a])].push_back(int32_t(e));
    std::vector<uint8_t> done(size_t(ne), 0);
    std::vector<int32_t> queue;
    queue.reserve(size_t(ne));
    queue.push_back(0);
    done[0] = 1;
    for (int a = 0; a < npe; a++)
        for (int d = 0; d < 3; d++) ijk[size_t(en[a]) * 3 + d] = kCorner[a][d];
    for (size_t qi = 0; qi < queue.size(); qi++) {
        const int32_t e = queue[qi];
        const int32_t* c = en + int64_t(e) * npe;
        // base point from an indexed corner (every queued element has one)
        int64_t b[3] = {0, 0, 0};
        int known = -1;
        for (int a = 0; a < npe && known < 0; a++)
            if (ijk[size_t(c[a]) * 3] != INT64_MIN) known = a;
        for (int d = 0; d < 3; d++) b[d] = ijk[size_t(c[known]) * 3 + d] - kCorner[known][d];
        for (int a = 0; a < npe; a++) {
            int64_t* t = &ijk[size_t(c[a]) * 3];
            for (int d = 0; d < 3; d++) {
                const int64_t v = b[d] + kCorner[a][d];
                if (t[0] != INT64_MIN && t[d] != v) { why = "elements are not lattice cells in libMesh order"; return false; }
            }
            for (int d = 0; d < 3; d++) t[d] = b[d] + kCorner[a][d];
            for (int32_t f : elems_of[size_t(c[a])])
                if (!done[size_t(f)]) {
                    done[size_t(f)] = 1;
                    queue.push_back(f);
                }
        }
    }
    if (int64_t(queue.size()) != ne) { why = "the mesh is not connected"; return false; }
    int64_t lo[3] = {INT64_MAX, INT64_MAX, INT64_MAX}, hi[3] = {INT64_MIN, INT64_MIN, INT64_MIN};
    for (int64_t r = 0; r < nn; r++) {
        if (ijk[size_t(r) * 3] == INT64_MIN) { why = "a node belongs to no element"; return false; }
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], ijk[size_t(r) * 3 + d]);
            hi[d] = std::max(hi[d], ijk[size_t(r) * 3 + d]);
        }
    }
    int n[3];
    for (int d = 0; d < 3; d++) {
        if (hi[d] - lo[d] > 2000000000LL) { why = "lattice too large"; return false; }
        n[d] = int(hi[d] - lo[d]);
    }
    if (dim == 2) n[2] = 0;
    L.dim = dim;
    L.npe = npe;
    L.nav = S.nav;
    L.nx = n[0]; L.ny = n[1]; L.nz = n[2];
    const int64_t nlat = int64_t(n[0] + 1) * (n[1] + 1) * (n[2] + 1);
    const int64_t ncell = int64_t(n[0]) * n[1] * (dim == 3 ? n[2] : 1);
    if (nlat != nn || ncell != ne) { why = "the mesh does not fill its lattice"; return false; }
    L.n_nodes = nn;
    L.n_elems = ne;
    L.rank_of_lat.assign(size_t(nlat), -1);
    for (int64_t r = 0; r < nn; r++) {
        const int64_t i = ijk[size_t(r) * 3] - lo[0], j = ijk[size_t(r) * 3 + 1] - lo[1],
                      k = ijk[size_t(r) * 3 + 2] - lo[2];
        const int64_t p = i + int64_t(n[0] + 1) * (j + int64_t(n[1] + 1) * k);
        if (L.rank_of_lat[size_t(p)] >= 0) { why = "coincident lattice points"; return false; }
        L.rank_of_lat[size_t(p)] = int32_t(r);
    }
    L.elem_of_cell.assign(size_t(ne), -1);
    for (int64_t e = 0; e < ne; e++) {
        const int32_t r0 = en[e * npe];
        const int64_t i0 = ijk[size_t(r0) * 3] - lo[0], j0 = ijk[size_t(r0) * 3 + 1] - lo[1],
                      k0 = ijk[size_t(r0) * 3 + 2] - lo[2];
        const int64_t cell = i0 + int64_t(n[0]) * (j0 + int64_t(n[1]) * k0);
        if (i0 >= n[0] || j0 >= n[1] || (dim == 3 && k0 >= n[2]) || L.elem_of_cell[size_t(cell)] >= 0) {
            why = "overlapping elements";
            return false;
        }
        L.elem_of_cell[size_t(cell)] = int32_t(e);
    }
    return true;
}

}  // namespace

MGHierarchy build_mg_hierarchy(const Structure& S) {
    MGHierarchy H;
    int dim = 0, physics = 0;
    if (S.npe == 8 && S.nav == 3) { dim = 3; physics = 1; }
    else if (S.npe == 4 && S.nav == 5) { dim = 2; physics = 2; }
    else if (S.npe == 8 && S.nav == 1) { dim = 3; physics = 3; }
    else { H.why = "multigrid is built for hex8 elasticity, the quad4 Mindlin plate and hex8 heat"; return H; }
    const int nav = S.nav, npe = S.npe;
    MGLevelHost L0;
    if (!detect_lattice(S, dim, L0, H.why)) {
        // displaced nodes: the lattice from the connectivity
        std::string why2;
        MGLevelHost T;
        if (!detect_lattice_topology(S, dim, T, why2)) {
            H.why = H.why + "; by connectivity: " + why2;
            return H;
        }
        L0 = std::move(T);
        H.why.clear();
    }
    H.levels.push_back(std::move(L0));
    // geometry of the problem mesh's lattice
    double lo[3] = {1e300, 1e300, 1e300}, h[3];
    for (int64_t r = 0; r < S.n_nodes; r++)
        for (int d = 0; d < 3; d++) lo[d] = std::min(lo[d], S.xyz_r[r * 3 + d]);
    const int opp = dim == 3 ? 6 : 2;
    for (int d = 0; d < 3; d++)
        h[d] = S.xyz_r[int64_t(S.econn[opp]) * 3 + d] - S.xyz_r[int64_t(S.econn[0]) * 3 + d];
    // constraint mask of the current finest-so-far level, by rank
    std::vector<uint8_t> con = S.constrained;
    for (int lvl = 1; lvl < 16; lvl++) {
        const MGLevelHost& F = H.levels.back();
        if (F.n_nodes * nav <= kMGCoarseTarget) break;
        if (F.nx % 2 || F.ny % 2 || (dim == 3 && F.nz % 2)) break;
        MGLevelHost C;
        C.dim = dim; C.nav = nav; C.npe = npe;
        C.nx = F.nx / 2; C.ny = F.ny / 2; C.nz = F.nz / 2;
        C.n_nodes = int64_t(C.nx + 1) * (C.ny + 1) * (C.nz + 1);
        C.n_elems = int64_t(C.nx) * C.ny * (dim == 3 ? C.nz : 1);
        const double scale = double(int64_t(1) << lvl);
        std::vector<double> xyz(size_t(C.n_nodes) * 3);
        std::vector<int64_t> base(size_t(C.n_nodes));
        std::vector<int64_t> cdofs;
        for (int k = 0; k <= C.nz; k++)
            for (int j = 0; j <= C.ny; j++)
                for (int i = 0; i <= C.nx; i++) {
                    const int64_t p = i + int64_t(C.nx + 1) * (j + int64_t(C.ny + 1) * k);
                    xyz[p * 3 + 0] = lo[0] + i * scale * h[0];
                    xyz[p * 3 + 1] = lo[1] + j * scale * h[1];
                    xyz[p * 3 + 2] = dim == 3 ? lo[2] + k * scale * h[2] : lo[2];
                    base[p] = p * nav;
                    const int64_t fp = 2 * i + int64_t(F.nx + 1) * (2 * j + int64_t(F.ny + 1) * 2 * k);
                    const int32_t fr = F.rank_of_lat[fp];
                    for (int v = 0; v < nav; v++)
                        if (con[size_t(fr) * nav + v]) cdofs.push_back(p * nav + v);
                }
        std::vector<int32_t> conn(size_t(C.n_elems) * npe);
        const int cz = dim == 3 ? C.nz : 1;
        for (int k = 0; k < cz; k++)
            for (int j = 0; j < C.ny; j++)
                for (int i = 0; i < C.nx; i++) {
                    const int64_t e = i + int64_t(C.nx) * (j + int64_t(C.ny) * k);
                    for (int a = 0; a < npe; a++)
                        conn[e * npe + a] = int32_t((i + kCorner[a][0]) +
                                                    int64_t(C.nx + 1) * ((j + kCorner[a][1]) +
                                                                         int64_t(C.ny + 1) * (k + kCorner[a][2])));
                }
        std::string err = build_structure(physics, C.n_nodes, C.n_elems, xyz.data(), conn.data(), nav, base.data(),
                                          int64_t(cdofs.size()), cdofs.data(), C.S);
        if (!err.empty()) { H.why = "coarse level: " + err; return H; }
        C.rank_of_lat.resize(size_t(C.n_nodes));
        for (int64_t p = 0; p < C.n_nodes; p++) C.rank_of_lat[p] = int32_t(p);
        C.elem_of_cell.resize(size_t(C.n_elems));
        for (int64_t e = 0; e < C.n_elems; e++) C.elem_of_cell[e] = int32_t(e);
        con = C.S.constrained;
        H.levels.push_back(std::move(C));
    }
    if (H.levels.size() < 2) { H.why = "the lattice does not coarsen (odd cell counts)"; return H; }
    H.coarsest_dofs = H.levels.back().n_nodes * nav;
    if (H.coarsest_dofs > kMGMaxCoarseDofs) {
        H.why = "coarsest level too large for the dense solve (" + std::to_string(H.coarsest_dofs) + " DOFs)";
        return H;
    }
    H.ok = true;
    return H;
}

void mg_class_child_products(const double* Khat, double* Gc) {
    // tri-linear weight of coarse corner A at fine local node a of child c
    auto q = [](int c, int a, int A) {
        double w = 1.0;
        for (int d = 0; d < 3; d++) {
            const int t = ((c >> d) & 1) + kCorner[a][d];
            w *= kCorner[A][d] ? 0.5 * t : 1.0 - 0.5 * t;
        }
        return w;
    };
    for (int c = 0; c < 8; c++) {
        double* G = Gc + c * 576;
        double Q[8][8];
        for (int a = 0; a < 8; a++)
            for (int A = 0; A < 8; A++) Q[a][A] = q(c, a, A);
        for (int A = 0; A < 8; A++)
            for (int B = 0; B < 8; B++)
                for (int v = 0; v < 3; v++)
                    for (int w = 0; w < 3; w++) {
                        double s = 0.0;
                        for (int a = 0; a < 8; a++) {
                            if (Q[a][A] == 0.0) continue;
                            double t = 0.0;
                            for (int b = 0; b < 8; b++) t += Khat[(a * 8 + b) * 9 + v * 3 + w] * Q[b][B];
                            s += Q[a][A] * t;
                        }
                        G[(A * 8 + B) * 9 + v * 3 + w] = s;
                    }
        // exact symmetry: row (A, v) > column (B, w) takes its mirror entry
        for (int A = 0; A < 8; A++)
            for (int B = 0; B < 8; B++)
                for (int v = 0; v < 3; v++)
                    for (int w = 0; w < 3; w++)
                        if (A * 3 + v > B * 3 + w) G[(A * 8 + B) * 9 + v * 3 + w] = G[(B * 8 + A) * 9 + w * 3 + v];
    }
}

void mg_dirty_coarse(const MGLevelHost& F, const MGLevelHost& C, const std::vector<uint8_t>& constrained,
                     std::vector<uint8_t>& dirty, std::vector<int32_t>& list) {
    dirty.assign(size_t(C.n_elems), 0);
    list.clear();
    const int64_t fx = F.nx + 1, fy = F.ny + 1;
    for (int64_t E = 0; E < C.n_elems; E++) {
        const int64_t I = E % C.nx, J = (E / C.nx) % C.ny, K = E / (int64_t(C.nx) * C.ny);
        bool d = false;
        for (int dk = 0; dk < 3 && !d; dk++)
            for (int dj = 0; dj < 3 && !d; dj++)
                for (int di = 0; di < 3 && !d; di++) {
                    const int64_t p = (2 * I + di) + fx * ((2 * J + dj) + fy * (2 * K + dk));
                    const int64_t r = F.rank_of_lat[size_t(p)];
                    for (int v = 0; v < 3; v++) d = d || constrained[size_t(r * 3 + v)] != 0;
                }
        if (d) {
            dirty[size_t(E)] = 1;
            list.push_back(int32_t(E));
        }
    }
}

}  // namespace mastg
