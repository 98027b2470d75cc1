// Structured workloads: the meshes, DOF maps, boundary conditions, loads and
// seeded design parameters of the benchmark configurations, produced as a
// mastg_problem_desc.  Meshes follow libMesh MeshTools::Generation::build_cube /
// build_square numbering (nodes i-fastest, elements i-fastest, libMesh local
// node and side order), which is what MAST's examples mesh with.
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "mastg/mastg.h"

struct mastg_workload {
    mastg_problem_desc d{};
    std::vector<double> xyz, loads, params;
    std::vector<int32_t> conn, side_elem, side_index, side_id;
    std::vector<int64_t> base, constrained;
};

namespace {

const int HEX_SIDE[6][4] = {{0, 3, 2, 1}, {0, 1, 5, 4}, {1, 2, 6, 5},
                            {2, 3, 7, 6}, {3, 0, 4, 7}, {4, 5, 6, 7}};
const int QUAD_SIDE[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};

void splitmix_uniform(uint64_t seed, int64_t n, double lo, double hi, double* out) {
    uint64_t s = seed;
    for (int64_t i = 0; i < n; i++) {
        s += 0x9E3779B97F4A7C15ULL;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        out[i] = lo + (hi - lo) * (double(z >> 11) * (1.0 / 9007199254740992.0));
    }
}

void cube(mastg_workload& w, int nx, int ny, int nz, double Lx, double Ly, double Lz) {
    const int64_t nn = int64_t(nx + 1) * (ny + 1) * (nz + 1), ne = int64_t(nx) * ny * nz;
    w.xyz.resize(nn * 3);
    int64_t id = 0;
    for (int k = 0; k <= nz; k++)
        for (int j = 0; j <= ny; j++)
            for (int i = 0; i <= nx; i++, id++) {
                w.xyz[id * 3 + 0] = (double(i) / double(nx)) * Lx;
                w.xyz[id * 3 + 1] = (double(j) / double(ny)) * Ly;
                w.xyz[id * 3 + 2] = (double(k) / double(nz)) * Lz;
            }
    auto idx = [&](int i, int j, int k) { return int32_t(i + (nx + 1) * (j + (ny + 1) * k)); };
    w.conn.resize(ne * 8);
    int64_t e = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++, e++) {
                int32_t* c = &w.conn[e * 8];
                c[0] = idx(i, j, k); c[1] = idx(i + 1, j, k); c[2] = idx(i + 1, j + 1, k);
                c[3] = idx(i, j + 1, k); c[4] = idx(i, j, k + 1); c[5] = idx(i + 1, j, k + 1);
                c[6] = idx(i + 1, j + 1, k + 1); c[7] = idx(i, j + 1, k + 1);
                auto side = [&](int s, int b) {
                    w.side_elem.push_back(int32_t(e)); w.side_index.push_back(s); w.side_id.push_back(b);
                };
                if (k == 0) side(0, 0);
                if (k == nz - 1) side(5, 5);
                if (j == 0) side(1, 1);
                if (j == ny - 1) side(3, 3);
                if (i == 0) side(4, 4);
                if (i == nx - 1) side(2, 2);
            }
    w.d.n_nodes = nn;
    w.d.n_elems = ne;
}

void square(mastg_workload& w, int nx, int ny, double Lx, double Ly) {
    const int64_t nn = int64_t(nx + 1) * (ny + 1), ne = int64_t(nx) * ny;
    w.xyz.resize(nn * 3);
    int64_t id = 0;
    for (int j = 0; j <= ny; j++)
        for (int i = 0; i <= nx; i++, id++) {
            w.xyz[id * 3 + 0] = (double(i) / double(nx)) * Lx;
            w.xyz[id * 3 + 1] = (double(j) / double(ny)) * Ly;
            w.xyz[id * 3 + 2] = 0.0;
        }
    auto idx = [&](int i, int j) { return int32_t(i + (nx + 1) * j); };
    w.conn.resize(ne * 4);
    int64_t e = 0;
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++, e++) {
            int32_t* c = &w.conn[e * 4];
            c[0] = idx(i, j); c[1] = idx(i + 1, j); c[2] = idx(i + 1, j + 1); c[3] = idx(i, j + 1);
            auto side = [&](int s, int b) {
                w.side_elem.push_back(int32_t(e)); w.side_index.push_back(s); w.side_id.push_back(b);
            };
            if (j == 0) side(0, 0);
            if (j == ny - 1) side(2, 2);
            if (i == 0) side(3, 3);
            if (i == nx - 1) side(1, 1);
        }
    w.d.n_nodes = nn;
    w.d.n_elems = ne;
}

// libMesh var-major first-touch DOF bases
void first_touch(mastg_workload& w, int npe, int nv) {
    w.base.assign(w.d.n_nodes, -1);
    int64_t next = 0;
    for (size_t i = 0; i < w.conn.size(); i++)
        if (w.base[w.conn[i]] < 0) { w.base[w.conn[i]] = next; next += nv; }
    for (auto& b : w.base)
        if (b < 0) { b = next; next += nv; }
    (void)npe;
}

// every variable in var_mask on every node of the sides with a boundary id in bid_mask
void dirichlet(mastg_workload& w, int dim, int nv, uint32_t bid_mask, uint32_t var_mask) {
    std::vector<uint8_t> mark(size_t(w.d.n_nodes) * nv, 0);
    const int npe = dim == 3 ? 8 : 4, nsn = dim == 3 ? 4 : 2;
    for (size_t s = 0; s < w.side_elem.size(); s++) {
        if (!((bid_mask >> w.side_id[s]) & 1u)) continue;
        for (int k = 0; k < nsn; k++) {
            const int ln = dim == 3 ? HEX_SIDE[w.side_index[s]][k] : QUAD_SIDE[w.side_index[s]][k];
            const int32_t n = w.conn[int64_t(w.side_elem[s]) * npe + ln];
            for (int v = 0; v < nv; v++)
                if ((var_mask >> v) & 1u) mark[size_t(n) * nv + v] = 1;
        }
    }
    for (int64_t n = 0; n < w.d.n_nodes; n++)
        for (int v = 0; v < nv; v++)
            if (mark[size_t(n) * nv + v]) w.constrained.push_back(w.base[n] + v);
}

// consistent nodal loads F_a = int_face N_a t_z dA of a uniform traction on the
// hex faces with boundary id `bid` (2x2 Gauss on the bilinear face)
void face_traction(mastg_workload& w, int nv, int bid, double tz) {
    const double g = 0.57735026918962576451;
    for (size_t s = 0; s < w.side_elem.size(); s++) {
        if (w.side_id[s] != bid) continue;
        const int32_t* c = &w.conn[int64_t(w.side_elem[s]) * 8];
        const int* sn = HEX_SIDE[w.side_index[s]];
        double P[4][3];
        for (int k = 0; k < 4; k++)
            for (int i = 0; i < 3; i++) P[k][i] = w.xyz[int64_t(c[sn[k]]) * 3 + i];
        for (int qj = 0; qj < 2; qj++)
            for (int qi = 0; qi < 2; qi++) {
                const double s1 = qi ? g : -g, s2 = qj ? g : -g;
                const double N[4] = {0.25 * (1 - s1) * (1 - s2), 0.25 * (1 + s1) * (1 - s2),
                                     0.25 * (1 + s1) * (1 + s2), 0.25 * (1 - s1) * (1 + s2)};
                const double d1[4] = {-0.25 * (1 - s2), 0.25 * (1 - s2), 0.25 * (1 + s2), -0.25 * (1 + s2)};
                const double d2[4] = {-0.25 * (1 - s1), -0.25 * (1 + s1), 0.25 * (1 + s1), 0.25 * (1 - s1)};
                double a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
                for (int k = 0; k < 4; k++)
                    for (int i = 0; i < 3; i++) { a[i] += P[k][i] * d1[k]; b[i] += P[k][i] * d2[k]; }
                const double cx = a[1] * b[2] - a[2] * b[1], cy = a[2] * b[0] - a[0] * b[2],
                             cz = a[0] * b[1] - a[1] * b[0];
                const double jxw = std::sqrt(cx * cx + cy * cy + cz * cz);
                for (int k = 0; k < 4; k++) w.loads[w.base[c[sn[k]]] + 2] += jxw * N[k] * tz;
            }
    }
}

}  // namespace

extern "C" {

int mastg_workload_create(int kind, int nx, int ny, int nz, uint64_t seed, double lo, double hi,
                          int mast_exact_vars, mastg_workload** out) {
    if (!out || nx <= 0) return MASTG_ERR_INVALID;
    *out = nullptr;
    mastg_workload* w = nullptr;
    try {
        w = new mastg_workload;
        mastg_problem_desc& d = w->d;
        if (kind == 1) {            // hex8 SIMP cantilever 1.0 x 0.5 x 0.5 m
            if (ny <= 0 || nz <= 0) { delete w; return MASTG_ERR_INVALID; }
            const int nv = mast_exact_vars ? 6 : 3;
            cube(*w, nx, ny, nz, 1.0, 0.5, 0.5);
            first_touch(*w, 8, nv);
            dirichlet(*w, 3, nv, 1u << 4, 0x3fu);        // x = 0 clamped
            w->loads.assign(size_t(d.n_nodes) * nv, 0.0);
            face_traction(*w, nv, 2, -1.0e6);            // t = -1e6 Pa in z on x = 1
            w->params.resize(d.n_elems);
            splitmix_uniform(seed, d.n_elems, lo, hi, w->params.data());
            d.physics = MASTG_HEX8_ELASTIC;
            d.n_vars = nv;
            d.E = 72e9; d.nu = 0.33; d.penalty = 3.0;
        } else if (kind == 2) {     // Mindlin plate 1 x 1 m
            const int nv = mast_exact_vars ? 6 : 5;
            square(*w, nx, ny > 0 ? ny : nx, 1.0, 1.0);
            first_touch(*w, 4, nv);
            dirichlet(*w, 2, nv, 0xfu, 0x7u);            // all edges: u, v, w
            w->params.resize(d.n_elems);
            splitmix_uniform(seed, d.n_elems, 0.002, 0.01, w->params.data());
            d.physics = MASTG_QUAD4_MINDLIN;
            d.n_vars = nv;
            d.E = 70e9; d.nu = 0.3; d.kappa = 5.0 / 6.0; d.offset = 0.0; d.pressure = 1.0e3;
        } else if (kind == 3) {     // heat conduction unit cube
            cube(*w, nx, nx, nx, 1.0, 1.0, 1.0);
            first_touch(*w, 8, 1);
            dirichlet(*w, 3, 1, 1u << 0, 1u);            // T = 0 on z = 0
            w->params.resize(d.n_elems);
            splitmix_uniform(seed, d.n_elems, 0.0, 1.0e6, w->params.data());
            d.physics = MASTG_HEX8_HEAT;
            d.n_vars = 1;
            d.k0 = 200.0; d.alpha = 1e-3; d.conv_h = 50.0; d.conv_tinf = 300.0;
            d.conv_bid_mask = 0x3eu;                     // boundary ids 1..5
        } else {
            delete w;
            return MASTG_ERR_INVALID;
        }
        d.xyz = w->xyz.data();
        d.conn = w->conn.data();
        d.node_dof_base = w->base.data();
        d.n_constrained = int64_t(w->constrained.size());
        d.constrained_dofs = w->constrained.data();
        d.n_sides = int64_t(w->side_elem.size());
        d.side_elem = w->side_elem.data();
        d.side_index = w->side_index.data();
        d.side_id = w->side_id.data();
        d.point_loads = w->loads.empty() ? nullptr : w->loads.data();
    } catch (const std::bad_alloc&) {
        delete w;
        return MASTG_ERR_OOM;
    }
    *out = w;
    return MASTG_OK;
}

const mastg_problem_desc* mastg_workload_desc(const mastg_workload* w) { return w ? &w->d : nullptr; }
const double* mastg_workload_params(const mastg_workload* w) {
    return w && !w->params.empty() ? w->params.data() : nullptr;
}
int mastg_workload_destroy(mastg_workload* w) {
    delete w;
    return MASTG_OK;
}

}  // extern "C"
