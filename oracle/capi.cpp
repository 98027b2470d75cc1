// Oracle (test infrastructure only): C entry points for the Python tests and
// the CPU-baseline leg of bench.py (loaded with ctypes from liboracle.so).
#include <algorithm>
#include "oracle.h"
#include <cstring>
#include <memory>
#include <omp.h>
#include <string>

using namespace oracle;

namespace {
struct Handle {
    Problem P;
    std::unique_ptr<CSR> K;
    std::vector<std::vector<int32_t>> node_elems;
    int threads = 0;
    CSR& csr() {
        if (!K) K.reset(new CSR(build_sparsity(P.mesh, P.dof)));
        return *K;
    }
};
Handle* H(void* h) { return static_cast<Handle*>(h); }
}  // namespace

extern "C" {

// kind: 1 cantilever hex8 (nx,ny,nz,lo,hi), 2 Mindlin plate (nx,ny), 3 heat cube (nx)
void* oracle_problem_new(int kind, int nx, int ny, int nz, uint64_t seed, double lo, double hi,
                         int mast_exact) {
    auto* h = new Handle;
    if (kind == 1) h->P = make_cantilever(nx, ny, nz, seed, lo, hi, mast_exact != 0);
    else if (kind == 2) h->P = make_plate(nx, ny, seed, mast_exact != 0);
    else h->P = make_heat(nx, seed);
    return h;
}
// A problem over caller arrays (any mesh and DofMap): kind 1 hex8 elasticity,
// 2 Mindlin plate, 3 heat; constrained: [n_dofs] 0/1; point_load [n_dofs] or
// null; material scalars as the structured builders' defaults, changed with
// oracle_set_scalar.  The plate pressure starts at 0.
void* oracle_problem_from_arrays(int kind, int64_t n_nodes, const double* xyz, int64_t n_elems,
                                 const int32_t* conn, int n_vars, const int64_t* node_base, int64_t n_dofs,
                                 const uint8_t* constrained, const double* params, const double* point_load) {
    auto* h = new Handle;
    Problem& P = h->P;
    P.phys = kind == 1 ? HEX8_ELASTIC : kind == 2 ? QUAD4_MINDLIN : HEX8_HEAT;
    P.mesh.dim = kind == 2 ? 2 : 3;
    P.mesh.npe = kind == 2 ? 4 : 8;
    P.mesh.n_nodes = n_nodes;
    P.mesh.n_elems = n_elems;
    P.mesh.xyz.assign(xyz, xyz + n_nodes * 3);
    P.mesh.conn.assign(conn, conn + n_elems * P.mesh.npe);
    P.dof.n_vars = n_vars;
    P.dof.n_dofs = n_dofs;
    P.dof.node_base.assign(node_base, node_base + n_nodes);
    P.dof.constrained.assign(constrained, constrained + n_dofs);
    P.param.assign(params, params + n_elems);
    if (point_load) P.point_load.assign(point_load, point_load + n_dofs);
    if (kind == 2) { P.E = 70e9; P.nu = 0.3; P.pressure = 0.0; }
    P.active_only = (kind == 1 && n_vars == 3) || (kind == 2 && n_vars == 5);
    return h;
}
// boundary sides (element, libMesh side, boundary id) of a problem built from
// arrays; kept in non-decreasing element order as the element loop expects
void oracle_set_bsides(void* h, int64_t n, const int32_t* elem, const int32_t* side, const int32_t* id) {
    Mesh& m = H(h)->P.mesh;
    std::vector<int64_t> ord(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; i++) ord[size_t(i)] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) { return elem[a] < elem[b]; });
    m.bs_elem.resize(size_t(n));
    m.bs_side.resize(size_t(n));
    m.bs_id.resize(size_t(n));
    for (int64_t i = 0; i < n; i++) {
        m.bs_elem[size_t(i)] = elem[ord[size_t(i)]];
        m.bs_side[size_t(i)] = side[ord[size_t(i)]];
        m.bs_id[size_t(i)] = id[ord[size_t(i)]];
    }
}
void oracle_problem_free(void* h) { delete H(h); }
void oracle_set_threads(void* h, int t) { H(h)->threads = t; }
int oracle_max_threads() { return omp_get_max_threads(); }

// set a scalar of the problem: E nu penalty kappa off k0 alpha conv_h conv_Tinf
// conv_mask pressure pres_mask rad_sigma rad_emiss rad_Tamb rad_T0 rad_mask; returns 0 on
// success, -1 for an unknown name
int oracle_set_scalar(void* h, const char* name, double v) {
    Problem& P = H(h)->P;
    const std::string n(name);
    if (n == "E") P.E = v;
    else if (n == "nu") P.nu = v;
    else if (n == "penalty") P.penalty = v;
    else if (n == "kappa") P.kappa = v;
    else if (n == "off") P.off = v;
    else if (n == "k0") P.k0 = v;
    else if (n == "alpha") P.alpha = v;
    else if (n == "conv_h") P.conv_h = v;
    else if (n == "conv_Tinf") P.conv_Tinf = v;
    else if (n == "conv_mask") P.conv_bid_mask = uint32_t(v);
    else if (n == "pressure") P.pressure = v;
    else if (n == "pres_mask") P.pres_bid_mask = uint32_t(v);
    else if (n == "rad_sigma") P.rad_sigma = v;
    else if (n == "rad_emiss") P.rad_emiss = v;
    else if (n == "rad_Tamb") P.rad_Tamb = v;
    else if (n == "rad_T0") P.rad_T0 = v;
    else if (n == "rad_mask") P.rad_bid_mask = uint32_t(v);
    else return -1;
    return 0;
}

int64_t oracle_n_nodes(void* h) { return H(h)->P.mesh.n_nodes; }
int64_t oracle_n_elems(void* h) { return H(h)->P.mesh.n_elems; }
int64_t oracle_n_dofs(void* h) { return H(h)->P.dof.n_dofs; }
int oracle_n_vars(void* h) { return H(h)->P.dof.n_vars; }
int oracle_npe(void* h) { return H(h)->P.mesh.npe; }
int64_t oracle_n_bsides(void* h) { return int64_t(H(h)->P.mesh.bs_elem.size()); }

void oracle_get_mesh(void* h, double* xyz, int32_t* conn) {
    const Mesh& m = H(h)->P.mesh;
    std::memcpy(xyz, m.xyz.data(), m.xyz.size() * sizeof(double));
    std::memcpy(conn, m.conn.data(), m.conn.size() * sizeof(int32_t));
}
void oracle_get_bsides(void* h, int32_t* elem, int32_t* side, int32_t* id) {
    const Mesh& m = H(h)->P.mesh;
    for (size_t s = 0; s < m.bs_elem.size(); s++) {
        elem[s] = m.bs_elem[s]; side[s] = m.bs_side[s]; id[s] = m.bs_id[s];
    }
}
void oracle_get_dofmap(void* h, int64_t* node_base, uint8_t* constrained) {
    const DofMap& d = H(h)->P.dof;
    std::memcpy(node_base, d.node_base.data(), d.node_base.size() * sizeof(int64_t));
    std::memcpy(constrained, d.constrained.data(), d.constrained.size());
}
void oracle_get_params(void* h, double* p) {
    std::memcpy(p, H(h)->P.param.data(), H(h)->P.param.size() * sizeof(double));
}
void oracle_set_params(void* h, const double* p) {
    std::memcpy(H(h)->P.param.data(), p, H(h)->P.param.size() * sizeof(double));
}
void oracle_get_point_load(void* h, double* f) {
    const auto& pl = H(h)->P.point_load;
    if (pl.empty()) std::memset(f, 0, H(h)->P.dof.n_dofs * sizeof(double));
    else std::memcpy(f, pl.data(), pl.size() * sizeof(double));
}

int64_t oracle_csr_nnz(void* h) { return H(h)->csr().nnz(); }
void oracle_csr(void* h, int64_t* row_ptr, int32_t* col) {
    CSR& K = H(h)->csr();
    std::memcpy(row_ptr, K.row_ptr.data(), K.row_ptr.size() * sizeof(int64_t));
    std::memcpy(col, K.col.data(), K.col.size() * sizeof(int32_t));
}
// one sparsity row computed from the elements around the node (no full CSR)
int64_t oracle_sparsity_row(void* h, int64_t dof, int32_t* out) {
    Handle* hh = H(h);
    if (hh->node_elems.empty()) hh->node_elems = node_to_elems(hh->P.mesh);
    auto row = sparsity_row(hh->P.mesh, hh->P.dof, hh->node_elems, dof);
    if (out) std::memcpy(out, row.data(), row.size() * sizeof(int32_t));
    return int64_t(row.size());
}

// residual (n_dofs) and Jacobian values in the oracle CSR order; X may be null (= 0)
void oracle_assemble(void* h, const double* X, double* R, double* Kval) {
    Handle* hh = H(h);
    CSR& K = hh->csr();
    assemble(hh->P, X, R, Kval ? &K : nullptr, hh->threads);
    if (Kval) std::memcpy(Kval, K.val.data(), K.val.size() * sizeof(double));
}

void oracle_sens_dots(void* h, const double* X, const double* lam, double* g) {
    sensitivity_dots(H(h)->P, X, lam, g, H(h)->threads);
}

// CG + Jacobi on the matrix from the last oracle_assemble call
int oracle_cg(void* h, const double* b, double* x, double rtol, int max_it, double* rel_res) {
    SolveStats st = cg_jacobi(H(h)->csr(), b, x, rtol, max_it, H(h)->threads);
    if (rel_res) *rel_res = st.rel_res;
    return st.converged ? st.iters : -st.iters;
}

void oracle_spmv(void* h, const double* x, double* y) { spmv(H(h)->csr(), x, y, H(h)->threads); }

int oracle_solve_elastic(void* h, double rtol, int max_it, double* u, double* C, double* dc,
                         double* rel_res) {
    ElasticResult r = solve_elastic(H(h)->P, rtol, max_it, H(h)->threads, &H(h)->csr());
    if (u) std::memcpy(u, r.u.data(), r.u.size() * sizeof(double));
    if (C) *C = r.compliance;
    if (dc) std::memcpy(dc, r.dc.data(), r.dc.size() * sizeof(double));
    if (rel_res) *rel_res = r.st.rel_res;
    return r.st.converged ? r.st.iters : -r.st.iters;
}

int oracle_solve_heat(void* h, double newton_rtol, double lin_rtol, double* T, double* meanT,
                      double* dJ, double* res_hist, int hist_cap) {
    HeatResult r = solve_heat(H(h)->P, newton_rtol, lin_rtol, H(h)->threads);
    if (T) std::memcpy(T, r.T.data(), r.T.size() * sizeof(double));
    if (meanT) *meanT = r.mean_T;
    if (dJ) std::memcpy(dJ, r.dJ.data(), r.dJ.size() * sizeof(double));
    if (res_hist)
        for (int i = 0; i < hist_cap && i < int(r.res_hist.size()); i++) res_hist[i] = r.res_hist[i];
    return r.newton_iters;
}

// -------------------------------------------------- single-element access ----
void oracle_material_D3(double E, double nu, double* D) { material_D3(E, nu, D); }
void oracle_material_dD3(double E, double nu, double dE, double dnu, double* D) {
    material_dD3(E, nu, dE, dnu, D);
}
void oracle_material_D2(double E, double nu, double* D) { material_D2_plane_stress(E, nu, D); }
void oracle_material_dD2(double E, double nu, double dE, double dnu, double* D) {
    material_dD2_plane_stress(E, nu, dE, dnu, D);
}
void oracle_section(double E, double nu, double h, double off, double kappa, double* A,
                    double* B, double* Db, double* S) {
    double Dps[9], G[4];
    material_D2_plane_stress(E, nu, Dps);
    material_G(E, nu, G);
    section_A(Dps, h, A); section_B(Dps, h, off, B); section_Db(Dps, h, off, Db);
    section_S(G, h, kappa, S);
}
// stresses [8][6] and JxW [8] of a hex8 at E, nu for the 24 translations u
void oracle_hex8_stress(const double* X, double E, double nu, const double* u, double* sigma, double* jxw) {
    double D[36];
    material_D3(E, nu, D);
    hex8_stress(reinterpret_cast<const double(*)[3]>(X), D, u, reinterpret_cast<double(*)[6]>(sigma), jxw);
}
void oracle_hex8_radiation(const double* X, int side, double sigma, double emiss, double Tamb,
                           double T0, const double* T, double* f, double* K) {
    hex8_radiation(reinterpret_cast<const double(*)[3]>(X), side, sigma, emiss, Tamb, T0, T, f, K);
}
// X: 8x3, u: 48, f: 48 (+=), K: 48x48 (+=, may be null)
void oracle_hex8_elastic(const double* X, double E, double nu, const double* u, double* f,
                         double* K) {
    double D[36];
    material_D3(E, nu, D);
    hex8_elastic(reinterpret_cast<const double(*)[3]>(X), D, u, f, K);
}
void oracle_hex8_elastic_sens(const double* X, const double* dD, const double* u, double* df) {
    hex8_elastic_sens(reinterpret_cast<const double(*)[3]>(X), dD, u, df);
}
// X: 4x3, u: 24; section from (E,nu,h,off,kappa)
void oracle_quad4_mindlin(const double* X, double E, double nu, double h, double off,
                          double kappa, const double* u, double* f, double* K) {
    PlateSection s;
    double Dps[9], G[4];
    material_D2_plane_stress(E, nu, Dps);
    material_G(E, nu, G);
    section_A(Dps, h, s.A); section_B(Dps, h, off, s.B); section_Db(Dps, h, off, s.Db);
    section_S(G, h, kappa, s.S);
    quad4_mindlin(reinterpret_cast<const double(*)[3]>(X), s, u, f, K, true);
}
void oracle_hex8_heat(const double* X, double k0, double alpha, const double* T, double* f,
                      double* K) {
    hex8_heat(reinterpret_cast<const double(*)[3]>(X), k0, alpha, T, f, K);
}
void oracle_hex8_pressure(const double* X, int side, double p, double* f) {
    hex8_pressure(reinterpret_cast<const double(*)[3]>(X), side, p, f);
}
void oracle_hex8_convection(const double* X, int side, double h, double Tinf, const double* T,
                            double* f, double* K) {
    hex8_convection(reinterpret_cast<const double(*)[3]>(X), side, h, Tinf, T, f, K);
}
void oracle_uniform(uint64_t seed, int64_t n, double lo, double hi, double* out) {
    uniform_fill(seed, n, lo, hi, out);
}

}  // extern "C"
