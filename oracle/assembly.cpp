// Oracle (test infrastructure only): problem set-up for the workloads, the
// NonlinearImplicitAssembly element loop with libMesh constraint semantics,
// and per-element adjoint sensitivity dots.
#include <stdexcept>
#include "oracle.h"
#include <cmath>
#include <cstring>
#include <algorithm>
#include <omp.h>

namespace oracle {

static void elem_coords(const Mesh& m, int64_t e, double X[][3]) {
    for (int a = 0; a < m.npe; a++) {
        int32_t n = m.conn[e * m.npe + a];
        for (int i = 0; i < 3; i++) X[a][i] = m.xyz[n * 3 + i];
    }
}

// Greedy element colouring: no two elements of one colour share a node, so a
// colour's scatter can run on all threads without write conflicts.
static std::vector<std::vector<int32_t>> colour_elements(const Mesh& m) {
    std::vector<uint64_t> used(m.n_nodes, 0);
    std::vector<std::vector<int32_t>> col;
    for (int64_t e = 0; e < m.n_elems; e++) {
        uint64_t mask = 0;
        for (int a = 0; a < m.npe; a++) mask |= used[m.conn[e * m.npe + a]];
        int c = 0;
        while ((mask >> c) & 1ull) c++;
        if (c >= 64) throw std::runtime_error("colouring overflow");
        if (int(col.size()) <= c) col.resize(c + 1);
        col[c].push_back(int32_t(e));
        for (int a = 0; a < m.npe; a++) used[m.conn[e * m.npe + a]] |= (1ull << c);
    }
    return col;
}

// plate section matrices for thickness h (and their d/dh when deriv)
static PlateSection plate_section(const Problem& P, double h, bool deriv) {
    PlateSection s;
    double Dps[9], G[4];
    material_D2_plane_stress(P.E, P.nu, Dps);
    material_G(P.E, P.nu, G);
    if (!deriv) {
        section_A(Dps, h, s.A);
        section_B(Dps, h, P.off, s.B);
        section_Db(Dps, h, P.off, s.Db);
        section_S(G, h, P.kappa, s.S);
    } else {
        // ExtensionStiffnessMatrix / ExtensionBendingStiffnessMatrix /
        // BendingStiffnessMatrix / TransverseStiffnessMatrix ::derivative with dh=1
        for (int k = 0; k < 9; k++) {
            s.A[k] = Dps[k] * 1.0;
            s.B[k] = Dps[k] * (1.0 * P.off);
            s.Db[k] = Dps[k] * (std::pow(h, 2) / 4. * 1.0 + 1.0 * std::pow(P.off, 2));
        }
        for (int k = 0; k < 4; k++) s.S[k] = G[k] * (1.0 * P.kappa);
    }
    return s;
}

// Element residual / Jacobian for problem physics, MAST-exact variable set
// restricted to the first nv variables. f: nv*npe, K: (nv*npe)^2.
static void element_calc(const Problem& P, int64_t e, const double* sol, double* f, double* K) {
    const Mesh& m = P.mesh;
    const int nv = P.dof.n_vars;
    double X[8][3];
    elem_coords(m, e, X);
    if (P.phys == HEX8_ELASTIC) {
        double D[36], u[48] = {0}, ff[48] = {0};
        static thread_local double KK[48 * 48];
        const double rho = P.param[e];
        material_D3(std::pow(rho, P.penalty) * P.E, P.nu, D);
        for (int i = 0; i < nv * 8; i++) u[i] = sol[i];
        if (K) std::memset(KK, 0, sizeof(KK));
        hex8_elastic(X, D, u, ff, K ? KK : nullptr);
        if (P.pres_bid_mask) {
            // pressure on this element's boundary sides in the mask
            // (the mesh lists sides in non-decreasing element order)
            auto lo = std::lower_bound(m.bs_elem.begin(), m.bs_elem.end(), int32_t(e));
            for (auto it = lo; it != m.bs_elem.end() && *it == e; ++it) {
                const size_t s = size_t(it - m.bs_elem.begin());
                if ((P.pres_bid_mask >> m.bs_id[s]) & 1u) hex8_pressure(X, m.bs_side[s], P.pressure, ff);
            }
        }
        for (int i = 0; i < nv * 8; i++) {
            f[i] += ff[i];
            if (K)
                for (int j = 0; j < nv * 8; j++) K[i * nv * 8 + j] += KK[i * 48 + j];
        }
    } else if (P.phys == QUAD4_MINDLIN) {
        double u[24] = {0}, ff[24] = {0};
        static thread_local double KK[24 * 24];
        for (int i = 0; i < nv * 4; i++) u[i] = sol[i];
        PlateSection s = plate_section(P, P.param[e], false);
        if (K) std::memset(KK, 0, sizeof(KK));
        double T[9], XL[4][3];
        if (nv == 6 && plate_local_frame(X, T, XL)) {
            // out of the x-y plane: the element in its local frame, state in,
            // vector and matrix out through T (the 6-variable system; the
            // 5-variable layout is the x-y-plane restriction)
            double ul[24], fl[24] = {0};
            static thread_local double KL[24 * 24];
            plate_to_local(T, u, ul);
            if (K) std::memset(KL, 0, sizeof(KL));
            quad4_mindlin(XL, s, ul, fl, K ? KL : nullptr, true);
            quad4_pressure(XL, P.pressure, fl);
            plate_to_global(T, fl, ff);
            if (K) plate_matrix_to_global(T, KL, KK);
        } else {
            quad4_mindlin(X, s, u, ff, K ? KK : nullptr, nv == 6);
            quad4_pressure(X, P.pressure, ff);
        }
        for (int i = 0; i < nv * 4; i++) {
            f[i] += ff[i];
            if (K)
                for (int j = 0; j < nv * 4; j++) K[i * nv * 4 + j] += KK[i * 24 + j];
        }
    } else {
        hex8_heat(X, P.k0, P.alpha, sol, f, K);
        // convection on this element's boundary sides in the mask
        // (the mesh lists sides in non-decreasing element order)
        auto lo = std::lower_bound(m.bs_elem.begin(), m.bs_elem.end(), int32_t(e));
        for (auto it = lo; it != m.bs_elem.end() && *it == e; ++it) {
            const size_t s = size_t(it - m.bs_elem.begin());
            if ((P.conv_bid_mask >> m.bs_id[s]) & 1u)
                hex8_convection(X, m.bs_side[s], P.conv_h, P.conv_Tinf, sol, f, K);
            if ((P.rad_bid_mask >> m.bs_id[s]) & 1u)
                hex8_radiation(X, m.bs_side[s], P.rad_sigma, P.rad_emiss, P.rad_Tamb, P.rad_T0, sol, f, K);
        }
        hex8_source(X, P.param[e], f);
    }
}

static void constrain(const Problem& P, const int64_t* dofs, int nd, double* f, double* K) {
    // libMesh DofMap::constrain_element_matrix_and_vector (libMesh 1.5.1,
    // src/base/dof_map_constraints.C; called by MAST per element at
    // src/base/nonlinear_implicit_assembly.cpp:174-183).  libMesh builds the
    // element constraint matrix C (build_constraint_matrix): for a homogeneous
    // Dirichlet DOF its row is zero, so K <- C^T K C zeroes that row and
    // column and F <- C^T F zeroes the entry; it then sets K(i,i) = 1 for every
    // constrained element DOF i.  The global diagonal of a constrained DOF is
    // therefore the number of elements that contain it.
    for (int i = 0; i < nd; i++) {
        if (!P.dof.constrained[dofs[i]]) continue;
        if (f) f[i] = 0.0;
        if (K) {
            for (int j = 0; j < nd; j++) { K[i * nd + j] = 0.0; K[j * nd + i] = 0.0; }
            K[i * nd + i] = 1.0;
        }
    }
}

void assemble(const Problem& P, const double* X, double* R, CSR* K, int n_threads) {
    const Mesh& m = P.mesh;
    const int nd = m.npe * P.dof.n_vars;
    if (R) std::fill(R, R + P.dof.n_dofs, 0.0);
    if (K) std::fill(K->val.begin(), K->val.end(), 0.0);
    auto colours = colour_elements(m);
    if (n_threads <= 0) n_threads = omp_get_max_threads();
    for (auto& cl : colours) {
#pragma omp parallel num_threads(n_threads)
        {
            std::vector<int64_t> dofs(nd);
            std::vector<double> sol(nd), f(nd), Ke(K ? nd * nd : 0);
#pragma omp for schedule(static)
            for (int64_t t = 0; t < int64_t(cl.size()); t++) {
                const int64_t e = cl[t];
                P.dof.elem_dofs(m, e, dofs.data());
                for (int i = 0; i < nd; i++) sol[i] = X ? X[dofs[i]] : 0.0;
                std::fill(f.begin(), f.end(), 0.0);
                if (K) std::fill(Ke.begin(), Ke.end(), 0.0);
                element_calc(P, e, sol.data(), f.data(), K ? Ke.data() : nullptr);
                constrain(P, dofs.data(), nd, f.data(), K ? Ke.data() : nullptr);
                if (R)
                    for (int i = 0; i < nd; i++) R[dofs[i]] += f[i];
                if (K)
                    for (int i = 0; i < nd; i++)
                        for (int j = 0; j < nd; j++) {
                            int64_t pos = K->find(dofs[i], dofs[j]);  // MatSetValues row search
                            if (pos < 0) throw std::runtime_error("entry outside sparsity");
                            K->val[pos] += Ke[i * nd + j];
                        }
            }
        }
    }
    // point loads: R -= F on unconstrained dofs (nonlinear_implicit_assembly.cpp:190-247)
    if (R && !P.point_load.empty())
        for (int64_t i = 0; i < P.dof.n_dofs; i++)
            if (!P.dof.constrained[i]) R[i] -= P.point_load[i];
}

void sensitivity_dots(const Problem& P, const double* X, const double* lambda, double* g,
                      int n_threads) {
    const Mesh& m = P.mesh;
    const int nv = P.dof.n_vars, nd = m.npe * nv;
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel num_threads(n_threads)
    {
        std::vector<int64_t> dofs(nd);
        std::vector<double> sol(nd), df(nd);
#pragma omp for schedule(static)
        for (int64_t e = 0; e < m.n_elems; e++) {
            P.dof.elem_dofs(m, e, dofs.data());
            for (int i = 0; i < nd; i++) sol[i] = X[dofs[i]];
            std::fill(df.begin(), df.end(), 0.0);
            double Xc[8][3];
            elem_coords(m, e, Xc);
            if (P.phys == HEX8_ELASTIC) {
                // SIMP: E_e = rho^p E0, dE/drho = p rho^(p-1) E0, dnu/drho = 0
                const double rho = P.param[e];
                const double Ee = std::pow(rho, P.penalty) * P.E;
                const double dE = P.penalty * std::pow(rho, P.penalty - 1.0) * P.E;
                double dD[36], u[48] = {0}, ff[48] = {0};
                material_dD3(Ee, P.nu, dE, 0.0, dD);
                for (int i = 0; i < nd; i++) u[i] = sol[i];
                hex8_elastic_sens(Xc, dD, u, ff);
                for (int i = 0; i < nd; i++) df[i] = ff[i];
            } else if (P.phys == QUAD4_MINDLIN) {
                double u[24] = {0}, ff[24] = {0};
                for (int i = 0; i < nd; i++) u[i] = sol[i];
                PlateSection ds = plate_section(P, P.param[e], true);
                double Xq[4][3], T[9], XL[4][3];
                for (int n = 0; n < 4; n++)
                    for (int d = 0; d < 3; d++) Xq[n][d] = Xc[n][d];
                if (nd == 24 && plate_local_frame(Xq, T, XL)) {
                    double ul[24], fl[24] = {0};
                    plate_to_local(T, u, ul);
                    quad4_mindlin_sens(XL, ds, ul, fl);
                    plate_to_global(T, fl, ff);
                } else {
                    quad4_mindlin_sens(Xc, ds, u, ff);
                }
                for (int i = 0; i < nd; i++) df[i] = ff[i];
            } else {
                // dR/dq_e = -int N (volume_heat_source_residual_sensitivity)
                hex8_source(Xc, 1.0, df.data());
            }
            constrain(P, dofs.data(), nd, df.data(), nullptr);
            double s = 0;
            for (int i = 0; i < nd; i++) s += lambda[dofs[i]] * df[i];
            g[e] = s;
        }
    }
}

// ----------------------------------------------------------- workloads ----
Problem make_cantilever(int nx, int ny, int nz, uint64_t seed, double rho_lo, double rho_hi,
                        bool mast_exact_vars) {
    Problem P;
    P.phys = HEX8_ELASTIC;
    P.mesh = build_cube(nx, ny, nz, 0.0, 1.0, 0.0, 0.5, 0.0, 0.5);
    P.dof = build_dofmap(P.mesh, mast_exact_vars ? 6 : 3);
    P.active_only = !mast_exact_vars;
    add_dirichlet(P.mesh, P.dof, 1u << 4, 0x3f);   // x=0 face (boundary id 4) clamped
    P.E = 72e9; P.nu = 0.33; P.penalty = 3.0;
    P.param.resize(P.mesh.n_elems);
    uniform_fill(seed, P.mesh.n_elems, rho_lo, rho_hi, P.param.data());
    // uniform traction t = (0,0,-1e6) Pa on the x=1 face (boundary id 2) as
    // consistent nodal point loads F_a = int N_a t dA (2x2 Gauss on each face)
    P.point_load.assign(P.dof.n_dofs, 0.0);
    const double tz = -1.0e6;
    for (size_t s = 0; s < P.mesh.bs_elem.size(); s++) {
        if (P.mesh.bs_id[s] != 2) continue;
        const int64_t e = P.mesh.bs_elem[s];
        double X[8][3];
        elem_coords(P.mesh, e, X);
        double fa[8] = {0};
        // the convection routine with h=1, Tinf=-1, T=0 integrates int N
        double T0[8] = {0};
        hex8_convection(X, P.mesh.bs_side[s], 1.0, -1.0, T0, fa, nullptr);
        for (int a = 0; a < 8; a++) {
            const int32_t n = P.mesh.conn[e * 8 + a];
            P.point_load[P.dof.node_base[n] + 2] += fa[a] * tz;
        }
    }
    return P;
}

Problem make_plate(int nx, int ny, uint64_t seed, bool mast_exact_vars) {
    Problem P;
    P.phys = QUAD4_MINDLIN;
    P.mesh = build_square(nx, ny, 0.0, 1.0, 0.0, 1.0);
    P.dof = build_dofmap(P.mesh, mast_exact_vars ? 6 : 5);
    P.active_only = !mast_exact_vars;
    add_dirichlet(P.mesh, P.dof, 0xf, 0x7);        // all edges: u, v, w = 0
    P.E = 70e9; P.nu = 0.3; P.kappa = 5.0 / 6.0; P.off = 0.0;
    P.pressure = 1.0e3;
    P.param.resize(P.mesh.n_elems);
    uniform_fill(seed, P.mesh.n_elems, 0.002, 0.01, P.param.data());
    return P;
}

Problem make_heat(int n, uint64_t seed) {
    Problem P;
    P.phys = HEX8_HEAT;
    P.mesh = build_cube(n, n, n, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0);
    P.dof = build_dofmap(P.mesh, 1);
    add_dirichlet(P.mesh, P.dof, 1u << 0, 1u);      // T = 0 on z=0 (boundary id 0)
    P.k0 = 200.0; P.alpha = 1e-3; P.conv_h = 50.0; P.conv_Tinf = 300.0;
    P.conv_bid_mask = 0x3e;                         // boundary ids 1..5
    P.param.resize(P.mesh.n_elems);
    uniform_fill(seed, P.mesh.n_elems, 0.0, 1.0e6, P.param.data());
    return P;
}

}  // namespace oracle
