// Oracle (test infrastructure only): Krylov solvers and the static / steady
// drivers (solve, compliance, adjoint sensitivities).
#include "oracle.h"
#include <cmath>
#include <algorithm>
#include <omp.h>

namespace oracle {

static int nthr(int n) { return n > 0 ? n : omp_get_max_threads(); }

void spmv(const CSR& A, const double* x, double* y, int n_threads) {
#pragma omp parallel for schedule(static) num_threads(nthr(n_threads))
    for (int64_t r = 0; r < A.n; r++) {
        double s = 0;
        for (int64_t k = A.row_ptr[r]; k < A.row_ptr[r + 1]; k++) s += A.val[k] * x[A.col[k]];
        y[r] = s;
    }
}

// fixed 4096-element chunks summed in chunk order: the result does not depend
// on the thread count or on the order threads finish (an OpenMP reduction
// clause combines per-thread sums in completion order)
static double dot(const double* a, const double* b, int64_t n, int t) {
    constexpr int64_t kChunk = 4096;
    const int64_t nc = (n + kChunk - 1) / kChunk;
    std::vector<double> part(size_t(std::max<int64_t>(nc, 1)), 0.0);
#pragma omp parallel for schedule(static) num_threads(nthr(t))
    for (int64_t c = 0; c < nc; c++) {
        double s = 0;
        const int64_t e = std::min(n, (c + 1) * kChunk);
        for (int64_t i = c * kChunk; i < e; i++) s += a[i] * b[i];
        part[size_t(c)] = s;
    }
    double s = 0;
    for (int64_t c = 0; c < nc; c++) s += part[size_t(c)];
    return s;
}

// PETSc KSPCG + PCJACOBI semantics, stopping on the true (unpreconditioned)
// relative residual ||b - A x||_2 / ||b||_2 <= rtol (x0 = 0).
SolveStats cg_jacobi(const CSR& A, const double* b, double* x, double rtol, int max_it,
                     int n_threads) {
    const int64_t n = A.n;
    const int t = nthr(n_threads);
    std::vector<double> r(n), z(n), p(n), q(n), dinv(n);
    for (int64_t i = 0; i < n; i++) {
        double d = A.val[A.find(i, i)];
        dinv[i] = d != 0 ? 1.0 / d : 1.0;
    }
    std::fill(x, x + n, 0.0);
    for (int64_t i = 0; i < n; i++) { r[i] = b[i]; z[i] = dinv[i] * r[i]; p[i] = z[i]; }
    const double bn = std::sqrt(dot(b, b, n, t));
    SolveStats st;
    if (bn == 0) { st.converged = true; return st; }
    double rz = dot(r.data(), z.data(), n, t);
    for (int it = 1; it <= max_it; it++) {
        spmv(A, p.data(), q.data(), t);
        const double alpha = rz / dot(p.data(), q.data(), n, t);
#pragma omp parallel for schedule(static) num_threads(t)
        for (int64_t i = 0; i < n; i++) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
            z[i] = dinv[i] * r[i];
        }
        const double rr = std::sqrt(dot(r.data(), r.data(), n, t));
        st.iters = it;
        st.rel_res = rr / bn;
        if (st.rel_res <= rtol) { st.converged = true; break; }
        const double rz1 = dot(r.data(), z.data(), n, t);
        const double beta = rz1 / rz;
        rz = rz1;
#pragma omp parallel for schedule(static) num_threads(t)
        for (int64_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    }
    return st;
}

// Right-preconditioned BiCGStab with Jacobi (PETSc KSPBCGS + PCJACOBI role)
SolveStats bicgstab_jacobi(const CSR& A, const double* b, double* x, double rtol, int max_it,
                           int n_threads) {
    const int64_t n = A.n;
    const int t = nthr(n_threads);
    std::vector<double> r(n), r0(n), p(n), v(n), s(n), tt(n), ph(n), sh(n), dinv(n);
    for (int64_t i = 0; i < n; i++) {
        double d = A.val[A.find(i, i)];
        dinv[i] = d != 0 ? 1.0 / d : 1.0;
    }
    std::fill(x, x + n, 0.0);
    for (int64_t i = 0; i < n; i++) { r[i] = b[i]; r0[i] = b[i]; p[i] = 0; v[i] = 0; }
    const double bn = std::sqrt(dot(b, b, n, t));
    SolveStats st;
    if (bn == 0) { st.converged = true; return st; }
    double rho = 1, alpha = 1, omega = 1;
    for (int it = 1; it <= max_it; it++) {
        const double rho1 = dot(r0.data(), r.data(), n, t);
        const double beta = (rho1 / rho) * (alpha / omega);
        rho = rho1;
        for (int64_t i = 0; i < n; i++) { p[i] = r[i] + beta * (p[i] - omega * v[i]); ph[i] = dinv[i] * p[i]; }
        spmv(A, ph.data(), v.data(), t);
        alpha = rho / dot(r0.data(), v.data(), n, t);
        for (int64_t i = 0; i < n; i++) { s[i] = r[i] - alpha * v[i]; sh[i] = dinv[i] * s[i]; }
        spmv(A, sh.data(), tt.data(), t);
        const double tts = dot(tt.data(), tt.data(), n, t);
        omega = tts > 0 ? dot(tt.data(), s.data(), n, t) / tts : 0.0;
        for (int64_t i = 0; i < n; i++) {
            x[i] += alpha * ph[i] + omega * sh[i];
            r[i] = s[i] - omega * tt[i];
        }
        st.iters = it;
        st.rel_res = std::sqrt(dot(r.data(), r.data(), n, t)) / bn;
        if (st.rel_res <= rtol) { st.converged = true; break; }
        if (omega == 0) break;
    }
    return st;
}

CSR transpose(const CSR& A) {
    CSR T;
    T.n = A.n;
    T.row_ptr.assign(A.n + 1, 0);
    for (int64_t k = 0; k < A.nnz(); k++) T.row_ptr[A.col[k] + 1]++;
    for (int64_t i = 0; i < A.n; i++) T.row_ptr[i + 1] += T.row_ptr[i];
    T.col.resize(A.nnz());
    T.val.resize(A.nnz());
    std::vector<int64_t> next(T.row_ptr.begin(), T.row_ptr.end() - 1);
    for (int64_t r = 0; r < A.n; r++)
        for (int64_t k = A.row_ptr[r]; k < A.row_ptr[r + 1]; k++) {
            int64_t p = next[A.col[k]]++;
            T.col[p] = int32_t(r);
            T.val[p] = A.val[k];
        }
    return T;
}

ElasticResult solve_elastic(const Problem& P, double rtol, int max_it, int n_threads, const CSR* pattern) {
    ElasticResult res;
    const int64_t n = P.dof.n_dofs;
    // the sparsity pattern is set up once per mesh (NonlinearSystem::init_data)
    CSR K = pattern ? *pattern : build_sparsity(P.mesh, P.dof);
    K.val.assign(K.col.size(), 0.0);
    std::vector<double> R(n), b(n);
    assemble(P, nullptr, R.data(), &K, n_threads);
    for (int64_t i = 0; i < n; i++) b[i] = -R[i];   // one Newton step from X = 0
    res.u.assign(n, 0.0);
    res.st = cg_jacobi(K, b.data(), res.u.data(), rtol, max_it, n_threads);
    // compliance = F . u (external loads; ComplianceOutput's -sum f_ext . u plus
    // the point loads, which ComplianceOutput does not see)
    double c = 0;
    for (int64_t i = 0; i < n; i++) c += b[i] * res.u[i];
    res.compliance = c;
    // self-adjoint: lambda = u, dC/dp = -u . dR/dp (+ 0 direct term)
    res.dc.assign(P.mesh.n_elems, 0.0);
    sensitivity_dots(P, res.u.data(), res.u.data(), res.dc.data(), n_threads);
    for (auto& g : res.dc) g = -g;
    return res;
}

HeatResult solve_heat(const Problem& P, double newton_rtol, double lin_rtol, int n_threads) {
    HeatResult res;
    const int64_t n = P.dof.n_dofs;
    CSR K = build_sparsity(P.mesh, P.dof);
    std::vector<double> R(n), b(n), dT(n);
    res.T.assign(n, 0.0);
    double r0 = -1;
    for (int it = 0; it < 50; it++) {
        assemble(P, res.T.data(), R.data(), &K, n_threads);
        double rn = 0;
        for (double v : R) rn += v * v;
        rn = std::sqrt(rn);
        if (r0 < 0) r0 = rn;
        res.res_hist.push_back(rn);
        if (rn <= newton_rtol * r0) break;
        for (int64_t i = 0; i < n; i++) b[i] = -R[i];
        bicgstab_jacobi(K, b.data(), dT.data(), lin_rtol, 20000, n_threads);
        for (int64_t i = 0; i < n; i++) res.T[i] += dT[i];
        res.newton_iters = it + 1;
    }
    // mean temperature J = (1/V) int T dV and dJ/dT = (1/V) int N (assembled, constrained)
    const Mesh& m = P.mesh;
    std::vector<double> dJdT(n, 0.0);
    double V = 0, JT = 0;
    std::vector<double> src(8);
    std::vector<int64_t> dofs(8);
    for (int64_t e = 0; e < m.n_elems; e++) {
        double X[8][3];
        for (int a = 0; a < 8; a++)
            for (int i = 0; i < 3; i++) X[a][i] = m.xyz[m.conn[e * 8 + a] * 3 + i];
        std::fill(src.begin(), src.end(), 0.0);
        hex8_source(X, -1.0, src.data());   // +int N
        P.dof.elem_dofs(m, e, dofs.data());
        for (int a = 0; a < 8; a++) {
            V += src[a];
            JT += src[a] * res.T[dofs[a]];
            if (!P.dof.constrained[dofs[a]]) dJdT[dofs[a]] += src[a];
        }
    }
    res.mean_T = JT / V;
    for (auto& v : dJdT) v /= V;
    // adjoint: J^T lambda = dJ/dT at the converged state
    assemble(P, res.T.data(), R.data(), &K, n_threads);
    CSR KT = transpose(K);
    std::vector<double> lam(n);
    bicgstab_jacobi(KT, dJdT.data(), lam.data(), lin_rtol, 20000, n_threads);
    // dJ/dq_e = -lambda . dR/dq_e
    res.dJ.assign(m.n_elems, 0.0);
    sensitivity_dots(P, res.T.data(), lam.data(), res.dJ.data(), n_threads);
    for (auto& g : res.dJ) g = -g;
    return res;
}

}  // namespace oracle
