// Stress outputs of hex8 elasticity: aggregates of the von Mises stress over
// the quadrature points, their derivative with respect to the state and their
// direct derivative with respect to the element densities.
//
//   stress at each quadrature point: sigma = E_e D(nu) B u_e, E_e = E0 rho_e^p
//     (StructuralElement3D::calculate_stress, src/elasticity/solid_element_3d.cpp:1147-1236)
//   von Mises: P = ((s0-s1)^2 + (s1-s2)^2 + (s2-s0)^2) / 2 + 3 (s3^2 + s4^2 + s5^2),
//     svm = P^(1/2) (StressStrainOutputBase::Data::von_Mises_stress,
//     src/elasticity/stress_output_base.cpp:183-230; derivative 0 where P = 0)
// Each aggregate is a function of two sums over the points, I = sum f_I(svm)
// JxW and V = sum f_V(svm) JxW, with s0 the reference stress:
//   KS (KSStressStrainOutput, src/elasticity/ks_stress_output.cpp:53-243):
//     f_I = exp(k (svm - s0) / s0), f_V = 1, value (1/k) ln(I / V)
//   p-norm (StressStrainOutputBase::functional_for_all_elems and its
//   derivatives, stress_output_base.cpp:852-1176): weight
//     w = exp(min(rho ((svm - s0) / s0)^pw, lim)) (lim = 100, the clamp of
//     _exp_arg_lim), f_I = (svm / s0)^p w, f_V = w, value s0 (I / V)^(1/p)
//   smooth ramp (SmoothRampStressStrainOutput, src/elasticity/
//   smooth_ramp_stress_output.cpp:36-241): f_I = (1 + (svm / s0)^p)^(1/p),
//     f_V = 1, value I - V
// The derivatives are sums over the points of c_q dsvm_q/d(.), c_q = the
// value's derivative in svm_q (the reference's num_sens / denom_sens terms);
// dsvm/drho_e|direct = (p / rho_e) svm.
//
// Two passes, a thread per element: the first sums I and V per work-group
// (fixed grid, fixed-order final reduction: deterministic), the second reads
// I and V from device memory and writes the element rows of the state
// derivative (constrained rows 0, assembled by the node-residual gather) and
// the direct derivative.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_util.h"
#include "kernels.h"

namespace mastg {

namespace {

constexpr int kKsBlocks = 1024;

__device__ __forceinline__ double ks_gp(int i) { return i ? 0.57735026918962576451 : -0.57735026918962576451; }

__device__ __forceinline__ void ks_dshape(int n, double xi, double eta, double zeta, double d[3]) {
    const double sx = (n == 1 || n == 2 || n == 5 || n == 6) ? 1.0 : -1.0;
    const double sy = (n == 2 || n == 3 || n == 6 || n == 7) ? 1.0 : -1.0;
    const double sz = (n >= 4) ? 1.0 : -1.0;
    const double fx = 0.5 * (1 + sx * xi), fy = 0.5 * (1 + sy * eta), fz = 0.5 * (1 + sz * zeta);
    d[0] = 0.5 * sx * fy * fz;
    d[1] = fx * 0.5 * sy * fz;
    d[2] = fx * fy * 0.5 * sz;
}

// shape gradients g[n][3] and JxW at quadrature point q (xi fastest)
__device__ __forceinline__ double ks_geometry(const double x[8][3], int q, double g[8][3]) {
    const double xi = ks_gp(q & 1), eta = ks_gp((q >> 1) & 1), zeta = ks_gp(q >> 2);
    double dN[8][3], J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int n = 0; n < 8; n++) {
        ks_dshape(n, xi, eta, zeta, dN[n]);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int r = 0; r < 3; r++) J[i * 3 + r] += x[n][i] * dN[n][r];
    }
    const double det = J[0] * (J[4] * J[8] - J[7] * J[5]) + J[3] * (J[7] * J[2] - J[1] * J[8]) +
                       J[6] * (J[1] * J[5] - J[4] * J[2]);
    const double id = 1.0 / det;
    // rows of J^-1 (d xi_r / d x_i): Ji[r][i]
    double Ji[3][3];
    Ji[0][0] = (J[4] * J[8] - J[5] * J[7]) * id;
    Ji[0][1] = (J[2] * J[7] - J[1] * J[8]) * id;
    Ji[0][2] = (J[1] * J[5] - J[2] * J[4]) * id;
    Ji[1][0] = (J[5] * J[6] - J[3] * J[8]) * id;
    Ji[1][1] = (J[0] * J[8] - J[2] * J[6]) * id;
    Ji[1][2] = (J[2] * J[3] - J[0] * J[5]) * id;
    Ji[2][0] = (J[3] * J[7] - J[4] * J[6]) * id;
    Ji[2][1] = (J[1] * J[6] - J[0] * J[7]) * id;
    Ji[2][2] = (J[0] * J[4] - J[1] * J[3]) * id;
#pragma unroll
    for (int n = 0; n < 8; n++)
#pragma unroll
        for (int i = 0; i < 3; i++) g[n][i] = dN[n][0] * Ji[0][i] + dN[n][1] * Ji[1][i] + dN[n][2] * Ji[2][i];
    return det;
}

// stress (xx yy zz xy yz zx) at one point and its von Mises value
__device__ __forceinline__ double ks_stress(const Hex8Args& A, const double g[8][3], const double u[8][3],
                                            double Ee, double s[6]) {
    double e[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int n = 0; n < 8; n++) {
        e[0] += g[n][0] * u[n][0];
        e[1] += g[n][1] * u[n][1];
        e[2] += g[n][2] * u[n][2];
        e[3] += g[n][1] * u[n][0] + g[n][0] * u[n][1];
        e[4] += g[n][2] * u[n][1] + g[n][1] * u[n][2];
        e[5] += g[n][2] * u[n][0] + g[n][0] * u[n][2];
    }
#pragma unroll
    for (int r = 0; r < 6; r++) {
        double v = 0.0;
#pragma unroll
        for (int c = 0; c < 6; c++) v += A.Dhat[r * 6 + c] * e[c];
        s[r] = Ee * v;
    }
    const double P = 0.5 * ((s[0] - s[1]) * (s[0] - s[1]) + (s[1] - s[2]) * (s[1] - s[2]) + (s[2] - s[0]) * (s[2] - s[0])) +
                     3.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
    return sqrt(P);
}

__device__ __forceinline__ void ks_load(const Hex8Args& A, int64_t e, double x[8][3], double u[8][3]) {
#pragma unroll
    for (int n = 0; n < 8; n++) {
        const int64_t r = A.econn[e * 8 + n];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            x[n][i] = A.xyz[r * 3 + i];
            u[n][i] = A.X[r * 3 + i];
        }
    }
}

// the point's contributions to I and V
__device__ __forceinline__ void agg_terms(const StressAgg& G, double svm, double jxw, double& dI, double& dV) {
    if (G.kind == kAggPNorm) {
        const double sp = pow((svm - G.s0) / G.s0, G.pw);
        const double w = G.rho * sp > G.lim ? exp(G.lim) : exp(G.rho * sp);
        dI = pow(svm / G.s0, G.p) * w * jxw;
        dV = w * jxw;
    } else if (G.kind == kAggRamp) {
        dI = pow(1.0 + pow(svm / G.s0, G.p), 1.0 / G.p) * jxw;
        dV = jxw;
    } else {
        dI = exp(G.p * (svm - G.s0) / G.s0) * jxw;
        dV = jxw;
    }
}

// c_q: the value's derivative in the point's von Mises stress
__device__ __forceinline__ double agg_coef(const StressAgg& G, double svm, double jxw, double I, double V) {
    if (G.kind == kAggPNorm) {
        const double x = svm / G.s0, y = (svm - G.s0) / G.s0;
        const double sps = pow(x, G.p), sp1s = pow(x, G.p - 1.0);
        const double spw = pow(y, G.pw), sp1w = pow(y, G.pw - 1.0);
        double num, den;
        if (G.rho * spw > G.lim) {
            num = G.p * sp1s * exp(G.lim);
            den = 0.0;
        } else {
            const double w = exp(G.rho * spw);
            num = w * (G.p * sp1s + sps * G.rho * G.pw * sp1w);
            den = w * G.rho * G.pw * sp1w;
        }
        return G.s0 / G.p * pow(I / V, 1.0 / G.p - 1.0) * (num / V - I / (V * V) * den) * jxw / G.s0;
    }
    if (G.kind == kAggRamp) {
        const double x = svm / G.s0;
        return pow(1.0 + pow(x, G.p), 1.0 / G.p - 1.0) * pow(x, G.p - 1.0) * jxw / G.s0;
    }
    return exp(G.p * (svm - G.s0) / G.s0) * jxw * (1.0 / I) / G.s0;
}

__global__ __launch_bounds__(256) void k_hex8_ks_sum(Hex8Args A, StressAgg G, double* __restrict__ part) {
    double I = 0.0, V = 0.0;
    for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < A.n_elems; e += int64_t(gridDim.x) * 256) {
        double x[8][3], u[8][3];
        ks_load(A, e, x, u);
        const double Ee = A.E0 * pow(A.param[e], A.penalty);
        for (int q = 0; q < 8; q++) {
            double g[8][3], s[6];
            const double jxw = ks_geometry(x, q, g);
            const double svm = ks_stress(A, g, u, Ee, s);
            double dI, dV;
            agg_terms(G, svm, jxw, dI, dV);
            I += dI;
            V += dV;
        }
    }
    block_sum_store(I, part, blockIdx.x);
    __syncthreads();
    block_sum_store(V, part + gridDim.x, blockIdx.x);
}

__global__ __launch_bounds__(256) void k_ks_reduce(const double* __restrict__ part, int nb, double* __restrict__ out) {
    const double I = reduce_partials(part, nb);
    __syncthreads();
    const double V = reduce_partials(part + nb, nb);
    if (threadIdx.x == 0) {
        out[0] = I;
        out[1] = V;
    }
}

__global__ __launch_bounds__(128) void k_hex8_ks_grad(Hex8Args A, StressAgg G, const double* __restrict__ IV,
                                                     double* __restrict__ dpar) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= A.n_elems) return;
    double x[8][3], u[8][3];
    ks_load(A, e, x, u);
    const double rho = A.param[e];
    const double Ee = A.E0 * pow(rho, A.penalty);
    const double I = IV[0], V = IV[1];
    double f[8][3];
#pragma unroll
    for (int n = 0; n < 8; n++) f[n][0] = f[n][1] = f[n][2] = 0.0;
    double direct = 0.0;
    for (int q = 0; q < 8; q++) {
        double g[8][3], s[6];
        const double jxw = ks_geometry(x, q, g);
        const double svm = ks_stress(A, g, u, Ee, s);
        const double w = agg_coef(G, svm, jxw, I, V);
        direct += w * svm;
        if (!(svm > 0.0)) continue;
        // dsvm/dsigma = dP/dsigma / (2 svm); d sigma / du = Ee D B (D symmetric)
        const double c = w * 0.5 / svm;
        const double gP[6] = {2 * s[0] - s[1] - s[2], 2 * s[1] - s[0] - s[2], 2 * s[2] - s[0] - s[1],
                              6 * s[3], 6 * s[4], 6 * s[5]};
        double t[6];
#pragma unroll
        for (int r = 0; r < 6; r++) {
            double v = 0.0;
#pragma unroll
            for (int cc = 0; cc < 6; cc++) v += A.Dhat[r * 6 + cc] * gP[cc];
            t[r] = c * Ee * v;
        }
#pragma unroll
        for (int n = 0; n < 8; n++) {
            f[n][0] += g[n][0] * t[0] + g[n][1] * t[3] + g[n][2] * t[5];
            f[n][1] += g[n][1] * t[1] + g[n][0] * t[3] + g[n][2] * t[4];
            f[n][2] += g[n][2] * t[2] + g[n][1] * t[4] + g[n][0] * t[5];
        }
    }
#pragma unroll
    for (int n = 0; n < 8; n++) {
        const int64_t r = A.econn[e * 8 + n];
#pragma unroll
        for (int v = 0; v < 3; v++) A.Rel[(e * 8 + n) * 3 + v] = (A.con && A.con[r * 3 + v]) ? 0.0 : f[n][v];
    }
    if (dpar) dpar[e] = A.penalty / rho * direct;
}

}  // namespace

void launch_hex8_ks_sum(const Hex8Args& a, const StressAgg& g, double* part, double* IV, hipStream_t s) {
    const int nb = int(std::min<int64_t>(kKsBlocks, std::max<int64_t>(1, (a.n_elems + 255) / 256)));
    hipLaunchKernelGGL(k_hex8_ks_sum, dim3(nb), dim3(256), 0, s, a, g, part);
    hipLaunchKernelGGL(k_ks_reduce, dim3(1), dim3(256), 0, s, part, nb, IV);
}

void launch_hex8_ks_grad(const Hex8Args& a, const StressAgg& g, const double* IV, double* dpar, hipStream_t s) {
    hipLaunchKernelGGL(k_hex8_ks_grad, dim3(std::max<int64_t>(1, (a.n_elems + 127) / 128)), dim3(128), 0, s, a, g,
                       IV, dpar);
}

}  // namespace mastg
