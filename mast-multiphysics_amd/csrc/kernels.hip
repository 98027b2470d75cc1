// gfx950 kernels of the assembly / solve / sensitivity path.
//
// Conventions (shared with the host code in context.hip):
//  * node "ranks": nodes ordered by caller DOF base; active DOF (r, v) = r*nav + v
//  * element matrices are materialised per node pair:
//      Kel[((e*npe + a)*npe + b)*nav*nav + v*nav + w]
//  * the global matrix is a scalar CSR whose nav rows of one node are
//    contiguous, so the values of node row r form one chunk
//      [blk_ptr[r]*nav*nav, blk_ptr[r+1]*nav*nav)  ordered (v, k, w).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"

#define Structure_kMaxClasses 8

namespace mastg {

// 2-point Gauss abscissa; the 2x2x2 rule has unit weights (libMesh QGauss,
// tensor order: first direction fastest)
#define GP 0.57735026918962576451

__device__ __forceinline__ double gp_coord(int bit) { return bit ? GP : -GP; }

// d/dxi of the hex8 shape function of node n (libMesh Hex8 node order)
__device__ __forceinline__ void hex8_dshape(int n, double xi, double eta, double zeta,
                                            double& dxi, double& deta, double& dzeta,
                                            double& N) {
    const double sx = (n == 1 || n == 2 || n == 5 || n == 6) ? 1.0 : -1.0;
    const double sy = (n == 2 || n == 3 || n == 6 || n == 7) ? 1.0 : -1.0;
    const double sz = (n >= 4) ? 1.0 : -1.0;
    const double fx = 0.5 * (1.0 + sx * xi), fy = 0.5 * (1.0 + sy * eta),
                 fz = 0.5 * (1.0 + sz * zeta);
    N = fx * fy * fz;
    dxi = 0.5 * sx * fy * fz;
    deta = fx * 0.5 * sy * fz;
    dzeta = fx * fy * 0.5 * sz;
}

// 3x3 map inverse in libMesh FEMap cofactor form; returns det, fills rows of
// the inverse: ixi = (dxi/dx, dxi/dy, dxi/dz), ieta, izeta
__device__ __forceinline__ double inv3(const double J[9], double ixi[3], double ieta[3],
                                       double izeta[3]) {
    // J[i*3+r] = d x_i / d xi_r
    const double dxdxi = J[0], dxdeta = J[1], dxdzeta = J[2];
    const double dydxi = J[3], dydeta = J[4], dydzeta = J[5];
    const double dzdxi = J[6], dzdeta = J[7], dzdzeta = J[8];
    const double jac = dxdxi * (dydeta * dzdzeta - dzdeta * dydzeta) +
                       dydxi * (dzdeta * dxdzeta - dxdeta * dzdzeta) +
                       dzdxi * (dxdeta * dydzeta - dydeta * dxdzeta);
    const double ij = 1.0 / jac;
    ixi[0] = (dydeta * dzdzeta - dzdeta * dydzeta) * ij;
    ixi[1] = (dzdeta * dxdzeta - dxdeta * dzdzeta) * ij;
    ixi[2] = (dxdeta * dydzeta - dydeta * dxdzeta) * ij;
    ieta[0] = (dzdxi * dydzeta - dydxi * dzdzeta) * ij;
    ieta[1] = (dxdxi * dzdzeta - dzdxi * dxdzeta) * ij;
    ieta[2] = (dydxi * dxdzeta - dxdxi * dydzeta) * ij;
    izeta[0] = (dydxi * dzdeta - dzdxi * dydeta) * ij;
    izeta[1] = (dzdxi * dxdeta - dxdxi * dzdeta) * ij;
    izeta[2] = (dxdxi * dydeta - dydxi * dxdeta) * ij;
    return jac;
}

// ---------------------------------------------------------------------------
// hex8 linear elasticity: K_e (and optionally the internal residual K_e u_e)
// StructuralElement3D::internal_residual, linear strain
// (src/elasticity/solid_element_3d.cpp:269-411): K = sum_qp JxW B^T D B.
// One wave per element, lane = (a, b) node pair: the lane builds the 3x3 block
// K_ab = sum_q c_q B_a^T (D B_b) from the shape gradients staged in LDS.
// libMesh constraint application is folded in: constrained rows / columns
// zeroed, unit diagonal (DofMap::constrain_element_matrix_and_vector).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_hex8_elastic_K(Hex8Args A) {
    __shared__ double sg[4][8][8][3];   // [wave][q][node][d]
    __shared__ double sc[4][8];         // [wave][q] JxW
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t e = int64_t(blockIdx.x) * 4 + wave;
    const bool valid = e < A.n_elems;
    const int64_t ee = valid ? e : 0;
    // ---- geometry: lane = (q, n)
    {
        const int q = lane >> 3, n = lane & 7;
        const int32_t rn = A.econn[ee * 8 + n];
        const double x = A.xyz[int64_t(rn) * 3 + 0], y = A.xyz[int64_t(rn) * 3 + 1],
                     z = A.xyz[int64_t(rn) * 3 + 2];
        double dxi, deta, dzeta, N;
        hex8_dshape(n, gp_coord(q & 1), gp_coord((q >> 1) & 1), gp_coord(q >> 2), dxi, deta,
                    dzeta, N);
        double J[9] = {x * dxi, x * deta, x * dzeta, y * dxi, y * deta, y * dzeta,
                       z * dxi, z * deta, z * dzeta};
#pragma unroll
        for (int m = 1; m < 8; m <<= 1)
#pragma unroll
            for (int i = 0; i < 9; i++) J[i] += __shfl_xor(J[i], m, 64);
        double ixi[3], ieta[3], izeta[3];
        const double jac = inv3(J, ixi, ieta, izeta);
        sg[wave][q][n][0] = dxi * ixi[0] + deta * ieta[0] + dzeta * izeta[0];
        sg[wave][q][n][1] = dxi * ixi[1] + deta * ieta[1] + dzeta * izeta[1];
        sg[wave][q][n][2] = dxi * ixi[2] + deta * ieta[2] + dzeta * izeta[2];
        if (n == 0) sc[wave][q] = jac * 1.0;
    }
    __syncthreads();
    const int a = lane >> 3, b = lane & 7;
#define D(i) A.Dhat[i]   // 6x6 material matrix at E = 1 (kernel argument)
    double K[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll 2
    for (int q = 0; q < 8; q++) {
        const double c = sc[wave][q];
        const double gax = sg[wave][q][a][0], gay = sg[wave][q][a][1], gaz = sg[wave][q][a][2];
        const double gbx = sg[wave][q][b][0], gby = sg[wave][q][b][1], gbz = sg[wave][q][b][2];
        double DB[6][3];
#pragma unroll
        for (int r = 0; r < 6; r++) {
            DB[r][0] = D(r * 6 + 0) * gbx + D(r * 6 + 3) * gby + D(r * 6 + 5) * gbz;
            DB[r][1] = D(r * 6 + 1) * gby + D(r * 6 + 3) * gbx + D(r * 6 + 4) * gbz;
            DB[r][2] = D(r * 6 + 2) * gbz + D(r * 6 + 4) * gby + D(r * 6 + 5) * gbx;
        }
#pragma unroll
        for (int w = 0; w < 3; w++) {
            K[0 * 3 + w] += c * (gax * DB[0][w] + gay * DB[3][w] + gaz * DB[5][w]);
            K[1 * 3 + w] += c * (gay * DB[1][w] + gax * DB[3][w] + gaz * DB[4][w]);
            K[2 * 3 + w] += c * (gaz * DB[2][w] + gay * DB[4][w] + gax * DB[5][w]);
        }
    }
#undef D
    if (!valid) return;
    const double Ee = A.E0 * pow(A.param[e], A.penalty);
#pragma unroll
    for (int i = 0; i < 9; i++) K[i] *= Ee;
    const int32_t ra = A.econn[e * 8 + a], rb = A.econn[e * 8 + b];
    // internal residual f_a = sum_b K_ab u_b (before constraints), constrained rows 0
    if (A.X) {
        double f[3];
        const double ub0 = A.X[int64_t(rb) * 3 + 0], ub1 = A.X[int64_t(rb) * 3 + 1],
                     ub2 = A.X[int64_t(rb) * 3 + 2];
#pragma unroll
        for (int v = 0; v < 3; v++) f[v] = K[v * 3 + 0] * ub0 + K[v * 3 + 1] * ub1 + K[v * 3 + 2] * ub2;
#pragma unroll
        for (int m = 1; m < 8; m <<= 1)
#pragma unroll
            for (int v = 0; v < 3; v++) f[v] += __shfl_xor(f[v], m, 64);
        if (b == 0)
#pragma unroll
            for (int v = 0; v < 3; v++)
                A.Rel[(e * 8 + a) * 3 + v] = (A.con && A.con[int64_t(ra) * 3 + v]) ? 0.0 : f[v];
    }
    if (A.Kel) {
        uint8_t ca[3] = {0, 0, 0}, cb[3] = {0, 0, 0};
        if (A.con)
#pragma unroll
            for (int v = 0; v < 3; v++) { ca[v] = A.con[int64_t(ra) * 3 + v]; cb[v] = A.con[int64_t(rb) * 3 + v]; }
        double* out = A.Kel + (e * 64 + lane) * 9;
#pragma unroll
        for (int v = 0; v < 3; v++)
#pragma unroll
            for (int w = 0; w < 3; w++) {
                double kv = K[v * 3 + w];
                if (ca[v] || cb[w]) kv = (a == b && v == w) ? 1.0 : 0.0;
                out[v * 3 + w] = kv;
            }
    }
}

// ---------------------------------------------------------------------------
// hex8 sensitivity dot: g_e = lambda_e . dR_e/drho_e with dR_e/drho_e =
// sum_qp JxW B^T D' B u_e, D' = dE/drho * D(E=1)
// (StructuralElement3D::internal_residual_sensitivity, solid_element_3d.cpp:566-788;
// StiffnessMatrix3D::derivative with dnu = 0, isotropic_material_property_card.cpp:533-567)
// constrained entries of dR/drho are zero (constrain_element_vector), so
// lambda is masked.  One thread per element.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(128) void k_hex8_elastic_sens(Hex8Args A, const double* __restrict__ lam,
                                                           double* __restrict__ g) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= A.n_elems) return;
    double x[8][3], u[8][3], l[8][3];
#pragma unroll
    for (int n = 0; n < 8; n++) {
        const int64_t r = A.econn[e * 8 + n];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            x[n][i] = A.xyz[r * 3 + i];
            u[n][i] = A.X[r * 3 + i];
            l[n][i] = A.con[r * 3 + i] ? 0.0 : lam[r * 3 + i];
        }
    }
    double acc = 0.0;
    for (int q = 0; q < 8; q++) {
        const double xi = gp_coord(q & 1), eta = gp_coord((q >> 1) & 1), zeta = gp_coord(q >> 2);
        double dN[8][3], J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int n = 0; n < 8; n++) {
            double N;
            hex8_dshape(n, xi, eta, zeta, dN[n][0], dN[n][1], dN[n][2], N);
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int r = 0; r < 3; r++) J[i * 3 + r] += x[n][i] * dN[n][r];
        }
        double ixi[3], ieta[3], izeta[3];
        const double jac = inv3(J, ixi, ieta, izeta);
        // strains (exx, eyy, ezz, gxy, gyz, gzx) of u and lambda
        double eu[6] = {0, 0, 0, 0, 0, 0}, el[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int n = 0; n < 8; n++) {
            const double gx = dN[n][0] * ixi[0] + dN[n][1] * ieta[0] + dN[n][2] * izeta[0];
            const double gy = dN[n][0] * ixi[1] + dN[n][1] * ieta[1] + dN[n][2] * izeta[1];
            const double gz = dN[n][0] * ixi[2] + dN[n][1] * ieta[2] + dN[n][2] * izeta[2];
            eu[0] += gx * u[n][0]; eu[1] += gy * u[n][1]; eu[2] += gz * u[n][2];
            eu[3] += gy * u[n][0] + gx * u[n][1];
            eu[4] += gz * u[n][1] + gy * u[n][2];
            eu[5] += gz * u[n][0] + gx * u[n][2];
            el[0] += gx * l[n][0]; el[1] += gy * l[n][1]; el[2] += gz * l[n][2];
            el[3] += gy * l[n][0] + gx * l[n][1];
            el[4] += gz * l[n][1] + gy * l[n][2];
            el[5] += gz * l[n][0] + gx * l[n][2];
        }
        double s = 0.0;
#pragma unroll
        for (int r = 0; r < 6; r++) {
            double sig = 0.0;
#pragma unroll
            for (int c = 0; c < 6; c++) sig += A.Dhat[r * 6 + c] * eu[c];
            s += el[r] * sig;
        }
        acc += jac * s;
    }
    const double rho = A.param[e];
    const double dE = A.penalty * pow(rho, A.penalty - 1.0) * A.E0;
    g[e] = dE * acc;
}

// ---------------------------------------------------------------------------
// Segmented-reduction scatter of materialised element blocks into the scalar
// CSR values (NonlinearImplicitAssembly add_matrix, nonlinear_implicit_assembly.cpp:182-183):
// one wave per node row; thread t of the node chunk owns value (v, k, w) and
// sums the element blocks of block k in a fixed order (no atomics, the result
// is bitwise reproducible).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_scatter_blocks(ScatterArgs S) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t r = int64_t(blockIdx.x) * 4 + wave;
    if (r >= S.n_nodes) return;
    const int nav = S.nav, nav2 = nav * nav;
    const int64_t k0 = S.blk_ptr[r];
    const int64_t deg = S.blk_ptr[r + 1] - k0;
    const int64_t cnt = deg * nav2;
    const int64_t base = k0 * nav2;
    for (int64_t t = lane; t < cnt; t += 64) {
        const int64_t v = t / (deg * nav);
        const int64_t rem = t - v * deg * nav;
        const int64_t k = rem / nav;
        const int64_t w = rem - k * nav;
        const int64_t blk = k0 + k;
        double s = 0.0;
        for (int32_t c = S.cptr[blk]; c < S.cptr[blk + 1]; c++)
            s += S.Kel[int64_t(S.cidx[c]) * nav2 + v * nav + w];
        S.val[base + t] = s;
    }
}

// R_active = sum of element residual rows - F (F on unconstrained DOFs only)
__global__ void k_node_residual(int64_t n_nodes, int nav, int npe, const int32_t* __restrict__ ninc_ptr,
                                const int32_t* __restrict__ ninc, const double* __restrict__ Rel,
                                const double* __restrict__ F, const uint8_t* __restrict__ con,
                                double* __restrict__ R) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n_nodes * nav) return;
    const int64_t r = i / nav;
    const int v = int(i - r * nav);
    double s = 0.0;
    if (Rel)
        for (int32_t k = ninc_ptr[r]; k < ninc_ptr[r + 1]; k++) s += Rel[int64_t(ninc[k]) * nav + v];
    if (F && !con[i]) s -= F[i];
    R[i] = s;
}

// ------------------------------------------------------------- vectors ----
__global__ void k_gather(int64_t n, const int64_t* __restrict__ idx, const double* __restrict__ src,
                         double* __restrict__ dst) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}
__global__ void k_scatter_vec(int64_t n, const int64_t* __restrict__ idx, const double* __restrict__ src,
                              double* __restrict__ dst) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[idx[i]] = src[i];
}
__global__ void k_fill(int64_t n, double v, double* __restrict__ x) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}
__global__ void k_axpby(int64_t n, double a, const double* __restrict__ x, double b,
                        double* __restrict__ y) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i] + b * y[i];
}
__global__ void k_jacobi_inv(int64_t n, const int64_t* __restrict__ diag_pos, const double* __restrict__ val,
                             double* __restrict__ dinv) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        const double d = val[diag_pos[i]];
        dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
    }
}

// block-level FP64 sum into partials[blockIdx.x] (blockDim 256)
__device__ __forceinline__ void block_sum_store(double v, double* partials, int slot) {
    __shared__ double red[4];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) partials[slot] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(256) void k_dot(int64_t n, const double* __restrict__ a,
                                             const double* __restrict__ b, double* __restrict__ partials) {
    double s = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        s += a[i] * b[i];
    block_sum_store(s, partials, blockIdx.x);
}

// deterministic final reduction of `n` partials (one block of 256)
__device__ __forceinline__ double reduce_partials(const double* __restrict__ p, int n) {
    __shared__ double red[4];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += p[i];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    const double t = (red[0] + red[1]) + (red[2] + red[3]);
    __syncthreads();
    return t;
}

__global__ __launch_bounds__(256) void k_reduce_to(const double* __restrict__ p, int n,
                                                   double* __restrict__ out) {
    const double s = reduce_partials(p, n);
    if (threadIdx.x == 0) *out = s;
}


// ---------------------------------------------------------------------------
// Matrix-free hex8 operator y = K x over congruence classes:
//   K = sum_e s_e P_e^T Khat_{class(e)} P_e   (s_e = E0 rho_e^p)
// with libMesh's constraint convention (constrained rows / columns removed,
// diagonal = number of elements at the node), i.e. the same matrix the
// segmented scatter assembles.  Two deterministic phases:
//   1. one thread per element gathers its 8 nodes once and writes the element
//      forces f_e = s_e Khat u_e to Fel (structure-of-arrays, coalesced); Khat
//      sits in LDS and, with one class, every lane reads the same entry (LDS
//      broadcast);
//   2. one thread per node sums the forces of its incident elements in a fixed
//      order, applies the constrained rows and the fused x.y partial sum.
// The input must be zero on constrained DOFs (CG directions are; the
// Jacobian-product entry point masks its input first).
// ---------------------------------------------------------------------------
template <bool ONE_CLASS>
__global__ __launch_bounds__(256) void k_hex8_mf_elem(CGArgs C, const double* __restrict__ x) {
    if (C.S && C.S->done) return;
    __shared__ double Kh[Structure_kMaxClasses * 576];
    const MFArgs& A = C.mf;
    const int ncl = ONE_CLASS ? 1 : A.n_classes;
    for (int i = threadIdx.x; i < ncl * 576; i += 256) Kh[i] = A.Khat[i];
    __syncthreads();
    const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (e >= A.n_elems) return;
    const int4* ec = reinterpret_cast<const int4*>(A.econn + e * 8);
    const int4 c0 = ec[0], c1 = ec[1];
    const int32_t rb[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
    double u[24];
#pragma unroll
    for (int lb = 0; lb < 8; lb++) {
        u[lb * 3 + 0] = x[int64_t(rb[lb]) * 3 + 0];
        u[lb * 3 + 1] = x[int64_t(rb[lb]) * 3 + 1];
        u[lb * 3 + 2] = x[int64_t(rb[lb]) * 3 + 2];
    }
    const double s = A.sE[e];
    const double* K = Kh + (ONE_CLASS ? 0 : int(A.eclass[e]) * 576);
    const int64_t ne = A.n_elems;
#pragma unroll
    for (int la = 0; la < 8; la++) {
        double f0 = 0.0, f1 = 0.0, f2 = 0.0;
#pragma unroll
        for (int lb = 0; lb < 8; lb++) {
            const double* k = K + (la * 8 + lb) * 9;
            f0 += k[0] * u[lb * 3 + 0] + k[1] * u[lb * 3 + 1] + k[2] * u[lb * 3 + 2];
            f1 += k[3] * u[lb * 3 + 0] + k[4] * u[lb * 3 + 1] + k[5] * u[lb * 3 + 2];
            f2 += k[6] * u[lb * 3 + 0] + k[7] * u[lb * 3 + 1] + k[8] * u[lb * 3 + 2];
        }
        A.Fel[(la * 3 + 0) * ne + e] = s * f0;
        A.Fel[(la * 3 + 1) * ne + e] = s * f1;
        A.Fel[(la * 3 + 2) * ne + e] = s * f2;
    }
}

__global__ __launch_bounds__(256) void k_hex8_mf_node(CGArgs C, const double* __restrict__ x,
                                                      double* __restrict__ y, int want_dot) {
    if (C.S && C.S->done) return;
    const MFArgs& A = C.mf;
    const int64_t r = int64_t(blockIdx.x) * 256 + threadIdx.x;
    double d = 0.0;
    if (r < A.n_nodes) {
        const int32_t k0 = A.ninc_ptr[r], k1 = A.ninc_ptr[r + 1];
        const int64_t ne = A.n_elems;
        double y0 = 0.0, y1 = 0.0, y2 = 0.0;
        for (int32_t k = k0; k < k1; k++) {
            const int32_t c = A.ninc[k];
            const int64_t e = c >> 3;
            const int la = c & 7;
            y0 += A.Fel[(la * 3 + 0) * ne + e];
            y1 += A.Fel[(la * 3 + 1) * ne + e];
            y2 += A.Fel[(la * 3 + 2) * ne + e];
        }
        const int m = A.nodecon[r];
        const double cnt = double(k1 - k0);
        const double x0 = x[r * 3 + 0], x1 = x[r * 3 + 1], x2 = x[r * 3 + 2];
        if (m & 1) y0 = cnt * x0;
        if (m & 2) y1 = cnt * x1;
        if (m & 4) y2 = cnt * x2;
        y[r * 3 + 0] = y0;
        y[r * 3 + 1] = y1;
        y[r * 3 + 2] = y2;
        d = y0 * x0 + y1 * x1 + y2 * x2;
    }
    if (want_dot) block_sum_store(d, C.part0, blockIdx.x);
}

__global__ void k_mask_constrained(int64_t n, const uint8_t* __restrict__ con, double* __restrict__ x) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n && con[i]) x[i] = 0.0;
}

__global__ void k_elem_scale(int64_t n, const double* __restrict__ rho, double E0, double p,
                             double* __restrict__ sE) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < n) sE[e] = E0 * pow(rho[e], p);
}

// ------------------------------------------------------- CG (FP64 CSR) ----
// y = A x over the scalar CSR with SW lanes per row (SW in 4..64, chosen by the
// autotuner), grid-stride over rows, and partial x.y per block.
template <int SW>
__global__ __launch_bounds__(256) void k_spmv_csr_dot(CGArgs C, const double* __restrict__ x,
                                                      double* __restrict__ y, int want_dot,
                                                      int in_cg_loop) {
    if (in_cg_loop && C.S->done) return;
    const int sub = threadIdx.x & (SW - 1);
    double dsum = 0.0;
    const int64_t rows_per_grid = (int64_t(gridDim.x) * 256) / SW;
    for (int64_t row = (int64_t(blockIdx.x) * 256 + threadIdx.x) / SW; row < C.n; row += rows_per_grid) {
        const int64_t k0 = C.row_ptr[row], k1 = C.row_ptr[row + 1];
        double s = 0.0;
#pragma unroll 4
        for (int64_t k = k0 + sub; k < k1; k += SW) s += C.val[k] * x[C.col[k]];
#pragma unroll
        for (int m = SW / 2; m >= 1; m >>= 1) s += __shfl_xor(s, m, SW);
        if (sub == 0) {
            y[row] = s;
            dsum += s * x[row];
        }
    }
    if (want_dot) block_sum_store(dsum, C.part0, blockIdx.x);
}

// alpha = rz / pq
__global__ __launch_bounds__(256) void k_cg_alpha(CGArgs C) {
    if (C.S->done) return;
    const double pq = reduce_partials(C.part0, C.nparts_op);
    if (threadIdx.x == 0) {
        C.S->pq = pq;
        C.S->alpha = C.S->rz / pq;
        if (!(pq > 0.0) || !isfinite(pq)) C.S->done = 3;   // breakdown / NaN
    }
}

// x += a p ; r -= a q ; z = dinv r ; partials (r.z, r.r)
__global__ __launch_bounds__(256) void k_cg_update(CGArgs C) {
    if (C.S->done) return;
    const double a = C.S->alpha;
    double rz = 0.0, rr = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < C.n; i += int64_t(gridDim.x) * 256) {
        C.x[i] += a * C.p[i];
        const double r = C.r[i] - a * C.q[i];
        C.r[i] = r;
        const double z = C.dinv[i] * r;
        C.z[i] = z;
        rz += r * z;
        rr += r * r;
    }
    block_sum_store(rz, C.part0, blockIdx.x);
    __syncthreads();
    block_sum_store(rr, C.part1, blockIdx.x);
}

// beta = rz_new / rz ; convergence on ||r||^2 <= tol2 * ||b||^2
__global__ __launch_bounds__(256) void k_cg_beta(CGArgs C) {
    if (C.S->done) return;
    const double rz = reduce_partials(C.part0, C.nparts);
    const double rr = reduce_partials(C.part1, C.nparts);
    if (threadIdx.x == 0) {
        CGScalars* S = C.S;
        S->beta = rz / S->rz;
        S->rz = rz;
        S->rr = rr;
        S->iter += 1;
        if (!isfinite(rr)) S->done = 4;
        else if (rr <= S->tol2 * S->bb) S->done = 1;
        else if (S->iter >= S->max_it) S->done = 2;
    }
}

// p = z + beta p
__global__ __launch_bounds__(256) void k_cg_p(CGArgs C) {
    if (C.S->done) return;
    const double b = C.S->beta;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < C.n; i += int64_t(gridDim.x) * 256)
        C.p[i] = C.z[i] + b * C.p[i];
}

// start: r = b - A x0 is given in r; z = dinv r; p = z; partials (r.z, b.b)
__global__ __launch_bounds__(256) void k_cg_start(CGArgs C, const double* __restrict__ b) {
    double rz = 0.0, bb = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < C.n; i += int64_t(gridDim.x) * 256) {
        const double r = C.r[i];
        const double z = C.dinv[i] * r;
        C.z[i] = z;
        C.p[i] = z;
        rz += r * z;
        bb += b[i] * b[i];
    }
    block_sum_store(rz, C.part0, blockIdx.x);
    __syncthreads();
    block_sum_store(bb, C.part1, blockIdx.x);
}

__global__ __launch_bounds__(256) void k_cg_start_scalars(CGArgs C, double tol2, int max_it) {
    const double rz = reduce_partials(C.part0, C.nparts);
    const double bb = reduce_partials(C.part1, C.nparts);
    if (threadIdx.x == 0) {
        CGScalars* S = C.S;
        S->rz = rz;
        S->bb = bb;
        S->rr = bb;
        S->iter = 0;
        S->tol2 = tol2;
        S->max_it = max_it;
        S->done = (bb == 0.0) ? 1 : (isfinite(bb) ? 0 : 4);
    }
}

// ---------------------------------------------------------- launchers ----
static inline unsigned blocks_for(int64_t n, int per) { return unsigned((n + per - 1) / per); }

void launch_hex8_K(const Hex8Args& a, hipStream_t s) {
    hipLaunchKernelGGL(k_hex8_elastic_K, dim3(blocks_for(a.n_elems, 4)), dim3(256), 0, s, a);
}
void launch_hex8_sens(const Hex8Args& a, const double* lam, double* g, hipStream_t s) {
    hipLaunchKernelGGL(k_hex8_elastic_sens, dim3(blocks_for(a.n_elems, 128)), dim3(128), 0, s, a, lam, g);
}
void launch_scatter(const ScatterArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_scatter_blocks, dim3(blocks_for(a.n_nodes, 4)), dim3(256), 0, s, a);
}
void launch_node_residual(int64_t n_nodes, int nav, int npe, const int32_t* ninc_ptr,
                          const int32_t* ninc, const double* Rel, const double* F,
                          const uint8_t* con, double* R, hipStream_t s) {
    hipLaunchKernelGGL(k_node_residual, dim3(blocks_for(n_nodes * nav, 256)), dim3(256), 0, s,
                       n_nodes, nav, npe, ninc_ptr, ninc, Rel, F, con, R);
}
void launch_gather(int64_t n, const int64_t* idx, const double* src, double* dst, hipStream_t s) {
    hipLaunchKernelGGL(k_gather, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, idx, src, dst);
}
void launch_scatter_vec(int64_t n, const int64_t* idx, const double* src, double* dst, hipStream_t s) {
    hipLaunchKernelGGL(k_scatter_vec, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, idx, src, dst);
}
void launch_fill(int64_t n, double v, double* x, hipStream_t s) {
    hipLaunchKernelGGL(k_fill, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, v, x);
}
void launch_axpby(int64_t n, double a, const double* x, double b, double* y, hipStream_t s) {
    hipLaunchKernelGGL(k_axpby, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, a, x, b, y);
}
void launch_jacobi_inv(int64_t n, const int64_t* diag_pos, const double* val, double* dinv,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_jacobi_inv, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, diag_pos, val, dinv);
}
void launch_dot(int64_t n, const double* a, const double* b, double* partials, int nparts,
                double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_dot, dim3(nparts), dim3(256), 0, s, n, a, b, partials);
    hipLaunchKernelGGL(k_reduce_to, dim3(1), dim3(256), 0, s, partials, nparts, out);
}
static void spmv_dispatch(const CGArgs& c, const double* x, double* y, int want_dot, int in_loop,
                          hipStream_t s) {
    switch (c.spmv_sw) {
        case 4: hipLaunchKernelGGL(k_spmv_csr_dot<4>, dim3(c.nparts), dim3(256), 0, s, c, x, y, want_dot, in_loop); break;
        case 8: hipLaunchKernelGGL(k_spmv_csr_dot<8>, dim3(c.nparts), dim3(256), 0, s, c, x, y, want_dot, in_loop); break;
        case 32: hipLaunchKernelGGL(k_spmv_csr_dot<32>, dim3(c.nparts), dim3(256), 0, s, c, x, y, want_dot, in_loop); break;
        case 64: hipLaunchKernelGGL(k_spmv_csr_dot<64>, dim3(c.nparts), dim3(256), 0, s, c, x, y, want_dot, in_loop); break;
        default: hipLaunchKernelGGL(k_spmv_csr_dot<16>, dim3(c.nparts), dim3(256), 0, s, c, x, y, want_dot, in_loop); break;
    }
}
void launch_spmv_csr(const CGArgs& c, const double* x, double* y, int want_dot, hipStream_t s) {
    spmv_dispatch(c, x, y, want_dot, 0, s);
}
void launch_cg_start(const CGArgs& c, const double* b, double tol2, int max_it, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_start, dim3(c.nparts), dim3(256), 0, s, c, b);
    hipLaunchKernelGGL(k_cg_start_scalars, dim3(1), dim3(256), 0, s, c, tol2, max_it);
}
void launch_elem_scale(int64_t n, const double* rho, double E0, double penalty, double* sE,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_elem_scale, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, rho, E0, penalty, sE);
}
int mf_node_blocks(int64_t n_nodes) { return int((n_nodes + 255) / 256); }
static void mf_dispatch(const CGArgs& c, const double* x_elem, const double* x_node, double* y,
                        int want_dot, hipStream_t s) {
    const unsigned ge = blocks_for(c.mf.n_elems, 256);
    if (c.mf.n_classes == 1)
        hipLaunchKernelGGL(k_hex8_mf_elem<true>, dim3(ge), dim3(256), 0, s, c, x_elem);
    else
        hipLaunchKernelGGL(k_hex8_mf_elem<false>, dim3(ge), dim3(256), 0, s, c, x_elem);
    hipLaunchKernelGGL(k_hex8_mf_node, dim3(mf_node_blocks(c.mf.n_nodes)), dim3(256), 0, s, c, x_node, y,
                       want_dot);
}
void launch_mf_apply(const CGArgs& c, const double* x_elem, const double* x_node, double* y,
                     int want_dot, hipStream_t s) {
    CGArgs d = c;
    d.S = nullptr;   // standalone application: ignore the CG termination flag
    mf_dispatch(d, x_elem, x_node, y, want_dot, s);
}
void launch_mask_constrained(int64_t n, const uint8_t* con, double* x, hipStream_t s) {
    hipLaunchKernelGGL(k_mask_constrained, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, con, x);
}
void launch_cg_iteration(const CGArgs& c, hipStream_t s) {
    if (c.use_mf)
        mf_dispatch(c, c.p, c.p, c.q, 1, s);
    else
        spmv_dispatch(c, c.p, c.q, 1, 1, s);
    hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(256), 0, s, c);
    hipLaunchKernelGGL(k_cg_update, dim3(c.nparts), dim3(256), 0, s, c);
    hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(256), 0, s, c);
    hipLaunchKernelGGL(k_cg_p, dim3(c.nparts), dim3(256), 0, s, c);
}

}  // namespace mastg
