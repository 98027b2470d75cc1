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
#include "device_util.h"
#include "plate_frame.h"

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
// K_ab = sum_q c_q B_a^T (D B_b) from the shape gradients staged in LDS, in
// the Lame form of the isotropic D (D = lam 1 1^T + mu diag(2, 2, 2, 1, 1, 1),
// lam = D[0][1], mu = D[3][3]; so also for dD/dnu):
//   (K_ab)_vw = sum_q c_q [lam g_a,v g_b,w + mu g_a,w g_b,v + mu delta_vw g_a . g_b]
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
    const double lam = A.Dhat[1], mu = A.Dhat[21];   // at E = 1 (kernel argument)
    // P_vw = sum_q c_q g_a,v g_b,w and S = trace
    double P[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll 2
    for (int q = 0; q < 8; q++) {
        const double c = sc[wave][q];
        const double ga[3] = {c * sg[wave][q][a][0], c * sg[wave][q][a][1], c * sg[wave][q][a][2]};
        const double gb[3] = {sg[wave][q][b][0], sg[wave][q][b][1], sg[wave][q][b][2]};
#pragma unroll
        for (int v = 0; v < 3; v++)
#pragma unroll
            for (int w = 0; w < 3; w++) P[v * 3 + w] = fma(ga[v], gb[w], P[v * 3 + w]);
    }
    const double S = P[0] + P[4] + P[8];
    double K[9];
#pragma unroll
    for (int v = 0; v < 3; v++)
#pragma unroll
        for (int w = 0; w < 3; w++) K[v * 3 + w] = lam * P[v * 3 + w] + mu * P[w * 3 + v] + (v == w ? mu * S : 0.0);
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
// hex8 K_e in FP32 on MFMA (v_mfma_f32_16x16x4f32, FP32 accumulation).
// Lane layout of the 16x16x4 instruction: A[i][k] from lane i + 16k, B[k][j]
// from lane j + 16k, D[4*(l/16)+r][l%16] in accumulator r.  Geometry in FP64,
// operands rounded to FP32.  Output: FP32 node-pair blocks, constrained,
// summed into the FP64 CSR by the scatter.  Entries carry ~1e-7 relative
// rounding; the rounded matrix no longer maps rigid-body motions exactly to
// zero, which moves compliance by 1e-4..1e-3 on ill-conditioned SIMP designs,
// so this path serves Jacobian exports and preconditioners, not the FP64 solve.
// ---------------------------------------------------------------------------
typedef float mastg_floatx4 __attribute__((ext_vector_type(4)));

// FP32 element matrices on the matrix cores.  With g_qa the physical gradient
// of N_a at point q and c_q its JxW, P^{kl}[a][b] = sum_q c_q g_qa,k g_qb,l is
// the 24 x 24 product (W G)^T G over the 8 points (rows (k, a), columns (l,
// b)): four 16 x 16 tiles of two v_mfma_f32_16x16x4f32 each.  The isotropic
// B^T D B then follows per entry (the Lame form of D = lam 1 1^T + mu diag(2,
// 2, 2, 1, 1, 1) in Voigt notation):
//   K_(a v)(b w) = E_e [lam P^{vw}[a][b] + mu P^{wv}[a][b] + mu delta_vw sum_k P^{kk}[a][b]]
// One wave per element; lanes (q, n) form the geometry in FP64 and round it.
__global__ __launch_bounds__(256) void k_hex8_elastic_K_f32(Hex8Args A, float* __restrict__ Kel32) {
    __shared__ float sg[4][8][8][3];   // [wave][q][node][d]
    __shared__ float sc[4][8];
    __shared__ float sP[4][24][25];    // [wave][(k, a)][(l, b)] (one float of padding per row)
    __shared__ uint8_t scon[4][8];     // constraint bits of the element's nodes
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t e = int64_t(blockIdx.x) * 4 + wave;
    const bool valid = e < A.n_elems;
    const int64_t ee = valid ? e : 0;
    {
        const int q = lane >> 3, n = lane & 7;
        const int32_t rn = A.econn[ee * 8 + n];
        const double x = A.xyz[int64_t(rn) * 3 + 0], y = A.xyz[int64_t(rn) * 3 + 1],
                     z = A.xyz[int64_t(rn) * 3 + 2];
        double dxi, deta, dzeta, N;
        hex8_dshape(n, gp_coord(q & 1), gp_coord((q >> 1) & 1), gp_coord(q >> 2), dxi, deta, dzeta, N);
        double J[9] = {x * dxi, x * deta, x * dzeta, y * dxi, y * deta, y * dzeta,
                       z * dxi, z * deta, z * dzeta};
#pragma unroll
        for (int m = 1; m < 8; m <<= 1)
#pragma unroll
            for (int i = 0; i < 9; i++) J[i] += __shfl_xor(J[i], m, 64);
        double ixi[3], ieta[3], izeta[3];
        const double jac = inv3(J, ixi, ieta, izeta);
        sg[wave][q][n][0] = float(dxi * ixi[0] + deta * ieta[0] + dzeta * izeta[0]);
        sg[wave][q][n][1] = float(dxi * ixi[1] + deta * ieta[1] + dzeta * izeta[1]);
        sg[wave][q][n][2] = float(dxi * ixi[2] + deta * ieta[2] + dzeta * izeta[2]);
        if (n == 0) sc[wave][q] = float(jac);
        if (q == 0)
            scon[wave][n] = A.con ? uint8_t(A.con[int64_t(rn) * 3] | (A.con[int64_t(rn) * 3 + 1] << 1) |
                                            (A.con[int64_t(rn) * 3 + 2] << 2))
                                  : uint8_t(0);
    }
    __syncthreads();
    {
        const int li = lane & 15, lk = lane >> 4;
#pragma unroll
        for (int tile = 0; tile < 4; tile++) {
            const int I = tile >> 1, Jt = tile & 1;
            const int ra = 16 * I + li, cb = 16 * Jt + li;   // this lane's A row / B column
            mastg_floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < 2; ks++) {
                const int q = 4 * ks + lk;
                const float av = ra < 24 ? sc[wave][q] * sg[wave][q][ra & 7][ra >> 3] : 0.f;
                const float bv = cb < 24 ? sg[wave][q][cb & 7][cb >> 3] : 0.f;
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
            }
            // acc[r]: row 16 I + 4 lk + r, column 16 Jt + li
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = 16 * I + 4 * lk + r, j = 16 * Jt + li;
                if (i < 24 && j < 24) sP[wave][i][j] = acc[r];
            }
        }
    }
    __syncthreads();
    if (!valid) return;
    const double Ee = A.E0 * pow(A.param[e], A.penalty);
    const float lam = float(Ee * A.Dhat[1]), mu = float(Ee * A.Dhat[21]);
    float* __restrict__ out = Kel32 + e * 576;
#pragma unroll
    for (int jj = 0; jj < 9; jj++) {
        // entry t = (a 8 + b) 9 + v 3 + w of the element's blocks
        const int t = lane + 64 * jj;
        const int ab = t / 9, vw = t - 9 * ab, a = ab >> 3, b = ab & 7, v = vw / 3, w = vw - 3 * v;
        const float tr = sP[wave][a][b] + sP[wave][8 + a][8 + b] + sP[wave][16 + a][16 + b];
        float kv = lam * sP[wave][v * 8 + a][w * 8 + b] + mu * sP[wave][w * 8 + a][v * 8 + b] + (v == w ? mu * tr : 0.f);
        if (((scon[wave][a] >> v) & 1) || ((scon[wave][b] >> w) & 1)) kv = (a == b && v == w) ? 1.0f : 0.0f;
        out[t] = kv;
    }
}

// segmented reduction of FP32 element blocks into the FP64 CSR values
__global__ __launch_bounds__(256) void k_scatter_blocks_f32(ScatterArgs S, const float* __restrict__ Kel32) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t r = int64_t(blockIdx.x) * 4 + wave;
    if (r >= S.n_nodes) return;
    const int nav = S.nav, nav2 = nav * nav;
    const int64_t k0 = S.blk_ptr[r];
    // a row's chunk is small (at most a few thousand values): 32-bit index
    // arithmetic, the divisors uniform across the wave
    const int deg = int(S.blk_ptr[r + 1] - k0);
    const int cnt = deg * nav2, dn = deg * nav;
    const int64_t base = k0 * nav2;
    for (int t = lane; t < cnt; t += 64) {
        const int v = t / dn;
        const int rem = t - v * dn;
        const int k = rem / nav;
        const int w = rem - k * nav;
        const int64_t blk = k0 + k;
        double s = 0.0;
        for (int32_t c = S.cptr[blk]; c < S.cptr[blk + 1]; c++)
            s += double(Kel32[int64_t(S.cidx[c]) * nav2 + v * nav + w]);
        S.val[base + t] = s;
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
// quad4 Mindlin plate: StructuralElement2D::internal_residual with the
// MindlinBendingOperator (src/elasticity/structural_element_2d.cpp:1075-1227,
// 1765-2019; src/elasticity/mindlin_bending_operator.cpp:57-361), linear strain,
// in the global x-y plane for an element in a z = const plane, else in the
// element's own frame (plate_frame.h).
//   membrane  eps = Bm u  (exx = du/dx, eyy = dv/dy, gxy = du/dy + dv/dx)
//   bending   kap = Bb u  (kx = d ty/dx, ky = -d tx/dy, kxy = d ty/dy - d tx/dx)
//   shear     gam = Bs u  (gxz = dw/dx + ty, gyz = dw/dy - tx) on the 1-point rule
//   K = sum_q c [Bm^T A Bm + Bm^T B Bb + Bb^T B^T Bm + Bb^T Db Bb] + c_s Bs^T S Bs
// with A = Dps h, B = Dps h off, Db = Dps (h^3/12 + h off^2), S = G h kappa I
// (Solid2DSectionProperty, solid_2d_section_element_property_card.cpp:105-157, 393-531).
// One wave = 4 elements x 16 node pairs; lane (a, b) builds the NAV x NAV block.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void quad4_shape(int n, double xi, double eta, double& N, double& dxi,
                                            double& deta) {
    const double sx = (n == 1 || n == 2) ? 1.0 : -1.0;
    const double sy = (n >= 2) ? 1.0 : -1.0;
    const double fx = 0.5 * (1.0 + sx * xi), fy = 0.5 * (1.0 + sy * eta);
    N = fx * fy;
    dxi = 0.5 * sx * fy;
    deta = fx * 0.5 * sy;
}

// physical gradients of the 4 shape functions at (xi, eta) from 4 node
// coordinates; returns the area Jacobian
__device__ __forceinline__ double quad4_grad(const double X[4][2], double xi, double eta,
                                            double g[4][2], double N[4]) {
    double dN[4][2];
    double j00 = 0, j01 = 0, j10 = 0, j11 = 0;   // d x_i / d xi_r
#pragma unroll
    for (int n = 0; n < 4; n++) {
        quad4_shape(n, xi, eta, N[n], dN[n][0], dN[n][1]);
        j00 += X[n][0] * dN[n][0]; j01 += X[n][0] * dN[n][1];
        j10 += X[n][1] * dN[n][0]; j11 += X[n][1] * dN[n][1];
    }
    const double det = j00 * j11 - j01 * j10;
    const double id = 1.0 / det;
    // inverse map: dxi/dx = j11/det, dxi/dy = -j01/det, deta/dx = -j10/det, deta/dy = j00/det
#pragma unroll
    for (int n = 0; n < 4; n++) {
        g[n][0] = (dN[n][0] * j11 - dN[n][1] * j10) * id;
        g[n][1] = (-dN[n][0] * j01 + dN[n][1] * j00) * id;
    }
    return det;
}

// NAV 5: every element in a z = const plane, active u v w tx ty.
// NAV 6: all six variables; an element out of the x-y plane is computed in
// its own frame on the surface map and its block goes to the global system as
// T6 K T6^T, T6 = diag(T, T); the Jacobian carries MAST's 1e-8 on the local
// theta_z diagonal (structural_element_2d.cpp:1215-1221), the residual does not.
template <int NAV>
__global__ __launch_bounds__(256) void k_quad4_mindlin_K(Quad4Args A) {
    const int sub = threadIdx.x >> 4, lane = threadIdx.x & 15;   // 16 elements per block
    const int64_t e = int64_t(blockIdx.x) * 16 + sub;
    if (e >= A.n_elems) return;
    const int a = lane >> 2, b = lane & 3;
    double X[4][2], XL[4][3], T[9];
    bool fr = false;
    {
        double XG[4][3];
#pragma unroll
        for (int n = 0; n < 4; n++) {
            const int64_t r = A.econn[e * 4 + n];
            X[n][0] = XG[n][0] = A.xyz[r * 3 + 0];
            X[n][1] = XG[n][1] = A.xyz[r * 3 + 1];
            XG[n][2] = A.xyz[r * 3 + 2];
        }
        if constexpr (NAV == 6) fr = quad4_frame(XG, T, XL);
    }
    auto grad = [&](double xi, double eta, double g[4][2], double N[4]) -> double {
        if constexpr (NAV == 6) return quad4_grad_surface(XL, xi, eta, g, N);
        else return quad4_grad(X, xi, eta, g, N);
    };
    // unit-section blocks: membrane (u,v)x(u,v), coupling (u,v)x(tx,ty),
    // bending (tx,ty)x(tx,ty), shear (w,tx,ty)x(w,tx,ty)
    double Km[4] = {0, 0, 0, 0}, Kc[4] = {0, 0, 0, 0}, Kd[4] = {0, 0, 0, 0}, Kt[4] = {0, 0, 0, 0};
    double Ks[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#define DP(i) A.Dps[i]
#pragma unroll
    for (int q = 0; q < 4; q++) {
        double g[4][2], N[4];
        const double c = grad(gp_coord(q & 1), gp_coord(q >> 1), g, N) * 1.0;
        const double ax = g[a][0], ay = g[a][1], bx = g[b][0], by = g[b][1];
        // Bm_b columns (u, v): [bx, 0, by], [0, by, bx]; Bb_b columns (tx, ty): [0, -by, -bx], [bx, 0, by]
        const double Bm_b[2][3] = {{bx, 0.0, by}, {0.0, by, bx}};
        const double Bb_b[2][3] = {{0.0, -by, -bx}, {bx, 0.0, by}};
        const double Bm_a[2][3] = {{ax, 0.0, ay}, {0.0, ay, ax}};
        const double Bb_a[2][3] = {{0.0, -ay, -ax}, {ax, 0.0, ay}};
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) {
                double mm = 0, mb = 0, bm = 0, bb = 0;
#pragma unroll
                for (int r = 0; r < 3; r++) {
                    double dmb = 0, dbb = 0, dmm = 0, dbm = 0;
#pragma unroll
                    for (int s2 = 0; s2 < 3; s2++) {
                        dmm += DP(r * 3 + s2) * Bm_b[j][s2];
                        dbb += DP(r * 3 + s2) * Bb_b[j][s2];
                    }
                    dmb = dbb;   // B = Dps * (h off): same unit matrix, other factor
                    dbm = dmm;
                    mm += Bm_a[i][r] * dmm;
                    mb += Bm_a[i][r] * dmb;
                    bm += Bb_a[i][r] * dbm;
                    bb += Bb_a[i][r] * dbb;
                }
                Km[i * 2 + j] += c * mm;
                Kc[i * 2 + j] += c * mb;
                Kt[i * 2 + j] += c * bm;
                Kd[i * 2 + j] += c * bb;
            }
    }
#undef DP
    {
        // transverse shear on the 1-point rule (weight 4)
        double g[4][2], N[4];
        const double c = grad(0.0, 0.0, g, N) * 4.0;
        // Bs columns (w, tx, ty): gxz = [gx, 0, N], gyz = [gy, -N, 0]
        const double Ba[2][3] = {{g[a][0], 0.0, N[a]}, {g[a][1], -N[a], 0.0}};
        const double Bb2[2][3] = {{g[b][0], 0.0, N[b]}, {g[b][1], -N[b], 0.0}};
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) Ks[i * 3 + j] += c * (Ba[0][i] * Bb2[0][j] + Ba[1][i] * Bb2[1][j]);
    }
    const double h = A.param[e];
    const double fA = h, fB = h * A.off, fD = h * h * h / 12.0 + h * A.off * A.off, fS = A.G * h * A.kappa;
    double K[NAV * NAV];
#pragma unroll
    for (int i = 0; i < NAV * NAV; i++) K[i] = 0.0;
    // vars: 0 u, 1 v, 2 w, 3 tx, 4 ty (5 tz: no stiffness in the local frame)
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            K[i * NAV + j] = fA * Km[i * 2 + j];
            K[i * NAV + 3 + j] = fB * Kc[i * 2 + j];
            K[(3 + i) * NAV + j] = fB * Kt[i * 2 + j];
            K[(3 + i) * NAV + 3 + j] = fD * Kd[i * 2 + j];
        }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) K[(2 + i) * NAV + 2 + j] += fS * Ks[i * 3 + j];
    if constexpr (NAV == 6) {
        if (fr) {
            // K <- T6 K T6^T: translations (0..2) and rotations (3..5) by T
            double W[36];
#pragma unroll
            for (int i = 0; i < 6; i++)
#pragma unroll
                for (int j = 0; j < 6; j++) {
                    const int bi = i / 3 * 3, ii = i % 3;
                    W[i * 6 + j] = T[ii * 3 + 0] * K[(bi + 0) * 6 + j] + T[ii * 3 + 1] * K[(bi + 1) * 6 + j] +
                                   T[ii * 3 + 2] * K[(bi + 2) * 6 + j];
                }
#pragma unroll
            for (int i = 0; i < 6; i++)
#pragma unroll
                for (int j = 0; j < 6; j++) {
                    const int bj = j / 3 * 3, jj = j % 3;
                    K[i * 6 + j] = W[i * 6 + bj + 0] * T[jj * 3 + 0] + W[i * 6 + bj + 1] * T[jj * 3 + 1] +
                                   W[i * 6 + bj + 2] * T[jj * 3 + 2];
                }
        }
    }
    const int32_t ra = A.econn[e * 4 + a], rb = A.econn[e * 4 + b];
    if (A.X) {
        double f[NAV];
#pragma unroll
        for (int v = 0; v < NAV; v++) {
            double t = 0.0;
#pragma unroll
            for (int w = 0; w < NAV; w++) t += K[v * NAV + w] * A.X[int64_t(rb) * NAV + w];
            f[v] = t;
        }
#pragma unroll
        for (int m = 1; m < 4; m <<= 1)
#pragma unroll
            for (int v = 0; v < NAV; v++) f[v] += __shfl_xor(f[v], m, 64);
        if (b == 0)
#pragma unroll
            for (int v = 0; v < NAV; v++)
                A.Rel[(e * 4 + a) * NAV + v] = (A.con && A.con[int64_t(ra) * NAV + v]) ? 0.0 : f[v];
    }
    if (A.Kel) {
        if constexpr (NAV == 6) {
            // 1e-8 on the local theta_z diagonal: 1e-8 v3 v3^T on the rotations
            if (a == b)
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++) K[(3 + i) * 6 + 3 + j] += 1.0e-8 * T[i * 3 + 2] * T[j * 3 + 2];
        }
        double* out = A.Kel + (e * 16 + lane) * (NAV * NAV);
#pragma unroll
        for (int v = 0; v < NAV; v++)
#pragma unroll
            for (int w = 0; w < NAV; w++) {
                double kv = K[v * NAV + w];
                if (A.con && (A.con[int64_t(ra) * NAV + v] || A.con[int64_t(rb) * NAV + w]))
                    kv = (a == b && v == w) ? 1.0 : 0.0;
                out[v * NAV + w] = kv;
            }
    }
}

// g_e = lambda_e . dR_e/dh_e with the section derivatives
// dA = Dps, dB = Dps off, dDb = Dps (h^2/4 + off^2), dS = G kappa
// (..._StiffnessMatrix::derivative, solid_2d_section_element_property_card.cpp:105-157, 393-531);
// NAV 6 out of the x-y plane: u and lambda taken to the element frame by T^T
template <int NAV>
__global__ __launch_bounds__(128) void k_quad4_mindlin_sens(Quad4Args A, const double* __restrict__ lam,
                                                            double* __restrict__ gout) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= A.n_elems) return;
    double X[4][2], XL[4][3], u[4][NAV], l[4][NAV];
    double XG[4][3];
#pragma unroll
    for (int n = 0; n < 4; n++) {
        const int64_t r = A.econn[e * 4 + n];
        X[n][0] = XG[n][0] = A.xyz[r * 3 + 0];
        X[n][1] = XG[n][1] = A.xyz[r * 3 + 1];
        XG[n][2] = A.xyz[r * 3 + 2];
#pragma unroll
        for (int v = 0; v < NAV; v++) {
            u[n][v] = A.X[r * NAV + v];
            l[n][v] = A.con[r * NAV + v] ? 0.0 : lam[r * NAV + v];
        }
    }
    if constexpr (NAV == 6) {
        double T[9];
        if (quad4_frame(XG, T, XL)) {
#pragma unroll
            for (int n = 0; n < 4; n++)
#pragma unroll
                for (int blk = 0; blk < 6; blk += 3) {
                    double tu[3], tl[3];
#pragma unroll
                    for (int j = 0; j < 3; j++) {
                        tu[j] = T[0 * 3 + j] * u[n][blk] + T[1 * 3 + j] * u[n][blk + 1] + T[2 * 3 + j] * u[n][blk + 2];
                        tl[j] = T[0 * 3 + j] * l[n][blk] + T[1 * 3 + j] * l[n][blk + 1] + T[2 * 3 + j] * l[n][blk + 2];
                    }
#pragma unroll
                    for (int j = 0; j < 3; j++) {
                        u[n][blk + j] = tu[j];
                        l[n][blk + j] = tl[j];
                    }
                }
        }
    }
    auto grad = [&](double xi, double eta, double g[4][2], double N[4]) -> double {
        if constexpr (NAV == 6) return quad4_grad_surface(XL, xi, eta, g, N);
        else return quad4_grad(X, xi, eta, g, N);
    };
    const double h = A.param[e];
    const double dA = 1.0, dB = A.off, dD = h * h / 4.0 + A.off * A.off, dS = A.G * A.kappa;
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        double g[4][2], N[4];
        const double c = grad(gp_coord(q & 1), gp_coord(q >> 1), g, N) * 1.0;
        double em[3] = {0, 0, 0}, kb[3] = {0, 0, 0}, lm[3] = {0, 0, 0}, lk[3] = {0, 0, 0};
#pragma unroll
        for (int n = 0; n < 4; n++) {
            const double gx = g[n][0], gy = g[n][1];
            em[0] += gx * u[n][0]; em[1] += gy * u[n][1]; em[2] += gy * u[n][0] + gx * u[n][1];
            kb[0] += gx * u[n][4]; kb[1] -= gy * u[n][3]; kb[2] += gy * u[n][4] - gx * u[n][3];
            lm[0] += gx * l[n][0]; lm[1] += gy * l[n][1]; lm[2] += gy * l[n][0] + gx * l[n][1];
            lk[0] += gx * l[n][4]; lk[1] -= gy * l[n][3]; lk[2] += gy * l[n][4] - gx * l[n][3];
        }
        double t = 0.0;
#pragma unroll
        for (int r = 0; r < 3; r++) {
            double Dem = 0, Dkb = 0;
#pragma unroll
            for (int s2 = 0; s2 < 3; s2++) {
                Dem += A.Dps[r * 3 + s2] * em[s2];
                Dkb += A.Dps[r * 3 + s2] * kb[s2];
            }
            t += lm[r] * (dA * Dem + dB * Dkb) + lk[r] * (dB * Dem + dD * Dkb);
        }
        acc += c * t;
    }
    {
        double g[4][2], N[4];
        const double c = grad(0.0, 0.0, g, N) * 4.0;
        double gu[2] = {0, 0}, gl[2] = {0, 0};
#pragma unroll
        for (int n = 0; n < 4; n++) {
            gu[0] += g[n][0] * u[n][2] + N[n] * u[n][4];
            gu[1] += g[n][1] * u[n][2] - N[n] * u[n][3];
            gl[0] += g[n][0] * l[n][2] + N[n] * l[n][4];
            gl[1] += g[n][1] * l[n][2] - N[n] * l[n][3];
        }
        acc += c * dS * (gl[0] * gu[0] + gl[1] * gu[1]);
    }
    gout[e] = acc;
}


// ---------------------------------------------------------------------------
// hex8 steady conduction: HeatConductionElementBase::internal_residual with the
// temperature-dependent conductance k(T) = k0 (1 + alpha T) evaluated at the
// quadrature-point temperature, its exact Jacobian including the dk/dT term
// (src/heat_conduction/heat_conduction_elem_base.cpp:53-149), convection on
// the element's marked sides (:1099-1154) and the volumetric source (:1709-1745):
//   R_a = sum_q c k(T_q) g_a.gradT_q + sum_sides int h N_a (T - Tinf) - q_e int N_a
//   J_ab = sum_q c [k(T_q) g_a.g_b + k' (gradT_q.g_a) N_b] + sum_sides int h N_a N_b
// One wave per element.  Lanes (q, n) form the geometry at quadrature point q;
// lane q then sums T_q and gradT_q over the 8 nodes (once per point, not per
// matrix entry, in node order); lanes (a, b) form J_ab and R_a from the
// per-point values in LDS.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void hex8_side_xi(int side, double s1, double s2, double& xi, double& eta,
                                             double& zeta, int& r0, int& r1) {
    switch (side) {
        case 0: xi = s1; eta = s2; zeta = -1; r0 = 0; r1 = 1; break;
        case 1: xi = s1; eta = -1; zeta = s2; r0 = 0; r1 = 2; break;
        case 2: xi = 1; eta = s1; zeta = s2; r0 = 1; r1 = 2; break;
        case 3: xi = s1; eta = 1; zeta = s2; r0 = 0; r1 = 2; break;
        case 4: xi = -1; eta = s1; zeta = s2; r0 = 1; r1 = 2; break;
        default: xi = s1; eta = s2; zeta = 1; r0 = 0; r1 = 1; break;
    }
}

// lane (q, n) of a wave: node n's coordinates, the physical gradient g and
// value N of N_n at quadrature point q of element ee, and det J at q (the
// Jacobian summed over the 8 node lanes of q)
__device__ __forceinline__ void heat_lane_geometry(const double* __restrict__ xyz, const int32_t* __restrict__ econn,
                                                   int64_t ee, int q, int n, double xn[3], double g[3], double& N,
                                                   double& jac) {
    const int32_t rn = econn[ee * 8 + n];
    xn[0] = xyz[int64_t(rn) * 3 + 0];
    xn[1] = xyz[int64_t(rn) * 3 + 1];
    xn[2] = xyz[int64_t(rn) * 3 + 2];
    double dxi, deta, dzeta;
    hex8_dshape(n, gp_coord(q & 1), gp_coord((q >> 1) & 1), gp_coord(q >> 2), dxi, deta, dzeta, N);
    double J[9] = {xn[0] * dxi, xn[0] * deta, xn[0] * dzeta, xn[1] * dxi, xn[1] * deta, xn[1] * dzeta,
                   xn[2] * dxi, xn[2] * deta, xn[2] * dzeta};
#pragma unroll
    for (int m = 1; m < 8; m <<= 1)
#pragma unroll
        for (int i = 0; i < 9; i++) J[i] += __shfl_xor(J[i], m, 64);
    double ixi[3], ieta[3], izeta[3];
    jac = inv3(J, ixi, ieta, izeta);
#pragma unroll
    for (int d = 0; d < 3; d++) g[d] = dxi * ixi[d] + deta * ieta[d] + dzeta * izeta[d];
}

// the geometry of each congruence class (its representative element): per
// class 8 x 8 x 3 gradients [q][n][d] then 8 JxW; one wave per class
__global__ __launch_bounds__(64) void k_hex8_heat_geometry(const double* __restrict__ xyz,
                                                           const int32_t* __restrict__ econn,
                                                           const int32_t* __restrict__ class_rep,
                                                           double* __restrict__ geo) {
    const int cls = blockIdx.x, lane = threadIdx.x, q = lane >> 3, n = lane & 7;
    double xn[3], g[3], N, jac;
    heat_lane_geometry(xyz, econn, class_rep[cls], q, n, xn, g, N, jac);
    double* out = geo + int64_t(cls) * kHeatGeo;
#pragma unroll
    for (int d = 0; d < 3; d++) out[(q * 8 + n) * 3 + d] = g[d];
    if (n == 0) out[192 + q] = jac * 1.0;
}

// CLASS: the element's geometry is its congruence class's (A.geo, A.eclass:
// no Jacobian per element); else formed per element
template <bool CLASS>
__global__ __launch_bounds__(256) void k_hex8_heat_K(HeatArgs A) {
    __shared__ double sg[4][8][8][3];
    __shared__ double sN[4][8][8];
    __shared__ double sc[4][8];
    __shared__ double sT[4][8];
    __shared__ double sX[4][8][3];
    __shared__ double sTq[4][8], sgT[4][8][3];   // T and grad T at each quadrature point
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t e = int64_t(blockIdx.x) * 4 + wave;
    const bool valid = e < A.n_elems;
    const int64_t ee = valid ? e : 0;
    {
        const int q = lane >> 3, n = lane & 7;
        if (CLASS) {
            const double* __restrict__ G = A.geo + int64_t(A.eclass[ee]) * kHeatGeo;
            sg[wave][q][n][0] = G[lane * 3 + 0];
            sg[wave][q][n][1] = G[lane * 3 + 1];
            sg[wave][q][n][2] = G[lane * 3 + 2];
            if (n == 0) sc[wave][q] = G[192 + q];
            double d0, d1, d2, N;
            hex8_dshape(n, gp_coord(q & 1), gp_coord((q >> 1) & 1), gp_coord(q >> 2), d0, d1, d2, N);
            sN[wave][q][n] = N;
            if (q == 0) {
                const int32_t rn = A.econn[ee * 8 + n];
                sT[wave][n] = A.X[rn];
                // node coordinates only for the faces of elements on marked sides
                const bool faces = (A.sides && A.sides[ee]) || (A.rsides && A.rsides[ee]);
                if (faces)
#pragma unroll
                    for (int d = 0; d < 3; d++) sX[wave][n][d] = A.xyz[int64_t(rn) * 3 + d];
            }
        } else {
            double xn[3], g[3], N, jac;
            heat_lane_geometry(A.xyz, A.econn, ee, q, n, xn, g, N, jac);
            if (q == 0) {
                sT[wave][n] = A.X[A.econn[ee * 8 + n]];
                sX[wave][n][0] = xn[0]; sX[wave][n][1] = xn[1]; sX[wave][n][2] = xn[2];
            }
            sg[wave][q][n][0] = g[0];
            sg[wave][q][n][1] = g[1];
            sg[wave][q][n][2] = g[2];
            sN[wave][q][n] = N;
            if (n == 0) sc[wave][q] = jac * 1.0;
        }
    }
    __syncthreads();
    if (lane < 8) {   // T_q = sum_n N_n T_n, gradT_q = sum_n g_n T_n
        const int q = lane;
        double Tq = 0.0, gT0 = 0.0, gT1 = 0.0, gT2 = 0.0;
#pragma unroll
        for (int n = 0; n < 8; n++) {
            const double t = sT[wave][n];
            Tq += sN[wave][q][n] * t;
            gT0 += sg[wave][q][n][0] * t;
            gT1 += sg[wave][q][n][1] * t;
            gT2 += sg[wave][q][n][2] * t;
        }
        sTq[wave][q] = Tq;
        sgT[wave][q][0] = gT0;
        sgT[wave][q][1] = gT1;
        sgT[wave][q][2] = gT2;
    }
    __syncthreads();
    if (!valid) return;
    const int a = lane >> 3, b = lane & 7;
    double K = 0.0, f = 0.0;
    const double dk = A.k0 * A.alpha;
#pragma unroll 2
    for (int q = 0; q < 8; q++) {
        const double Tq = sTq[wave][q], gT0 = sgT[wave][q][0], gT1 = sgT[wave][q][1], gT2 = sgT[wave][q][2];
        const double c = sc[wave][q];
        const double kk = A.k0 * (1.0 + A.alpha * Tq);
        const double ga0 = sg[wave][q][a][0], ga1 = sg[wave][q][a][1], ga2 = sg[wave][q][a][2];
        const double gb0 = sg[wave][q][b][0], gb1 = sg[wave][q][b][1], gb2 = sg[wave][q][b][2];
        K += c * kk * (ga0 * gb0 + ga1 * gb1 + ga2 * gb2);
        K += c * (gT0 * ga0 + gT1 * ga1 + gT2 * ga2) * dk * sN[wave][q][b];
        if (b == 0) {
            f += c * kk * (ga0 * gT0 + ga1 * gT1 + ga2 * gT2);
            f -= c * sN[wave][q][a] * (A.param[e] * A.src_scale);
        }
    }
    // convection and radiation on marked sides (2x2 Gauss on the face)
    const int cm = A.sides ? A.sides[e] : 0;
    const int rm = A.rsides ? A.rsides[e] : 0;
    const int sm = cm | rm;
    if (sm)
        for (int sd = 0; sd < 6; sd++) {
            if (!((sm >> sd) & 1)) continue;
            const bool conv = (cm >> sd) & 1, rad = (rm >> sd) & 1;
            for (int qq = 0; qq < 4; qq++) {
                double xi, eta, zeta;
                int r0, r1;
                hex8_side_xi(sd, gp_coord(qq & 1), gp_coord(qq >> 1), xi, eta, zeta, r0, r1);
                double Na = 0, Nb = 0, T = 0, t0[3] = {0, 0, 0}, t1[3] = {0, 0, 0};
                // not unrolled: the unrolled face loop set the register
                // budget of the whole kernel (160 VGPRs, 3 waves per SIMD,
                // against 70 and 7), for the few elements on convecting faces
#pragma unroll 1
                for (int n = 0; n < 8; n++) {
                    double d[3], N;
                    hex8_dshape(n, xi, eta, zeta, d[0], d[1], d[2], N);
                    if (n == a) Na = N;
                    if (n == b) Nb = N;
                    T += N * sT[wave][n];
#pragma unroll
                    for (int i = 0; i < 3; i++) {
                        t0[i] += sX[wave][n][i] * d[r0];
                        t1[i] += sX[wave][n][i] * d[r1];
                    }
                }
                const double g11 = t0[0] * t0[0] + t0[1] * t0[1] + t0[2] * t0[2];
                const double g12 = t0[0] * t1[0] + t0[1] * t1[1] + t0[2] * t1[2];
                const double g22 = t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2];
                const double jxw = sqrt(g11 * g22 - g12 * g12) * 1.0;
                if (conv) {
                    K += jxw * (Na * Nb) * A.h;
                    if (b == 0) f += jxw * Na * A.h * (T - A.Tinf);
                }
                if (rad) {   // surface_radiation_residual (heat_conduction_elem_base.cpp:1402-1463)
                    const double ts = T + A.rad_T0, ta = A.rad_Tamb + A.rad_T0;
                    const double ts3 = ts * ts * ts;
                    K += jxw * (Na * Nb) * A.rad_se * 4.0 * ts3;
                    if (b == 0) f += jxw * Na * A.rad_se * (ts3 * ts - ta * ta * ta * ta);
                }
            }
        }
    const int32_t ra = A.econn[e * 8 + a], rb = A.econn[e * 8 + b];
    const bool ca = A.con && A.con[ra], cb = A.con && A.con[rb];
    if (A.Rel && b == 0) A.Rel[e * 8 + a] = ca ? 0.0 : f;
    if (A.Kel) A.Kel[e * 64 + lane] = (ca || cb) ? ((a == b) ? 1.0 : 0.0) : K;
}

// g_e = scale * lambda_e . int N (d R/d q_e = -int N: volume_heat_source_residual_sensitivity,
// heat_conduction_elem_base.cpp:1749-1791); constrained entries of lambda ignored
__global__ __launch_bounds__(128) void k_hex8_source_dot(HeatArgs A, const double* __restrict__ lam,
                                                         double scale, double* __restrict__ g) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= A.n_elems) return;
    double x[8][3], l[8];
#pragma unroll
    for (int n = 0; n < 8; n++) {
        const int64_t r = A.econn[e * 8 + n];
#pragma unroll
        for (int i = 0; i < 3; i++) x[n][i] = A.xyz[r * 3 + i];
        l[n] = (A.con && A.con[r]) ? 0.0 : lam[r];
    }
    double acc = 0.0;
    for (int q = 0; q < 8; q++) {
        const double xi = gp_coord(q & 1), eta = gp_coord((q >> 1) & 1), zeta = gp_coord(q >> 2);
        double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, N[8];
#pragma unroll
        for (int n = 0; n < 8; n++) {
            double d[3];
            hex8_dshape(n, xi, eta, zeta, d[0], d[1], d[2], N[n]);
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int r = 0; r < 3; r++) J[i * 3 + r] += x[n][i] * d[r];
        }
        double ixi[3], ieta[3], izeta[3];
        const double jac = inv3(J, ixi, ieta, izeta);
        double t = 0.0;
#pragma unroll
        for (int n = 0; n < 8; n++) t += N[n] * l[n];
        acc += jac * t;
    }
    g[e] = scale * acc;
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
    // a row's chunk is small (at most a few thousand values): 32-bit index
    // arithmetic, the divisors uniform across the wave
    const int deg = int(S.blk_ptr[r + 1] - k0);
    const int cnt = deg * nav2, dn = deg * nav;
    const int64_t base = k0 * nav2;
    for (int t = lane; t < cnt; t += 64) {
        const int v = t / dn;
        const int rem = t - v * dn;
        const int k = rem / nav;
        const int w = rem - k * nav;
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

__global__ __launch_bounds__(256) void k_dot(int64_t n, const double* __restrict__ a,
                                             const double* __restrict__ b, double* __restrict__ partials) {
    double s = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        s += a[i] * b[i];
    block_sum_store(s, partials, blockIdx.x);
}

__global__ __launch_bounds__(256) void k_reduce_to(const double* __restrict__ p, int n,
                                                   double* __restrict__ out) {
    const double s = reduce_partials(p, n);
    if (threadIdx.x == 0) *out = s;
}

// partial sums of sum_i w_i x[idx[i]] over the i with skip[i] == 0 (skip may
// be null): x is read through the active -> caller DOF map, so a caller's
// vector needs no staging copy; fixed grid-stride assignment, deterministic
__global__ __launch_bounds__(256) void k_gather_dot(int64_t n, const double* __restrict__ w,
                                                    const uint8_t* __restrict__ skip,
                                                    const int64_t* __restrict__ idx,
                                                    const double* __restrict__ x, double* __restrict__ partials) {
    double s = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        if (!skip || !skip[i]) s += w[i] * x[idx[i]];
    block_sum_store(s, partials, blockIdx.x);
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
    const double rz = reduce_partials(C.part0, C.nvec);
    const double rr = reduce_partials(C.part1, C.nvec);
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
    const double rz = reduce_partials(C.part0, C.nvec);
    const double bb = reduce_partials(C.part1, C.nvec);
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


// reliable residual replacement (mixed-precision CG): r = rt, z = dinv r
__global__ __launch_bounds__(256) void k_cg_replace(CGArgs C, const double* __restrict__ rt, int restart) {
    double rz = 0.0, rr = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < C.n; i += int64_t(gridDim.x) * 256) {
        const double r = rt[i];
        C.r[i] = r;
        const double z = C.dinv[i] * r;
        C.z[i] = z;
        if (restart) C.p[i] = z;
        rz += r * z;
        rr += r * r;
    }
    block_sum_store(rz, C.part0, blockIdx.x);
    __syncthreads();
    block_sum_store(rr, C.part1, blockIdx.x);
}

__global__ __launch_bounds__(256) void k_cg_replace_scalars(CGArgs C) {
    const double rz = reduce_partials(C.part0, C.nvec);
    const double rr = reduce_partials(C.part1, C.nvec);
    if (threadIdx.x == 0) {
        CGScalars* S = C.S;
        S->rz = rz;
        S->rr = rr;
        if (!isfinite(rr)) S->done = 4;
        else if (rr <= S->tol2 * S->bb) S->done = 1;
        else if (S->iter >= S->max_it) S->done = 2;
        else S->done = 0;
    }
}

// y = A x over node blocks: SW lanes per node row, NAV rows per node; values
// (the FP64 CSR values or their FP32 copy) in the scalar CSR order (row v of
// node r: deg*NAV entries, block-major), read with one column index per block;
// products and sums in FP64; optional partial sums of y.x
template <typename T, int NAV, int SW>
__global__ __launch_bounds__(256) void k_spmv_nb_dot(CGArgs C, const T* __restrict__ vals,
                                                     const double* __restrict__ x, double* __restrict__ y,
                                                     int want_dot, int in_cg_loop) {
    if (in_cg_loop && C.S->done) return;
    const int sub = threadIdx.x & (SW - 1);
    double dsum = 0.0;
    const int64_t nodes_per_grid = (int64_t(gridDim.x) * 256) / SW;
    const int64_t n_nodes = C.n / NAV;
    for (int64_t r = (int64_t(blockIdx.x) * 256 + threadIdx.x) / SW; r < n_nodes; r += nodes_per_grid) {
        const int64_t k0 = C.blk_ptr[r];
        const int deg = int(C.blk_ptr[r + 1] - k0);
        const T* __restrict__ V = vals + k0 * (NAV * NAV);
        double s[NAV];
#pragma unroll
        for (int v = 0; v < NAV; v++) s[v] = 0.0;
        for (int k = sub; k < deg; k += SW) {
            const int64_t cb = int64_t(C.blk_col[k0 + k]) * NAV;
            double xw[NAV];
#pragma unroll
            for (int w = 0; w < NAV; w++) xw[w] = x[cb + w];
#pragma unroll
            for (int v = 0; v < NAV; v++)
#pragma unroll
                for (int w = 0; w < NAV; w++) s[v] += double(V[(int64_t(v) * deg + k) * NAV + w]) * xw[w];
        }
#pragma unroll
        for (int v = 0; v < NAV; v++)
#pragma unroll
            for (int m = SW / 2; m >= 1; m >>= 1) s[v] += __shfl_xor(s[v], m, SW);
        if (sub == 0)
#pragma unroll
            for (int v = 0; v < NAV; v++) {
                y[r * NAV + v] = s[v];
                dsum += s[v] * x[r * NAV + v];
            }
    }
    if (want_dot) block_sum_store(dsum, C.part0, blockIdx.x);
}

__global__ void k_to_f32(int64_t n, const double* __restrict__ src, float* __restrict__ dst) {
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        dst[i] = float(src[i]);
}

// Plain device copy (16-byte lanes): the reference kernel of known traffic that
// hardware byte counters are scaled against.
__global__ __launch_bounds__(256) void k_copy_f64(int64_t n2, const double2* __restrict__ src,
                                                  double2* __restrict__ dst) {
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n2; i += int64_t(gridDim.x) * 256)
        dst[i] = src[i];
}

// ------------------------------------------------- BiCGStab (FP64 CSR) ----
// y = A x with SW lanes per row; optional partials sum(y.w1) -> part0 and
// sum(y.y) -> part1 (per block).  Stops immediately once the solver is done.
template <int SW>
__global__ __launch_bounds__(256) void k_spmv_csr_gen(BiCGArgs B, const double* __restrict__ x,
                                                      double* __restrict__ y, const double* __restrict__ w1,
                                                      int want_yy) {
    if (B.S->done) return;
    const int sub = threadIdx.x & (SW - 1);
    double d1 = 0.0, d2 = 0.0;
    const int64_t rows_per_grid = (int64_t(gridDim.x) * 256) / SW;
    for (int64_t row = (int64_t(blockIdx.x) * 256 + threadIdx.x) / SW; row < B.n; row += rows_per_grid) {
        const int64_t k0 = B.row_ptr[row], k1 = B.row_ptr[row + 1];
        double s = 0.0;
#pragma unroll 4
        for (int64_t k = k0 + sub; k < k1; k += SW) s += B.val[k] * x[B.col[k]];
#pragma unroll
        for (int m = SW / 2; m >= 1; m >>= 1) s += __shfl_xor(s, m, SW);
        if (sub == 0) {
            y[row] = s;
            if (w1) d1 += s * w1[row];
            d2 += s * s;
        }
    }
    if (w1) block_sum_store(d1, B.part0, blockIdx.x);
    if (want_yy) {
        __syncthreads();
        block_sum_store(d2, B.part1, blockIdx.x);
    }
}

__global__ __launch_bounds__(256) void k_bicg_start(BiCGArgs B, const double* __restrict__ b) {
    double bb = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < B.n; i += int64_t(gridDim.x) * 256) {
        const double v = b[i];
        B.x[i] = 0.0; B.r[i] = v; B.r0[i] = v; B.p[i] = 0.0; B.v[i] = 0.0;
        bb += v * v;
    }
    block_sum_store(bb, B.part0, blockIdx.x);
}

__global__ __launch_bounds__(256) void k_bicg_start_s(BiCGArgs B, double tol2, int max_it) {
    const double bb = reduce_partials(B.part0, B.nvec);
    if (threadIdx.x == 0) {
        BiCGScalars* S = B.S;
        S->bb = bb; S->rr = bb; S->rho_old = 1.0; S->alpha = 1.0; S->omega = 1.0;
        S->iter = 0; S->max_it = max_it; S->tol2 = tol2;
        S->done = (bb == 0.0) ? 1 : (isfinite(bb) ? 0 : 4);
    }
}

// partial r0.r
__global__ __launch_bounds__(256) void k_bicg_rho(BiCGArgs B) {
    if (B.S->done) return;
    double d = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < B.n; i += int64_t(gridDim.x) * 256)
        d += B.r0[i] * B.r[i];
    block_sum_store(d, B.part0, blockIdx.x);
}

// rho, beta; p = r + beta (p - omega v); ph = dinv p  (every block reduces the
// partials itself: deterministic, one launch less)
__global__ __launch_bounds__(256) void k_bicg_p(BiCGArgs B) {
    if (B.S->done) return;
    const double rho = reduce_partials(B.part0, B.nvec);
    const double beta = (rho / B.S->rho_old) * (B.S->alpha / B.S->omega);
    const double om = B.S->omega;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < B.n; i += int64_t(gridDim.x) * 256) {
        const double p = B.r[i] + beta * (B.p[i] - om * B.v[i]);
        B.p[i] = p;
        B.ph[i] = B.dinv[i] * p;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        B.S->beta = beta;
        B.S->rho = rho;
    }
}

// alpha = rho / r0.v
__global__ __launch_bounds__(256) void k_bicg_alpha(BiCGArgs B) {
    if (B.S->done) return;
    const double r0v = reduce_partials(B.part0, B.nparts);
    if (threadIdx.x == 0) {
        B.S->alpha = B.S->rho / r0v;
        if (!(r0v != 0.0) || !isfinite(r0v)) B.S->done = 3;
    }
}

// s = r - alpha v ; sh = dinv s
__global__ __launch_bounds__(256) void k_bicg_s(BiCGArgs B) {
    if (B.S->done) return;
    const double al = B.S->alpha;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < B.n; i += int64_t(gridDim.x) * 256) {
        const double sv = B.r[i] - al * B.v[i];
        B.s[i] = sv;
        B.sh[i] = B.dinv[i] * sv;
    }
}

// omega = (t.s)/(t.t); x += alpha ph + omega sh; r = s - omega t; partial r.r
__global__ __launch_bounds__(256) void k_bicg_xr(BiCGArgs B) {
    if (B.S->done) return;
    const double ts = reduce_partials(B.part0, B.nparts);
    const double tt = reduce_partials(B.part1, B.nparts);
    const double om = tt > 0.0 ? ts / tt : 0.0;
    const double al = B.S->alpha;
    double rr = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < B.n; i += int64_t(gridDim.x) * 256) {
        B.x[i] += al * B.ph[i] + om * B.sh[i];
        const double rv = B.s[i] - om * B.t[i];
        B.r[i] = rv;
        rr += rv * rv;
    }
    __syncthreads();
    block_sum_store(rr, B.part0 + B.nparts, blockIdx.x);
    if (blockIdx.x == 0 && threadIdx.x == 0) B.S->omega = om;
}

__global__ __launch_bounds__(256) void k_bicg_check(BiCGArgs B) {
    if (B.S->done) return;
    const double rr = reduce_partials(B.part0 + B.nparts, B.nvec);
    if (threadIdx.x == 0) {
        BiCGScalars* S = B.S;
        S->rr = rr;
        S->iter += 1;
        S->rho_old = S->rho;
        if (!isfinite(rr)) S->done = 4;
        else if (rr <= S->tol2 * S->bb) S->done = 1;
        else if (S->omega == 0.0) S->done = 3;
        else if (S->iter >= S->max_it) S->done = 2;
    }
}

__global__ void k_gather_values(int64_t n, const int32_t* __restrict__ idx, const double* __restrict__ src,
                                double* __restrict__ dst) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

// ---------------------------------------------------------- launchers ----
static inline unsigned blocks_for(int64_t n, int per) { return unsigned((n + per - 1) / per); }

void launch_hex8_K(const Hex8Args& a, hipStream_t s) {
    hipLaunchKernelGGL(k_hex8_elastic_K, dim3(blocks_for(a.n_elems, 4)), dim3(256), 0, s, a);
}
void launch_hex8_K_f32(const Hex8Args& a, float* Kel32, hipStream_t s) {
    hipLaunchKernelGGL(k_hex8_elastic_K_f32, dim3(blocks_for(a.n_elems, 4)), dim3(256), 0, s, a, Kel32);
}
void launch_scatter_f32(const ScatterArgs& a, const float* Kel32, hipStream_t s) {
    hipLaunchKernelGGL(k_scatter_blocks_f32, dim3(blocks_for(a.n_nodes, 4)), dim3(256), 0, s, a, Kel32);
}
void launch_hex8_sens(const Hex8Args& a, const double* lam, double* g, hipStream_t s) {
    hipLaunchKernelGGL(k_hex8_elastic_sens, dim3(blocks_for(a.n_elems, 128)), dim3(128), 0, s, a, lam, g);
}
void launch_quad4_K(const Quad4Args& a, hipStream_t s) {
    if (a.nav == 6)
        hipLaunchKernelGGL(k_quad4_mindlin_K<6>, dim3(blocks_for(a.n_elems, 16)), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(k_quad4_mindlin_K<5>, dim3(blocks_for(a.n_elems, 16)), dim3(256), 0, s, a);
}
void launch_quad4_sens(const Quad4Args& a, const double* lam, double* g, hipStream_t s) {
    if (a.nav == 6)
        hipLaunchKernelGGL(k_quad4_mindlin_sens<6>, dim3(blocks_for(a.n_elems, 128)), dim3(128), 0, s, a, lam, g);
    else
        hipLaunchKernelGGL(k_quad4_mindlin_sens<5>, dim3(blocks_for(a.n_elems, 128)), dim3(128), 0, s, a, lam, g);
}
void launch_heat_K(const HeatArgs& a, hipStream_t s) {
    if (a.geo)
        hipLaunchKernelGGL(k_hex8_heat_K<true>, dim3(blocks_for(a.n_elems, 4)), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(k_hex8_heat_K<false>, dim3(blocks_for(a.n_elems, 4)), dim3(256), 0, s, a);
}
void launch_heat_geometry(int n_classes, const double* xyz, const int32_t* econn, const int32_t* class_rep,
                          double* geo, hipStream_t s) {
    if (n_classes <= 0) return;
    hipLaunchKernelGGL(k_hex8_heat_geometry, dim3(n_classes), dim3(64), 0, s, xyz, econn, class_rep, geo);
}
void launch_heat_source_dot(const HeatArgs& a, const double* lam, double scale, double* g, hipStream_t s) {
    hipLaunchKernelGGL(k_hex8_source_dot, dim3(blocks_for(a.n_elems, 128)), dim3(128), 0, s, a, lam, scale, g);
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
void launch_gather_dot(int64_t n, const double* w, const uint8_t* skip, const int64_t* idx, const double* x,
                       double* partials, int nparts, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_gather_dot, dim3(nparts), dim3(256), 0, s, n, w, skip, idx, x, partials);
    hipLaunchKernelGGL(k_reduce_to, dim3(1), dim3(256), 0, s, partials, nparts, out);
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
// node-block FP32 SpMV: 8 lanes per node row (a hex8 interior node has 27
// blocks, a quad4 one 9), grid of the CSR SpMV
template <typename T>
static void nb_launch(const CGArgs& c, const T* vals, const double* x, double* y, int want_dot, int in_loop,
                      hipStream_t s) {
    switch (c.nav) {
        case 1: hipLaunchKernelGGL((k_spmv_nb_dot<T, 1, 8>), dim3(c.nparts), dim3(256), 0, s, c, vals, x, y,
                                   want_dot, in_loop); break;
        case 5: hipLaunchKernelGGL((k_spmv_nb_dot<T, 5, 8>), dim3(c.nparts), dim3(256), 0, s, c, vals, x, y,
                                   want_dot, in_loop); break;
        case 6: hipLaunchKernelGGL((k_spmv_nb_dot<T, 6, 8>), dim3(c.nparts), dim3(256), 0, s, c, vals, x, y,
                                   want_dot, in_loop); break;
        default: hipLaunchKernelGGL((k_spmv_nb_dot<T, 3, 8>), dim3(c.nparts), dim3(256), 0, s, c, vals, x, y,
                                    want_dot, in_loop); break;
    }
}
static void nb32_dispatch(const CGArgs& c, const double* x, double* y, int want_dot, int in_loop,
                          hipStream_t s) {
    nb_launch<float>(c, c.val32, x, y, want_dot, in_loop, s);
}
// the FP64 operator over node blocks (plates: no column index per value)
static void nb64_dispatch(const CGArgs& c, const double* x, double* y, int want_dot, int in_loop,
                          hipStream_t s) {
    nb_launch<double>(c, c.val, x, y, want_dot, in_loop, s);
}
void launch_spmv_nb32(const CGArgs& c, const double* x, double* y, int want_dot, hipStream_t s) {
    nb32_dispatch(c, x, y, want_dot, 0, s);
}
void launch_spmv_nb64(const CGArgs& c, const double* x, double* y, int want_dot, hipStream_t s) {
    nb64_dispatch(c, x, y, want_dot, 0, s);
}
void launch_to_f32(int64_t n, const double* src, float* dst, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_to_f32, dim3(std::min<int64_t>(blocks_for(n, 256), 8192)), dim3(256), 0, s, n, src, dst);
}
void launch_copy_f64(int64_t n, const double* src, double* dst, hipStream_t s) {
    if (n <= 0 || n % 2) return;
    hipLaunchKernelGGL(k_copy_f64, dim3(std::min<int64_t>(blocks_for(n / 2, 256), 8192)), dim3(256), 0, s, n / 2,
                       reinterpret_cast<const double2*>(src), reinterpret_cast<double2*>(dst));
}
void launch_cg_replace(const CGArgs& c, const double* rt, int restart, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_replace, dim3(c.nvec), dim3(256), 0, s, c, rt, restart);
    hipLaunchKernelGGL(k_cg_replace_scalars, dim3(1), dim3(256), 0, s, c);
}
void launch_cg_start(const CGArgs& c, const double* b, double tol2, int max_it, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_start, dim3(c.nvec), dim3(256), 0, s, c, b);
    hipLaunchKernelGGL(k_cg_start_scalars, dim3(1), dim3(256), 0, s, c, tol2, max_it);
}
static void spmv_gen(const BiCGArgs& B, const double* x, double* y, const double* w1, int want_yy,
                     hipStream_t s) {
    switch (B.spmv_sw) {
        case 4: hipLaunchKernelGGL(k_spmv_csr_gen<4>, dim3(B.nparts), dim3(256), 0, s, B, x, y, w1, want_yy); break;
        case 8: hipLaunchKernelGGL(k_spmv_csr_gen<8>, dim3(B.nparts), dim3(256), 0, s, B, x, y, w1, want_yy); break;
        case 32: hipLaunchKernelGGL(k_spmv_csr_gen<32>, dim3(B.nparts), dim3(256), 0, s, B, x, y, w1, want_yy); break;
        case 64: hipLaunchKernelGGL(k_spmv_csr_gen<64>, dim3(B.nparts), dim3(256), 0, s, B, x, y, w1, want_yy); break;
        default: hipLaunchKernelGGL(k_spmv_csr_gen<16>, dim3(B.nparts), dim3(256), 0, s, B, x, y, w1, want_yy); break;
    }
}
void launch_bicg_start(const BiCGArgs& a, const double* b, double tol2, int max_it, hipStream_t s) {
    hipLaunchKernelGGL(k_bicg_start, dim3(a.nvec), dim3(256), 0, s, a, b);
    hipLaunchKernelGGL(k_bicg_start_s, dim3(1), dim3(256), 0, s, a, tol2, max_it);
}
void launch_bicg_part(const BiCGArgs& a, int part, hipStream_t s) {
    if (part == 0) {
        hipLaunchKernelGGL(k_bicg_rho, dim3(a.nvec), dim3(256), 0, s, a);
        hipLaunchKernelGGL(k_bicg_p, dim3(a.nvec), dim3(256), 0, s, a);
    } else if (part == 1) {
        spmv_gen(a, a.ph, a.v, a.r0, 0, s);
        hipLaunchKernelGGL(k_bicg_alpha, dim3(1), dim3(256), 0, s, a);
        hipLaunchKernelGGL(k_bicg_s, dim3(a.nvec), dim3(256), 0, s, a);
    } else {
        spmv_gen(a, a.sh, a.t, a.s, 1, s);
        hipLaunchKernelGGL(k_bicg_xr, dim3(a.nvec), dim3(256), 0, s, a);
        hipLaunchKernelGGL(k_bicg_check, dim3(1), dim3(256), 0, s, a);
    }
}
void launch_bicg_iteration(const BiCGArgs& a, hipStream_t s) {
    for (int part = 0; part < 3; part++) launch_bicg_part(a, part, s);
}
void launch_gather_values(int64_t n, const int32_t* idx, const double* src, double* dst, hipStream_t s) {
    hipLaunchKernelGGL(k_gather_values, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, idx, src, dst);
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
void launch_mf_elem(const CGArgs& c, const double* x, hipStream_t s) {
    const unsigned ge = blocks_for(c.mf.n_elems, 256);
    if (c.mf.n_classes == 1)
        hipLaunchKernelGGL(k_hex8_mf_elem<true>, dim3(ge), dim3(256), 0, s, c, x);
    else
        hipLaunchKernelGGL(k_hex8_mf_elem<false>, dim3(ge), dim3(256), 0, s, c, x);
}
void launch_cg_op_alpha(const CGArgs& c, hipStream_t s) {
    if (c.use_mf == 3)
        launch_lattice_apply(c, c.p, c.q, s);
    else if (c.use_mf == 1)
        mf_dispatch(c, c.p, c.p, c.q, 1, s);
    else if (c.use_mf == 2)
        nb32_dispatch(c, c.p, c.q, 1, 1, s);
    else if (c.use_mf == 4)
        nb64_dispatch(c, c.p, c.q, 1, 1, s);
    else
        spmv_dispatch(c, c.p, c.q, 1, 1, s);
    hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(256), 0, s, c);
}
void launch_cg_p(const CGArgs& c, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_p, dim3(c.nvec), dim3(256), 0, s, c);
}
void launch_cg_iteration(const CGArgs& c, hipStream_t s) {
    if (c.use_mf == 1)
        mf_dispatch(c, c.p, c.p, c.q, 1, s);
    else if (c.use_mf == 2)
        nb32_dispatch(c, c.p, c.q, 1, 1, s);
    else
        spmv_dispatch(c, c.p, c.q, 1, 1, s);
    hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(256), 0, s, c);
    hipLaunchKernelGGL(k_cg_update, dim3(c.nvec), dim3(256), 0, s, c);
    hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(256), 0, s, c);
    hipLaunchKernelGGL(k_cg_p, dim3(c.nvec), dim3(256), 0, s, c);
}

}  // namespace mastg
