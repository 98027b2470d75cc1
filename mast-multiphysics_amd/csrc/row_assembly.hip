// Fused assembly of hex8 elasticity by CSR row owner: element Jacobian
// blocks, the internal residual, the scatter into the global CSR (libMesh
// constraint convention) and the inverse diagonal in one pass, without the
// element matrices ever reaching memory.
//
// StructuralElement3D::internal_residual, linear strain (src/elasticity/
// solid_element_3d.cpp:269-411): K_e = sum_q JxW_q B_q^T D B_q, here with the
// isotropic D of IsotropicMaterialProperty (isotropic_material_property_card.cpp:
// 512-528) written through its Lame constants, D = lam 1 1^T + mu (2 I on the
// normal strains, I on the engineering shears), so the 3 x 3 block of nodes
// a, b is
//     K_ab[i][j] = E_e sum_q JxW_q (lam g_a,i g_b,j + mu g_a,j g_b,i + mu delta_ij g_a . g_b)
// (g = shape-function gradients at q) -- the same matrix as the dense
// B^T D B product, 9 + 3 multiply-adds per (block, point) instead of 81.
// NonlinearImplicitAssembly::residual_and_jacobian (src/base/
// nonlinear_implicit_assembly.cpp:111-186) adds every element block into the
// rows of its first node; here the owner of a row (one work-group = one wave
// per node) computes exactly the element blocks of that row -- each
// (element, a, b) block is computed once in the whole assembly -- and sums
// them in a fixed order (the node's incidence order), so the assembly is
// deterministic without atomics.  The row's values are written once,
// contiguous (rows r*3..r*3+2 of node r are one chunk ordered (v, block, w)).
//
// Per wave: lanes (s, q) form the geometry of element slot s at quadrature
// point q (Jacobian, its inverse, the 8 shape gradients) into LDS; lanes
// (s, b) then form block (a_s, b) of element s, its constraint treatment
// (zero constrained rows and columns, unit diagonal per element) and its
// residual contribution K_ab u_b (before constraints, as the reference's
// element vector), and add the block into the row accumulator in slot
// order.  A node with more than 8 elements takes several such passes.
#include <hip/hip_runtime.h>

#include "device_util.h"
#include "kernels.h"

namespace mastg {

namespace {

constexpr int kRowMaxBlocks = 64;   // neighbour blocks per row the LDS accumulator holds

__device__ __forceinline__ double ra_gp(int i) { return i ? 0.57735026918962576451 : -0.57735026918962576451; }

__global__ __launch_bounds__(64) void k_hex8_row_assemble(Hex8RowArgs A) {
    __shared__ double g[8][8][8][3];          // [slot][q][node][d] shape gradients
    __shared__ double cq[8][8];               // [slot][q] JxW
    __shared__ double acc[kRowMaxBlocks * 9];  // [block][v][w] row accumulator
    __shared__ int32_t bcol[kRowMaxBlocks];
    __shared__ int32_t selem[8], sla[8];
    const int lane = threadIdx.x;
    const int64_t r = blockIdx.x;
    const int64_t b0 = A.blk_ptr[r];
    const int deg = int(A.blk_ptr[r + 1] - b0);
    const int32_t i0 = A.ninc_ptr[r], ne = A.ninc_ptr[r + 1] - i0;
    if (lane < deg) bcol[lane] = A.blk_col[b0 + lane];
    for (int t = lane; t < deg * 9; t += 64) acc[t] = 0.0;
    double rv0 = 0.0, rv1 = 0.0, rv2 = 0.0;   // residual of row r from this lane's blocks
    const double lam = A.lam, mu = A.mu;
    for (int c0 = 0; c0 < ne; c0 += 8) {
        const int ns = min(8, ne - c0);
        if (lane < ns) {
            const int32_t inc = A.ninc[i0 + c0 + lane];
            selem[lane] = inc >> 3;
            sla[lane] = inc & 7;
        }
        __syncthreads();
        // ---- geometry: lane = (s, q)
        {
            const int s = lane >> 3, q = lane & 7;
            if (s < ns) {
                const int64_t e = selem[s];
                double x[8][3];
#pragma unroll
                for (int n = 0; n < 8; n++) {
                    const int64_t rn = A.econn[e * 8 + n];
#pragma unroll
                    for (int d = 0; d < 3; d++) x[n][d] = A.xyz[rn * 3 + d];
                }
                const double xi = ra_gp(q & 1), eta = ra_gp((q >> 1) & 1), zeta = ra_gp(q >> 2);
                double dN[8][3], J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
                for (int n = 0; n < 8; n++) {
                    const double sx = (n == 1 || n == 2 || n == 5 || n == 6) ? 1.0 : -1.0;
                    const double sy = (n == 2 || n == 3 || n == 6 || n == 7) ? 1.0 : -1.0;
                    const double sz = (n >= 4) ? 1.0 : -1.0;
                    const double fx = 0.5 * (1 + sx * xi), fy = 0.5 * (1 + sy * eta), fz = 0.5 * (1 + sz * zeta);
                    dN[n][0] = 0.5 * sx * fy * fz;
                    dN[n][1] = fx * 0.5 * sy * fz;
                    dN[n][2] = fx * fy * 0.5 * sz;
#pragma unroll
                    for (int d = 0; d < 3; d++)
#pragma unroll
                        for (int k = 0; k < 3; k++) J[d * 3 + k] += x[n][d] * dN[n][k];
                }
                const double det = J[0] * (J[4] * J[8] - J[7] * J[5]) + J[3] * (J[7] * J[2] - J[1] * J[8]) +
                                   J[6] * (J[1] * J[5] - J[4] * J[2]);
                const double id = 1.0 / det;
                // (J^-1)[k][d] = d xi_k / d x_d
                const double i00 = (J[4] * J[8] - J[5] * J[7]) * id, i01 = (J[2] * J[7] - J[1] * J[8]) * id,
                             i02 = (J[1] * J[5] - J[2] * J[4]) * id, i10 = (J[5] * J[6] - J[3] * J[8]) * id,
                             i11 = (J[0] * J[8] - J[2] * J[6]) * id, i12 = (J[2] * J[3] - J[0] * J[5]) * id,
                             i20 = (J[3] * J[7] - J[4] * J[6]) * id, i21 = (J[1] * J[6] - J[0] * J[7]) * id,
                             i22 = (J[0] * J[4] - J[1] * J[3]) * id;
#pragma unroll
                for (int n = 0; n < 8; n++) {
                    g[s][q][n][0] = dN[n][0] * i00 + dN[n][1] * i10 + dN[n][2] * i20;
                    g[s][q][n][1] = dN[n][0] * i01 + dN[n][1] * i11 + dN[n][2] * i21;
                    g[s][q][n][2] = dN[n][0] * i02 + dN[n][1] * i12 + dN[n][2] * i22;
                }
                cq[s][q] = det;
            }
        }
        __syncthreads();
        // ---- blocks: lane = (s, b)
        const int s = lane >> 3, b = lane & 7;
        double K[9];
        int k = -1;
        if (s < ns) {
            const int a = sla[s];
            const int64_t e = selem[s];
            double P[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, S = 0.0;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const double c = cq[s][q];
                const double ga0 = c * g[s][q][a][0], ga1 = c * g[s][q][a][1], ga2 = c * g[s][q][a][2];
                const double gb0 = g[s][q][b][0], gb1 = g[s][q][b][1], gb2 = g[s][q][b][2];
                P[0] += ga0 * gb0; P[1] += ga0 * gb1; P[2] += ga0 * gb2;
                P[3] += ga1 * gb0; P[4] += ga1 * gb1; P[5] += ga1 * gb2;
                P[6] += ga2 * gb0; P[7] += ga2 * gb1; P[8] += ga2 * gb2;
                S += ga0 * gb0 + ga1 * gb1 + ga2 * gb2;
            }
            const double Ee = A.E0 * pow(A.param[e], A.penalty);
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++)
                    K[i * 3 + j] = Ee * (lam * P[i * 3 + j] + mu * P[j * 3 + i] + (i == j ? mu * S : 0.0));
            const int32_t rb = A.econn[e * 8 + b];
            if (A.X) {
                const double u0 = A.X[int64_t(rb) * 3], u1 = A.X[int64_t(rb) * 3 + 1], u2 = A.X[int64_t(rb) * 3 + 2];
                rv0 += K[0] * u0 + K[1] * u1 + K[2] * u2;
                rv1 += K[3] * u0 + K[4] * u1 + K[5] * u2;
                rv2 += K[6] * u0 + K[7] * u1 + K[8] * u2;
            }
            // libMesh constraint convention on the element block
            if (A.con) {
#pragma unroll
                for (int v = 0; v < 3; v++)
#pragma unroll
                    for (int w = 0; w < 3; w++)
                        if (A.con[r * 3 + v] || A.con[int64_t(rb) * 3 + w])
                            K[v * 3 + w] = (rb == int32_t(r) && v == w) ? 1.0 : 0.0;
            }
            // position of the neighbour in the row (sorted block columns)
            int lo = 0, hi = deg - 1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (bcol[mid] < rb) lo = mid + 1;
                else hi = mid;
            }
            k = lo;
        }
        // add the blocks into the row in slot order (within a slot the 8
        // blocks are distinct)
        for (int t = 0; t < ns; t++) {
            if (s == t)
#pragma unroll
                for (int i = 0; i < 9; i++) acc[k * 9 + i] += K[i];
            __syncthreads();
        }
    }
    __syncthreads();   // the accumulator is complete (also for a node without elements)
    // residual of row r: fixed-order reduction over the wave
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        rv0 += __shfl_xor(rv0, m, 64);
        rv1 += __shfl_xor(rv1, m, 64);
        rv2 += __shfl_xor(rv2, m, 64);
    }
    if (A.R && lane < 3) {
        const int64_t i = r * 3 + lane;
        double v = lane == 0 ? rv0 : lane == 1 ? rv1 : rv2;
        if (A.F && !(A.con && A.con[i])) v -= A.F[i];
        A.R[i] = (A.con && A.con[i]) ? 0.0 : v;
    }
    // the row's values: (v, block, w) order, one contiguous chunk
    const int64_t start = A.row_ptr[r * 3];
    const int n3 = 3 * deg;
    for (int t = lane; t < 3 * n3; t += 64) {
        const int v = t / n3, rem = t - v * n3, kb = rem / 3, w = rem - kb * 3;
        A.val[start + t] = acc[kb * 9 + v * 3 + w];
    }
    if (A.dinv && lane < 3) {
        int kd = 0;
        while (kd < deg && bcol[kd] != int32_t(r)) kd++;
        A.dinv[r * 3 + lane] = kd < deg ? 1.0 / acc[kd * 9 + lane * 4] : 0.0;
    }
}

}  // namespace

int hex8_row_assembly_max_blocks() { return kRowMaxBlocks; }

void launch_hex8_row_assemble(const Hex8RowArgs& a, hipStream_t s) {
    if (a.n_nodes <= 0) return;
    hipLaunchKernelGGL(k_hex8_row_assemble, dim3(unsigned(a.n_nodes)), dim3(64), 0, s, a);
}

}  // namespace mastg
