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
// When the mesh's elements fall into a few congruence classes (the same node
// coordinates relative to the first node), K_e = E_e Khat_class with Khat the
// class representative's matrix at E = 1, and a leaner kernel (no geometry,
// no staged blocks) assembles the rows: k_hex8_row_assemble_class below.
// NonlinearImplicitAssembly::residual_and_jacobian (src/base/
// nonlinear_implicit_assembly.cpp:111-186) adds every element block into the
// rows of its first node; here the owner of a row (one work-group = one wave
// per node) forms exactly the element blocks of that row -- each (element,
// a, b) block once in the whole assembly -- and sums them in a fixed order (the
// node's incidence order), so the assembly is deterministic without atomics.
//
// Per wave and pass over up to 8 incident elements (slots s): for the general
// form, lanes (s, q) evaluate the geometry of slot s at quadrature point q
// (Jacobian, its inverse, the 8 shape gradients) into LDS; then lanes (s, b)
// form block (a_s, b) of slot s, its residual contribution K_ab u_b (before
// constraints, as the reference's element vector) and its constraint
// treatment (zero constrained rows and columns, unit diagonal per element),
// and stage it in LDS with the position of neighbour b in the row.  Each lane
// then adds, for the row values it owns, the staged blocks of the slots that
// touch that neighbour, in slot order, into registers.  The row's values
// (rows r*3..r*3+2 of node r: one chunk ordered (v, block, w)) are written
// once, coalesced, by their owners.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_util.h"
#include "kernels.h"

#ifndef MASTG_ROW_GRID
#define MASTG_ROW_GRID (int64_t(32768))
#endif

namespace mastg {

namespace {

constexpr int kRowMaxBlocks = 64;   // neighbour blocks per row (9 row values per lane)
constexpr int kRowSmallBlocks = 28;  // rows of up to 28 blocks: 4 values per lane, fewer registers

__device__ __forceinline__ double ra_gp(int i) { return i ? 0.57735026918962576451 : -0.57735026918962576451; }

// a value that is the same in every lane, moved to scalar registers
__device__ __forceinline__ int uniform_i(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ double uniform_d(double x) {
    const uint64_t u = uint64_t(__double_as_longlong(x));
    const uint32_t lo = uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(u))));
    const uint32_t hi = uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(u >> 32))));
    return __longlong_as_double((long long)((uint64_t(hi) << 32) | lo));
}

template <int NV>
__global__ __launch_bounds__(64) void k_hex8_row_assemble(Hex8RowArgs A) {
    __shared__ double g[8][8][8][3];                  // [slot][q][node][d] shape gradients
    __shared__ double cq[8][8];                       // [slot][q] JxW
    __shared__ double kst[64 * 9];                    // [slot * 8 + b][v * 3 + w] staged element blocks
    __shared__ int8_t cont[NV * 64 / 9 * 8];          // [block][slot] local node b, or -1
    __shared__ int32_t bcol[NV * 64 / 9];
    __shared__ int32_t selem[8], sla[8];
    __shared__ int sdiag;
    const int lane = threadIdx.x;
    const int64_t r = xcd_block(blockIdx.x, gridDim.x);   // consecutive nodes on one XCD (shared L2)
    const int64_t b0 = A.blk_ptr[r];
    const int deg = int(A.blk_ptr[r + 1] - b0);
    const int32_t i0 = A.ninc_ptr[r], ne = A.ninc_ptr[r + 1] - i0;
    if (lane == 0) sdiag = -1;
    __syncthreads();
    if (lane < deg) {
        const int32_t cb = A.blk_col[b0 + lane];
        bcol[lane] = cb;
        if (cb == int32_t(r)) sdiag = lane;
    }
    __syncthreads();   // bcol and sdiag are read below (also by a node without elements)
    // the row values this lane owns: t = lane + 64 j in the chunk's (v, block,
    // w) order, decoded once into block * 16 + v * 3 + w
    const int n3 = max(3 * deg, 1), nval = 9 * deg;
    double acc[NV];
    int eix[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) {
        const int t = lane + 64 * j;
        const int v = t / n3, rem = t - v * n3, kb = rem / 3, w = rem - kb * 3;
        eix[j] = kb * 16 + v * 3 + w;
        acc[j] = 0.0;
    }
    double rv0 = 0.0, rv1 = 0.0, rv2 = 0.0;   // residual of row r from this lane's blocks
    const double lam = A.lam, mu = A.mu;
    for (int c0 = 0; c0 < ne; c0 += 8) {
        const int ns = min(8, ne - c0);
        if (lane < ns) {
            const int32_t inc = A.ninc[i0 + c0 + lane];
            selem[lane] = inc >> 3;
            sla[lane] = inc & 7;
        }
        for (int t = lane; t < deg * 8; t += 64) cont[t] = int8_t(-1);
        __syncthreads();
        {
            // ---- geometry: lane = (s, q)
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
            __syncthreads();
        }
        // ---- blocks: lane = (s, b)
        const int s = lane >> 3, b = lane & 7;
        if (s < ns) {
            const int a = sla[s];
            const int64_t e = selem[s];
            const int32_t rb = A.econn[e * 8 + b];
            const double Ee = A.sE[e];
            double K[9];
            {
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
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++)
                        K[i * 3 + j] = Ee * (lam * P[i * 3 + j] + mu * P[j * 3 + i] + (i == j ? mu * S : 0.0));
            }
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
#pragma unroll
            for (int i = 0; i < 9; i++) kst[lane * 9 + i] = K[i];
            cont[lo * 8 + s] = int8_t(b);   // an element's 8 nodes are distinct: one b per (block, slot)
        }
        __syncthreads();
        // ---- the owned row values add the staged blocks in slot order
#pragma unroll
        for (int j = 0; j < NV; j++) {
            if (lane + 64 * j < nval) {
                const int8_t* cb = cont + (eix[j] >> 4) * 8;
                const double* kj = kst + (eix[j] & 15);
                double sum = acc[j];
                for (int t = 0; t < ns; t++) {
                    const int bb = cb[t];
                    if (bb >= 0) sum += kj[(t * 8 + bb) * 9];
                }
                acc[j] = sum;
            }
        }
        __syncthreads();   // the next pass rewrites the staged blocks and slots
    }
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
    // the row's values, written by their owners (consecutive lanes, consecutive addresses)
    const int64_t start = A.row_ptr[r * 3];
    const int kd = sdiag;
#pragma unroll
    for (int j = 0; j < NV; j++) {
        if (lane + 64 * j < nval) {
            A.val[start + lane + 64 * j] = acc[j];
            const int iw = eix[j] & 15;
            if (A.dinv && (eix[j] >> 4) == kd && iw % 4 == 0) A.dinv[r * 3 + iw / 4] = 1.0 / acc[j];
        }
    }
    if (A.dinv && kd < 0 && lane < 3) A.dinv[r * 3 + lane] = 0.0;
}

// The same assembly when the elements fall into congruence classes (K_e =
// sE[e] Khat_class): the constraint convention depends only on the (row,
// column) variable pair, not on the element, so each owned row value sums the
// unconstrained contributions sE[e] Khat[class][a_s][b][v][w] of the slots
// touching its neighbour (slot order), counts them, and is replaced once at
// the end by the convention's value (the count on the diagonal, 0 elsewhere)
// when its row or column variable is constrained -- the same value, bit for
// bit, as masking each element block before the sum.  The residual row is
// the unconstrained row times the state.  The class matrices are read from
// LDS: with one class the whole matrix, loaded once per work-group (coalesced);
// with several, per pass the 8 rows Khat[class_s][a_s][:] of the slots, loaded
// by the lanes (s, b) that place neighbour b of slot s in the row.  Each
// work-group runs over a contiguous range of nodes, so the one-class load is
// amortised.
// NP = 2: 5 waves per SIMD at 93 VGPRs without spills (the compiler's own
// choice is 99 VGPRs, 4 waves); NP = 3 spills under that bound and keeps its own
template <int NP, bool ONE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(NP == 2 ? 5 : 1, 8))) void k_hex8_row_assemble_class(
    Hex8RowArgs A) {
    // a lane owns NP (row variable v, neighbour block k) pairs p = v deg + k,
    // i.e. the 3 consecutive row values 3 p .. 3 p + 2 (w = 0, 1, 2)
    constexpr int MAXB = NP * 64 / 3;
    __shared__ double kh[576];                   // ONE: Khat[a][b][v][w]; else [slot][b][v][w] rows
    __shared__ __align__(8) int8_t cont[MAXB * 8];  // [block][slot] local node b, or -1
    __shared__ int32_t bcol[MAXB];
    __shared__ double sEe[8];
    __shared__ int32_t sko[8];                   // kh offset of slot s's row: a_s * 72 (ONE) or s * 72
    __shared__ int sdiag;
    const int lane = threadIdx.x;
    if (ONE)
        for (int t = lane; t < 576; t += 64) kh[t] = A.Khat[t];
    // a contiguous run of nodes per work-group, the runs of one XCD adjacent
    // (work-groups are dealt to the 8 XCDs in turn): neighbouring rows gather
    // the same state and flags through the same L2
    const int64_t nb = gridDim.x, per = (A.n_nodes + nb - 1) / nb;
    const int64_t rbeg = xcd_block(blockIdx.x, nb) * per, rend = min(A.n_nodes, rbeg + per);
    for (int64_t r = rbeg; r < rend; r++) {
        const int64_t b0 = A.blk_ptr[r];
        const int deg = int(A.blk_ptr[r + 1] - b0);
        const int32_t i0 = A.ninc_ptr[r], ne = A.ninc_ptr[r + 1] - i0;
        if (lane == 0) sdiag = -1;
        __syncthreads();
        if (lane < deg) {
            const int32_t cb = A.blk_col[b0 + lane];
            bcol[lane] = cb;
            if (cb == int32_t(r)) sdiag = lane;
        }
        const int npair = 3 * deg, dg = max(deg, 1);
        double acc[NP][3];
        int pk[NP], pv[NP], cnt[NP];
#pragma unroll
        for (int j = 0; j < NP; j++) {
            const int p = lane + 64 * j;
            pv[j] = p / dg;
            pk[j] = p - pv[j] * dg;
            cnt[j] = 0;
            acc[j][0] = acc[j][1] = acc[j][2] = 0.0;
        }
        __syncthreads();   // bcol and sdiag
        // the neighbours' state and constraint flags are loaded now, in flight
        // while the element slots are placed
        double xs[NP][3];
        uint8_t cw[NP];
#pragma unroll
        for (int j = 0; j < NP; j++) {
            xs[j][0] = xs[j][1] = xs[j][2] = 0.0;
            cw[j] = 0;
            if (lane + 64 * j < npair) {
                const int64_t col = int64_t(bcol[pk[j]]) * 3;
                if (A.X) {
                    xs[j][0] = A.X[col];
                    xs[j][1] = A.X[col + 1];
                    xs[j][2] = A.X[col + 2];
                }
                if (A.con) cw[j] = uint8_t(A.con[col] | (A.con[col + 1] << 1) | (A.con[col + 2] << 2));
            }
        }
        for (int c0 = 0; c0 < ne; c0 += 8) {
            const int ns = min(8, ne - c0);
            for (int t = lane; t < deg * 8; t += 64) cont[t] = int8_t(-1);
            __syncthreads();   // bcol (first pass) and the cleared slots; the previous pass's readers are done
            {
                const int s = lane >> 3, b = lane & 7;
                if (s < ns) {
                    const int32_t inc = A.ninc[i0 + c0 + s];
                    const int64_t e = inc >> 3;
                    const int a = inc & 7;
                    const int32_t rb = A.econn[e * 8 + b];
                    if (b == 0) {
                        sEe[s] = A.sE[e];
                        sko[s] = (ONE ? a : s) * 72;
                    }
                    if (!ONE) {
                        const double* __restrict__ src = A.Khat + (int64_t(A.eclass[e]) * 64 + a * 8 + b) * 9;
#pragma unroll
                        for (int i = 0; i < 9; i++) kh[(s * 8 + b) * 9 + i] = src[i];
                    }
                    int lo = 0, hi = deg - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (bcol[mid] < rb) lo = mid + 1;
                        else hi = mid;
                    }
                    cont[lo * 8 + s] = int8_t(b);   // an element's 8 nodes are distinct: one b per (block, slot)
                }
            }
            __syncthreads();
            // slot scales and row offsets are the same for every lane (scalar registers)
            double se[8];
            int so[8];
#pragma unroll
            for (int t = 0; t < 8; t++) {
                se[t] = uniform_d(t < ns ? sEe[t] : 0.0);
                so[t] = uniform_i(t < ns ? sko[t] : 0);
            }
#pragma unroll
            for (int j = 0; j < NP; j++) {
                if (lane + 64 * j < npair) {
                    const uint64_t row = *reinterpret_cast<const uint64_t*>(cont + pk[j] * 8);
                    const int vo = pv[j] * 3;
                    double s0 = acc[j][0], s1 = acc[j][1], s2 = acc[j][2];
                    int n = cnt[j];
                    // slot order, branch-free: a slot that does not touch the
                    // neighbour reads an in-bounds row with the scale 0 (adds an
                    // exact zero)
#pragma unroll
                    for (int t = 0; t < 8; t++) {
                        const int bb = int(int8_t((row >> (8 * t)) & 0xff));
                        const bool on = t < ns && bb >= 0;
                        const double* k3 = kh + so[t] + (on ? bb : 0) * 9 + vo;
                        const double m = on ? se[t] : 0.0;
                        s0 = fma(m, k3[0], s0);
                        s1 = fma(m, k3[1], s1);
                        s2 = fma(m, k3[2], s2);
                        n += on ? 1 : 0;
                    }
                    acc[j][0] = s0;
                    acc[j][1] = s1;
                    acc[j][2] = s2;
                    cnt[j] = n;
                }
            }
            __syncthreads();   // the next pass rewrites the slots
        }
        const int kd = sdiag;
        double rv0 = 0.0, rv1 = 0.0, rv2 = 0.0;
        const int64_t start = A.row_ptr[r * 3];
        const uint8_t cr = A.con ? uint8_t(A.con[r * 3] | (A.con[r * 3 + 1] << 1) | (A.con[r * 3 + 2] << 2)) : 0;
#pragma unroll
        for (int j = 0; j < NP; j++) {
            const int p = lane + 64 * j;
            if (p < npair) {
                const int k = pk[j], v = pv[j];
                if (A.X) {
                    const double q = acc[j][0] * xs[j][0] + acc[j][1] * xs[j][1] + acc[j][2] * xs[j][2];
                    rv0 += v == 0 ? q : 0.0;
                    rv1 += v == 1 ? q : 0.0;
                    rv2 += v == 2 ? q : 0.0;
                }
#pragma unroll
                for (int w = 0; w < 3; w++) {
                    double val = acc[j][w];
                    if (A.con && (((cr >> v) & 1) || ((cw[j] >> w) & 1))) val = (k == kd && v == w) ? double(cnt[j]) : 0.0;
                    A.val[start + 3 * p + w] = val;
                    if (A.dinv && k == kd && v == w) A.dinv[r * 3 + v] = 1.0 / val;
                }
            }
        }
        if (A.dinv && kd < 0 && lane < 3) A.dinv[r * 3 + lane] = 0.0;
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
        __syncthreads();   // the shared arrays are rewritten for the next node
    }
}

// ---- the class assembly over a row plan ----
// The row plan row_pos[r * 64 + s * 8 + b] holds, for node r, the position in
// its sorted block row of local node b of the element in incidence slot s (-1
// past the node's incidences).  It depends on the mesh only and is built once
// (k_hex8_row_plan: the searches the kernel above does per pass).  With it a
// node's data is two loads deep -- its block columns, incidences and plan
// row, then the neighbours' state and constraint flags and the slots' scales
// -- and k_hex8_row_assemble_plan keeps both levels ahead of the node it
// assembles: the gathers of the next node and the index loads of the node
// after it are in flight while a node is summed and written.  The work-group
// runs over at most 32 consecutive nodes, whose pointers a lane each loads
// once up front.  The sums, their order and the constraint treatment are
// those of k_hex8_row_assemble_class.

__global__ __launch_bounds__(256) void k_hex8_row_plan(Hex8RowArgs A, int8_t* pos, uint8_t* cmask) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= A.n_nodes * 64) return;
    const int64_t r = t >> 6;
    const int s = int(t >> 3) & 7, b = int(t) & 7;
    if (s == 0 && b == 0)
        cmask[r] = A.con ? uint8_t(A.con[r * 3] | (A.con[r * 3 + 1] << 1) | (A.con[r * 3 + 2] << 2)) : 0;
    const int32_t i0 = A.ninc_ptr[r], ne = A.ninc_ptr[r + 1] - i0;
    int p = -1;
    if (s < ne) {
        const int32_t rb = A.econn[int64_t(A.ninc[i0 + s] >> 3) * 8 + b];
        const int64_t b0 = A.blk_ptr[r];
        int64_t lo = b0, hi = A.blk_ptr[r + 1] - 1;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (A.blk_col[mid] < rb) lo = mid + 1;
            else hi = mid;
        }
        p = int(lo - b0);
    }
    pos[t] = int8_t(p);
}

__device__ __forceinline__ int lane_i(int x, int j) { return __builtin_amdgcn_readlane(x, j); }
__device__ __forceinline__ int64_t lane_i64(int64_t x, int j) {
    const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(uint64_t(x))), j));
    const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(uint64_t(x) >> 32)), j));
    return int64_t((uint64_t(hi) << 32) | lo);
}
__device__ __forceinline__ double lane_d(double x, int j) {
    return __longlong_as_double((long long)lane_i64(int64_t(__double_as_longlong(x)), j));
}

// an empty statement that takes v in a vector register: the load that
// produced it is complete here
__device__ __forceinline__ void ready(double v) { asm volatile("" ::"v"(v)); }
__device__ __forceinline__ void ready(int v) { asm volatile("" ::"v"(v)); }

constexpr int kPlanRun = 32;   // nodes per work-group at most (a lane holds each one's pointers)

template <int NP, bool ONE>
__global__ __launch_bounds__(64) void k_hex8_row_assemble_plan(Hex8RowArgs A) {
    constexpr int MAXB = NP * 64 / 3;
    constexpr int kZero = 64;   // kh row of zeros: the slots that do not touch a neighbour
    __shared__ double kh[65 * 9];                    // ONE: Khat[a][b][v][w]; else [slot][b][v][w]; + zeros
    __shared__ __align__(8) uint8_t cont[MAXB * 8];  // [block][slot] kh row of the slot's block, or kZero
    // the run's residual and inverse diagonal, written out once at its end
    // (node by node, 24-byte pieces of lines that leave L2 in between)
    __shared__ double sres[kPlanRun * 3], sdinv[kPlanRun * 3];
    const int lane = threadIdx.x;
    for (int t = lane; t < 65 * 9; t += 64) kh[t] = (ONE && t < 576) ? A.Khat[t] : 0.0;
    for (int t = lane; t < MAXB * 8; t += 64) cont[t] = kZero;
    const int64_t nb = gridDim.x, per = (A.n_nodes + nb - 1) / nb;
    const int64_t rbeg = xcd_block(blockIdx.x, nb) * per;
    const int cnt = int(min(A.n_nodes, rbeg + per) - rbeg);
    if (cnt <= 0) return;
    // the run's pointers: lane i holds node rbeg + i's
    int64_t pb = 0, prow = 0;
    int32_t pn = 0;
    int pcr = 0;
    if (lane <= cnt) {
        pb = A.blk_ptr[rbeg + lane];
        pn = A.ninc_ptr[rbeg + lane];
    }
    if (lane < cnt) {
        prow = A.row_ptr[(rbeg + lane) * 3];
        pcr = A.cmask[rbeg + lane];
    }
    // level 1 of node jj: block column (lane k), incidence of slot lane / 8,
    // plan entry (lane = slot * 8 + b), load row value (lanes 0..2)
    auto load1 = [&](int jj, int& bc, int& inc, int& ps, double& f) {
        const int64_t b0 = lane_i64(pb, jj);
        const int deg = int(lane_i64(pb, jj + 1) - b0);
        const int i0 = lane_i(pn, jj), ne = lane_i(pn, jj + 1) - i0;
        bc = lane < deg ? A.blk_col[b0 + lane] : -1;
        inc = (lane >> 3) < ne ? A.ninc[i0 + (lane >> 3)] : -1;
        ps = A.row_pos[(rbeg + jj) * 64 + lane];
        f = (A.F && lane < 3) ? A.F[(rbeg + jj) * 3 + lane] : 0.0;
    };
    // level 2: the neighbour's state and flags (lane k), the slot's scale and class
    auto load2 = [&](int bc, int inc, double* x, int& cw, double& se, int& cl) {
        x[0] = x[1] = x[2] = 0.0;
        cw = 0;
        se = 0.0;
        cl = 0;
        if (bc >= 0) {
            const int64_t col = int64_t(bc) * 3;
            if (A.X) {
                x[0] = A.X[col];
                x[1] = A.X[col + 1];
                x[2] = A.X[col + 2];
            }
            cw = A.cmask[bc];
        }
        if (inc >= 0) {
            se = A.sE[inc >> 3];
            if (!ONE) cl = A.eclass[inc >> 3];
        }
    };
    int bcA, incA, psA, bcB, incB, psB, cwA, clA;
    double fA, fB, xA[3], seA;
    load1(0, bcA, incA, psA, fA);
    load2(bcA, incA, xA, cwA, seA, clA);
    load1(min(1, cnt - 1), bcB, incB, psB, fB);
    for (int j = 0; j < cnt; j++) {
        // in flight while node j is assembled: node j + 1's gathers, node j + 2's indices
        double xB[3], seB;
        int cwB, clB, bcC, incC, psC;
        double fC;
        load2(bcB, incB, xB, cwB, seB, clB);
        load1(min(j + 2, cnt - 1), bcC, incC, psC, fC);

        const int64_t r = rbeg + j;
        const int64_t b0 = lane_i64(pb, j);
        const int deg = int(lane_i64(pb, j + 1) - b0);
        const int ne = lane_i(pn, j + 1) - lane_i(pn, j);
        const int64_t start = lane_i64(prow, j);
        const int cr = lane_i(pcr, j);
        const uint64_t dm = __ballot(bcA == int32_t(r));
        const int kd = dm ? __builtin_ctzll(dm) : -1;
        const int sl = lane >> 3, b = lane & 7, a = incA & 7;
        if (incA >= 0) {
            // an element's 8 nodes are distinct: one entry per (block, slot)
            cont[psA * 8 + sl] = uint8_t((ONE ? a : sl) * 8 + b);
            if (!ONE) {
                const double* __restrict__ src = A.Khat + (int64_t(clA) * 64 + a * 8 + b) * 9;
#pragma unroll
                for (int i = 0; i < 9; i++) kh[(sl * 8 + b) * 9 + i] = src[i];
            }
        }
        __syncthreads();
        // slot scales (scalar registers; 0 past the node's incidences)
        double se[8];
#pragma unroll
        for (int t = 0; t < 8; t++) se[t] = lane_d(seA, t * 8);
        const int npair = 3 * deg, dg = max(deg, 1);
        double acc[NP][3];
        int pk[NP], pv[NP];
#pragma unroll
        for (int jp = 0; jp < NP; jp++) {
            const int p = lane + 64 * jp;
            pv[jp] = p / dg;
            pk[jp] = p - pv[jp] * dg;
            acc[jp][0] = acc[jp][1] = acc[jp][2] = 0.0;
            if (p < npair) {
                // slot order; a slot that does not touch the neighbour adds
                // its scale times the zero row (an exact zero)
                const uint64_t row = *reinterpret_cast<const uint64_t*>(cont + pk[jp] * 8);
                const int vo = pv[jp] * 3;
                double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
                for (int t = 0; t < 8; t++) {
                    const double* k3 = kh + int((row >> (8 * t)) & 0xff) * 9 + vo;
                    s0 = fma(se[t], k3[0], s0);
                    s1 = fma(se[t], k3[1], s1);
                    s2 = fma(se[t], k3[2], s2);
                }
                acc[jp][0] = s0;
                acc[jp][1] = s1;
                acc[jp][2] = s2;
            }
        }
        __syncthreads();   // the slots are read; clear this node's entries for the next
        if (incA >= 0) cont[psA * 8 + sl] = kZero;
        double rv0 = 0.0, rv1 = 0.0, rv2 = 0.0, vals[NP][3];
#pragma unroll
        for (int jp = 0; jp < NP; jp++) {
            const int p = lane + 64 * jp;
            const int k = pk[jp], v = pv[jp];
            // the neighbour's state and flags, from the lane that gathered them
            const double x0 = __shfl(xA[0], k, 64), x1 = __shfl(xA[1], k, 64), x2 = __shfl(xA[2], k, 64);
            const int cw = __shfl(cwA, k, 64);
            if (p < npair && A.X) {
                const double q = acc[jp][0] * x0 + acc[jp][1] * x1 + acc[jp][2] * x2;
                rv0 += v == 0 ? q : 0.0;
                rv1 += v == 1 ? q : 0.0;
                rv2 += v == 2 ? q : 0.0;
            }
            // constrained row or column: 0, and on the diagonal the number
            // of elements that put a unit there (all ne touch the node itself)
#pragma unroll
            for (int w = 0; w < 3; w++) {
                double val = acc[jp][w];
                if (A.con && (((cr >> v) & 1) || ((cw >> w) & 1))) val = (k == kd && v == w) ? double(ne) : 0.0;
                vals[jp][w] = val;
            }
        }
#pragma unroll
        for (int m = 1; m < 64; m <<= 1) {
            rv0 += __shfl_xor(rv0, m, 64);
            rv1 += __shfl_xor(rv1, m, 64);
            rv2 += __shfl_xor(rv2, m, 64);
        }
        const double fr = fA;
        // the next node's data is waited for here, before this node's stores
        // are issued: the memory counter is in order, so a wait after them
        // (where the compiler moves the loop's register copies) would wait for them
        ready(xB[0]); ready(xB[1]); ready(xB[2]); ready(seB); ready(cwB); ready(clB);
        ready(bcC); ready(incC); ready(psC); ready(fC);
        bcA = bcB; incA = incB; psA = psB; fA = fB;
        xA[0] = xB[0]; xA[1] = xB[1]; xA[2] = xB[2];
        cwA = cwB; seA = seB; clA = clB;
        bcB = bcC; incB = incC; psB = psC; fB = fC;
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int jp = 0; jp < NP; jp++) {
            const int p = lane + 64 * jp;
            const int k = pk[jp], v = pv[jp];
            if (p < npair) {
#pragma unroll
                for (int w = 0; w < 3; w++) A.val[start + 3 * p + w] = vals[jp][w];
                if (k == kd) {
                    const double d = v == 0 ? vals[jp][0] : v == 1 ? vals[jp][1] : vals[jp][2];
                    sdinv[j * 3 + v] = 1.0 / d;
                }
            }
        }
        if (kd < 0 && lane < 3) sdinv[j * 3 + lane] = 0.0;
        if (lane < 3) {
            const bool ci = (cr >> lane) & 1;
            double v = lane == 0 ? rv0 : lane == 1 ? rv1 : rv2;
            if (A.F && !ci) v -= fr;
            sres[j * 3 + lane] = ci ? 0.0 : v;
        }
    }
    __syncthreads();
    for (int t = lane; t < cnt * 3; t += 64) {
        if (A.R) A.R[rbeg * 3 + t] = sres[t];
        if (A.dinv) A.dinv[rbeg * 3 + t] = sdinv[t];
    }
}

}  // namespace

int hex8_row_assembly_max_blocks() { return kRowMaxBlocks; }

void launch_hex8_row_assemble(const Hex8RowArgs& a, int max_blocks_per_row, hipStream_t s) {
    if (a.n_nodes <= 0) return;
    const dim3 g(unsigned(a.n_nodes));
    const bool small = max_blocks_per_row <= kRowSmallBlocks;
    if (a.Khat && a.eclass && a.row_pos) {
        // runs of at most kPlanRun consecutive nodes per work-group
        const dim3 gp(unsigned(std::max<int64_t>(std::min<int64_t>(a.n_nodes, MASTG_ROW_GRID),
                                                 (a.n_nodes + kPlanRun - 1) / kPlanRun)));
        const bool two = max_blocks_per_row <= 42;
        if (a.n_classes == 1) {
            if (two) hipLaunchKernelGGL((k_hex8_row_assemble_plan<2, true>), gp, dim3(64), 0, s, a);
            else hipLaunchKernelGGL((k_hex8_row_assemble_plan<3, true>), gp, dim3(64), 0, s, a);
        } else {
            if (two) hipLaunchKernelGGL((k_hex8_row_assemble_plan<2, false>), gp, dim3(64), 0, s, a);
            else hipLaunchKernelGGL((k_hex8_row_assemble_plan<3, false>), gp, dim3(64), 0, s, a);
        }
        return;
    }
    if (a.Khat && a.eclass) {
        const dim3 gc(unsigned(std::min<int64_t>(a.n_nodes, MASTG_ROW_GRID)));
        // pairs per lane: 2 up to 42 blocks per row, 3 up to 64
        const bool two = max_blocks_per_row <= 42;
        if (a.n_classes == 1) {
            if (two) hipLaunchKernelGGL((k_hex8_row_assemble_class<2, true>), gc, dim3(64), 0, s, a);
            else hipLaunchKernelGGL((k_hex8_row_assemble_class<3, true>), gc, dim3(64), 0, s, a);
        } else {
            if (two) hipLaunchKernelGGL((k_hex8_row_assemble_class<2, false>), gc, dim3(64), 0, s, a);
            else hipLaunchKernelGGL((k_hex8_row_assemble_class<3, false>), gc, dim3(64), 0, s, a);
        }
    } else {
        if (small) hipLaunchKernelGGL((k_hex8_row_assemble<4>), g, dim3(64), 0, s, a);
        else hipLaunchKernelGGL((k_hex8_row_assemble<9>), g, dim3(64), 0, s, a);
    }
}

void launch_hex8_row_plan(const Hex8RowArgs& a, int8_t* pos, uint8_t* cmask, hipStream_t s) {
    if (a.n_nodes <= 0) return;
    const int64_t n = a.n_nodes * 64;
    hipLaunchKernelGGL(k_hex8_row_plan, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, a, pos, cmask);
}

}  // namespace mastg
