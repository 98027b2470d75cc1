// Geometric multigrid for hex8 elasticity on lattice meshes: Galerkin coarse
// element matrices, Chebyshev/Jacobi smoothing fused into the level operators,
// trilinear restriction / prolongation, dense coarsest solve, and the
// preconditioned CG kernels around the V-cycle.
//
// Every reduction is a fixed-order sum (no atomics), so a solve is bitwise
// reproducible for a given launch configuration.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "kernels.h"
#include "device_util.h"

namespace mastg {

// libMesh hex8 local node -> cell corner offset
__constant__ int8_t kMGCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                       {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

// tri/bilinear weight of coarse corner A at fine local node a of child cell c
// (c bit d = the child's offset along dimension d inside the coarse cell)
template <int DIM>
__device__ __forceinline__ double mg_q(int c, int a, int A) {
    double w = 1.0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const int t = ((c >> d) & 1) + kMGCorner[a][d];   // 0, 1, 2 fine cells
        w *= kMGCorner[A][d] ? 0.5 * t : 1.0 - 0.5 * t;
    }
    return w;
}

// ---- Galerkin coarse element matrices: K_E = sum_c Q_c^T K_c Q_c ---------
// One wave per coarse element, 4 per block; lane (A, B) < NPE^2 owns one
// NAV x NAV block.  Fine blocks are masked (the rows and columns of
// constrained fine DOFs removed, as P has no row for them), the result is made
// exactly symmetric and gets the libMesh constraint convention for
// constrained coarse DOFs (rows / columns zero, unit diagonal).
template <int DIM, int NAV>
__global__ __launch_bounds__(256) void k_mg_galerkin(MGGalerkinArgs G) {
    constexpr int NPE = 1 << DIM, NCH = 1 << DIM, NB = NAV * NAV, SZ = NPE * NPE * NB;
    __shared__ double sK[4][SZ];
    __shared__ double sT[4][SZ];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const bool lane_on = lane < NPE * NPE;
    const int ln = lane_on ? lane : 0;
    const int hi = ln / NPE, lo = ln % NPE;
    // coarse element of this wave: the listed ones, or all of them
    const int64_t w = int64_t(blockIdx.x) * 4 + wave;
    const bool valid = G.elems ? w < G.n_elems_list : w < G.n_coarse;
    const int64_t E = !valid ? 0 : (G.elems ? int64_t(G.elems[w]) : w);
    // 32-bit index arithmetic (coarse element counts are below 2^31)
    const uint32_t Eu = uint32_t(E), cnx = uint32_t(G.cnx), Eq = Eu / cnx;
    const int64_t I = int64_t(Eu - Eq * cnx), J = int64_t(Eq % uint32_t(G.cny)), K = int64_t(Eq / uint32_t(G.cny));
    double acc[NB];
#pragma unroll
    for (int i = 0; i < NB; i++) acc[i] = 0.0;
    for (int c = 0; c < NCH; c++) {
        const int64_t fi = 2 * I + (c & 1), fj = 2 * J + ((c >> 1) & 1), fk = DIM == 3 ? 2 * K + ((c >> 2) & 1) : 0;
        const int64_t cell = fi + int64_t(G.fnx) * (fj + int64_t(G.fny) * fk);
        const int64_t fe = G.f_elem_of_cell ? int64_t(G.f_elem_of_cell[cell]) : cell;
        if (lane_on) {   // masked fine block (a, b) = (hi, lo)
            const int32_t ra = G.f_econn[fe * NPE + hi], rb = G.f_econn[fe * NPE + lo];
            const double sc = G.Khat ? G.sE_cell[cell] : 1.0;
            const double* kb = G.Khat ? G.Khat + ln * NB : G.Kf + (fe * NPE * NPE + ln) * NB;
#pragma unroll
            for (int v = 0; v < NAV; v++)
#pragma unroll
                for (int w = 0; w < NAV; w++) {
                    double x = sc * kb[v * NAV + w];
                    if (G.f_con[int64_t(ra) * NAV + v] || G.f_con[int64_t(rb) * NAV + w]) x = 0.0;
                    sK[wave][ln * NB + v * NAV + w] = x;
                }
        }
        __syncthreads();
        if (lane_on) {   // T[a][B] = sum_b K[a][b] Q_c[b][B], lane = (a, B)
            double t[NB];
#pragma unroll
            for (int i = 0; i < NB; i++) t[i] = 0.0;
            for (int b = 0; b < NPE; b++) {
                const double q = mg_q<DIM>(c, b, lo);
                if (q != 0.0) {
                    const double* kb = &sK[wave][(hi * NPE + b) * NB];
#pragma unroll
                    for (int i = 0; i < NB; i++) t[i] += kb[i] * q;
                }
            }
#pragma unroll
            for (int i = 0; i < NB; i++) sT[wave][ln * NB + i] = t[i];
        }
        __syncthreads();
        if (lane_on)   // K_E[A][B] += sum_a Q_c[a][A] T[a][B], lane = (A, B)
            for (int a = 0; a < NPE; a++) {
                const double q = mg_q<DIM>(c, a, hi);
                if (q != 0.0) {
                    const double* tb = &sT[wave][(a * NPE + lo) * NB];
#pragma unroll
                    for (int i = 0; i < NB; i++) acc[i] += q * tb[i];
                }
            }
        __syncthreads();
    }
    // exact symmetry: lanes below the diagonal take the transpose of their partner
    if (lane_on)
#pragma unroll
        for (int i = 0; i < NB; i++) sK[wave][ln * NB + i] = acc[i];
    __syncthreads();
    if (!valid || !lane_on) return;
    double out[NB];
    if (hi > lo) {
        const double* tb = &sK[wave][(lo * NPE + hi) * NB];
#pragma unroll
        for (int v = 0; v < NAV; v++)
#pragma unroll
            for (int w = 0; w < NAV; w++) out[v * NAV + w] = tb[w * NAV + v];
    } else if (hi == lo) {
#pragma unroll
        for (int v = 0; v < NAV; v++)
#pragma unroll
            for (int w = 0; w < NAV; w++) out[v * NAV + w] = 0.5 * (acc[v * NAV + w] + acc[w * NAV + v]);
    } else {
#pragma unroll
        for (int i = 0; i < NB; i++) out[i] = acc[i];
    }
    const int32_t cA = G.c_econn[E * NPE + hi], cB = G.c_econn[E * NPE + lo];
    double* o = G.Kc + (E * NPE * NPE + ln) * NB;
#pragma unroll
    for (int v = 0; v < NAV; v++)
#pragma unroll
        for (int w = 0; w < NAV; w++) {
            double x = out[v * NAV + w];
            if (G.c_con[int64_t(cA) * NAV + v] || G.c_con[int64_t(cB) * NAV + w]) x = (hi == lo && v == w) ? 1.0 : 0.0;
            o[v * NAV + w] = x;
        }
}

// Coarse elements of a one-class hex8 lattice with no constrained DOF among
// their 27 fine nodes: K_E = sum_c s_c Gc[c] (Gc exactly symmetric, so K_E is);
// one wave per coarse element, 9 of the 576 entries per lane, coalesced stores.
__global__ __launch_bounds__(256) void k_mg_galerkin_class(MGGalerkinArgs G) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t E = int64_t(blockIdx.x) * 4 + wave;
    if (E >= G.n_coarse || G.dirty[E]) return;
    const int64_t I = E % G.cnx, J = (E / G.cnx) % G.cny, K = E / (int64_t(G.cnx) * G.cny);
    double sc[8];
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const int64_t fi = 2 * I + (c & 1), fj = 2 * J + ((c >> 1) & 1), fk = 2 * K + (c >> 2);
        sc[c] = G.sE_cell[fi + int64_t(G.fnx) * (fj + int64_t(G.fny) * fk)];
    }
    double* __restrict__ o = G.Kc + E * 576;
#pragma unroll
    for (int t = 0; t < 9; t++) {
        const int i = lane + 64 * t;
        double acc = 0.0;
#pragma unroll
        for (int c = 0; c < 8; c++) acc = fma(sc[c], G.Gc[c * 576 + i], acc);
        o[i] = acc;
    }
}

// ---- row information and eigenvalue bounds ----------------------------------
__global__ __launch_bounds__(256) void k_mg_max(const double* __restrict__ part, int n, double* __restrict__ out) {
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) m = fmax(m, part[i]);
    block_max_store(m, out, 0);
}

// Chebyshev coefficients on [lam/ratio, lam] for every level
__global__ void k_mg_cheb(const double* __restrict__ lam, double* __restrict__ cheb, int n_levels, int k,
                          double ratio) {
    const int l = threadIdx.x;
    if (l >= n_levels) return;
    const double b = lam[l] > 0.0 ? lam[l] : 1.0, a = b / ratio;
    const double theta = 0.5 * (b + a), delta = 0.5 * (b - a), sigma = theta / delta;
    double* c = cheb + l * kMGChebStride;
    c[0] = 1.0 / theta;
    c[1] = 0.0;
    double rho = 1.0 / sigma;
    for (int i = 1; i < k; i++) {
        const double rn = 1.0 / (2.0 * sigma - rho);
        c[2 * i] = rn * rho;
        c[2 * i + 1] = 2.0 * rn / delta;
        rho = rn;
    }
}

// ---- smoothing ---------------------------------------------------------------
__device__ __forceinline__ void mg_epilogue(const MGLevelArgs& L, int mode, int64_t i, double y,
                                            const double* __restrict__ din, double* __restrict__ dout,
                                            double c1, double c2, double ith) {
    if (mode == MG_RESID_INIT) {
        const double rs = L.r[i] - y;
        L.res[i] = rs;
        dout[i] = L.dinv[i] * rs * ith;
    } else {
        const double dv = din[i];
        L.x[i] += dv;
        const double rs = L.res[i] - y;
        L.res[i] = rs;
        if (mode == MG_STEP) dout[i] = c1 * dv + c2 * L.dinv[i] * rs;
    }
}

__global__ __launch_bounds__(256) void k_mg_pre_init(MGLevelArgs L, const int* __restrict__ done) {
    if (done && *done) return;
    const double ith = L.cheb[0];
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < L.n_nodes * L.nav; i += int64_t(gridDim.x) * 256) {
        const double r = L.r[i];
        L.x[i] = 0.0;
        L.res[i] = r;
        L.da[i] = L.dinv[i] * r * ith;
    }
}

// the same start with block Jacobi (5 variables per node): da = ith Dinv_b r,
// a thread per DOF (row v of node n's inverse block), so that a wave's loads
// of the blocks and of r cover contiguous memory
__global__ __launch_bounds__(256) void k_mg_pre_init_blk5(MGLevelArgs L, const int* __restrict__ done) {
    if (done && *done) return;
    const double ith = L.cheb[0];
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < L.n_nodes * 5; i += int64_t(gridDim.x) * 256) {
        const int64_t n = i / 5;
        const double* rn = L.r + n * 5;
        const float* B = L.dinvb + i * 5;   // row v = i - 5 n of node n's block
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < 5; w++) t = fma(double(B[w]), rn[w], t);
        L.x[i] = 0.0;
        L.res[i] = rn[i - n * 5];
        L.da[i] = ith * t;
    }
}

// inverse of each node's 5 x 5 diagonal block (Gauss-Jordan without pivoting:
// the blocks of an SPD matrix are SPD; constrained variables keep their unit
// diagonal convention), from the FP64 node-block values (layout (v, k, w))
__global__ __launch_bounds__(256) void k_mg_block_dinv5(MGLevelArgs L, float* __restrict__ dinvb) {
    const int64_t n = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (n >= L.n_nodes) return;
    const int64_t k0 = L.blk_ptr[n];
    const int deg = int(L.blk_ptr[n + 1] - k0);
    int kd = 0;
    for (int k = 0; k < deg; k++)
        if (L.blk_col[k0 + k] == int32_t(n)) kd = k;
    const double* V = L.val + k0 * 25;
    double a[5][5];
#pragma unroll
    for (int v = 0; v < 5; v++)
#pragma unroll
        for (int w = 0; w < 5; w++) a[v][w] = V[(v * deg + kd) * 5 + w];
#pragma unroll
    for (int k = 0; k < 5; k++) {
        const double p = 1.0 / a[k][k];
#pragma unroll
        for (int w = 0; w < 5; w++) a[k][w] = w == k ? p : a[k][w] * p;
#pragma unroll
        for (int v = 0; v < 5; v++) {
            if (v == k) continue;
            const double f = a[v][k];
#pragma unroll
            for (int w = 0; w < 5; w++) a[v][w] = w == k ? -f * p : fma(-f, a[k][w], a[v][w]);
        }
    }
#pragma unroll
    for (int v = 0; v < 5; v++)
#pragma unroll
        for (int w = 0; w < 5; w++) dinvb[n * 25 + v * 5 + w] = float(a[v][w]);
}

// the same start on the FP32 state of the hex8 problem level
__global__ __launch_bounds__(256) void k_mg_pre_init32(MGLevelArgs L, const int* __restrict__ done) {
    if (done && *done) return;
    const float ith = float(L.cheb[0]);
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < L.n_nodes * L.nav; i += int64_t(gridDim.x) * 256) {
        const float r = float(L.r[i]);
        L.x32[i] = 0.f;
        L.res32[i] = r;
        L.da32[i] = L.dinv32[i] * r * ith;
    }
}

__device__ __forceinline__ double mg_val(double v) { return v; }
__device__ __forceinline__ double mg_val(float v) { return double(v); }
__device__ __forceinline__ double mg_val(uint16_t v) { return double(__half2float(__ushort_as_half(v))); }

// FP16 copy of a node-block matrix, each block row (node r, variable v: the
// deg x NAV values V[(v deg + k) NAV + w]) divided by its largest magnitude
template <int NAV>
__global__ __launch_bounds__(256) void k_to_f16_rows(int64_t n_nodes, const int64_t* __restrict__ blk_ptr,
                                                     const double* __restrict__ val, uint16_t* __restrict__ val16,
                                                     float* __restrict__ rs16) {
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n_nodes * NAV) return;
    const int64_t r = i / NAV, v = i - r * NAV;
    const int64_t k0 = blk_ptr[r];
    const int64_t deg = blk_ptr[r + 1] - k0;
    const int64_t b = k0 * NAV * NAV + v * deg * NAV, n = deg * NAV;
    double m = 0.0;
    for (int64_t t = 0; t < n; t++) m = fmax(m, fabs(val[b + t]));
    const float d = m > 0.0 ? float(m) : 1.0f;
    rs16[i] = d;
    for (int64_t t = 0; t < n; t++) val16[b + t] = __half_as_ushort(__float2half(float(val[b + t] / double(d))));
}

// node-block operator of a level (values T = double, their FP32 copy or the
// row-scaled FP16 copy; products and sums in FP64), 8 lanes per node, fused
// epilogue
template <typename T, int NAV>
__global__ __launch_bounds__(256) void k_mg_bsr(MGLevelArgs L, const T* __restrict__ vals, int mode, int step,
                                                const double* __restrict__ din, double* __restrict__ dout,
                                                const int* __restrict__ done) {
    if (done && *done) return;
    const int sub = threadIdx.x & 7;
    const double* __restrict__ in = mode == MG_RESID_INIT ? L.x : din;
    const double ith = L.cheb[0];
    const double c1 = step > 0 ? L.cheb[2 * step] : 0.0, c2 = step > 0 ? L.cheb[2 * step + 1] : 0.0;
    const int64_t nodes_per_grid = (int64_t(gridDim.x) * 256) / 8;
    for (int64_t r = (int64_t(blockIdx.x) * 256 + threadIdx.x) / 8; r < L.n_nodes; r += nodes_per_grid) {
        const int64_t k0 = L.blk_ptr[r];
        const int deg = int(L.blk_ptr[r + 1] - k0);
        const T* __restrict__ V = vals + k0 * (NAV * NAV);
        double sv[NAV];
#pragma unroll
        for (int v = 0; v < NAV; v++) sv[v] = 0.0;
        for (int k = sub; k < deg; k += 8) {
            const int64_t cb = int64_t(L.blk_col[k0 + k]) * NAV;
            double xw[NAV];
#pragma unroll
            for (int w = 0; w < NAV; w++) xw[w] = in[cb + w];
#pragma unroll
            for (int v = 0; v < NAV; v++) {
                const T* vr = V + (int64_t(v) * deg + k) * NAV;
#pragma unroll
                for (int w = 0; w < NAV; w++) sv[v] += mg_val(vr[w]) * xw[w];
            }
        }
#pragma unroll
        for (int v = 0; v < NAV; v++)
#pragma unroll
            for (int m = 4; m >= 1; m >>= 1) sv[v] += __shfl_xor(sv[v], m, 8);
        if constexpr (sizeof(T) == 2)
#pragma unroll
            for (int v = 0; v < NAV; v++) sv[v] *= double(L.rs16[r * NAV + v]);
        if (NAV == 5 && L.dinvb) {
            // block Jacobi: every lane of the node's group holds the 5 rows y;
            // each forms the 5 new residuals and lane v < 5 its direction entry
            // Dinv_b[v] . res
            const int64_t i0 = r * NAV;
            double rs[NAV];
#pragma unroll
            for (int w = 0; w < NAV; w++) rs[w] = (mode == MG_RESID_INIT ? L.r[i0 + w] : L.res[i0 + w]) - sv[w];
            if (sub < NAV) {
                const int64_t i = i0 + sub;
                double rsv = rs[0];
#pragma unroll
                for (int w = 1; w < NAV; w++) rsv = sub == w ? rs[w] : rsv;
                const float* B = L.dinvb + r * 25 + sub * 5;
                double z = 0.0;
#pragma unroll
                for (int w = 0; w < NAV; w++) z = fma(double(B[w]), rs[w], z);
                if (mode == MG_RESID_INIT) {
                    L.res[i] = rsv;
                    dout[i] = z * ith;
                } else {
                    const double dv = din[i];
                    L.x[i] += dv;
                    L.res[i] = rsv;
                    if (mode == MG_STEP) dout[i] = c1 * dv + c2 * z;
                }
            }
        } else if (sub < NAV) {
            double y = sv[0];
#pragma unroll
            for (int v = 1; v < NAV; v++) y = sub == v ? sv[v] : y;
            mg_epilogue(L, mode, r * NAV + sub, y, din, dout, c1, c2, ith);
        }
    }
}

// ---- a coarse hex8 level as a symmetric 27-point stencil --------------------
// The level's node blocks are exactly symmetric (B(r, m) = B(m, r)^T), so each
// node keeps half of them (FP32, 123 entries, entry e of node r at e N + r):
// the 13 blocks towards the neighbours o = 14..26 (o = (dx+1) + 3 (dy+1) +
// 9 (dz+1)) at e = (o - 14) 9 + 3 v + w, and the 6 entries v <= w of its own
// block at e = 117 + kSymIdx[v][w].  The block towards a neighbour m with
// o < 13 is the transpose of m's block towards r, read at index m (coalesced:
// m = r + a fixed offset).  Offsets outside the lattice hold 0.
__constant__ int8_t kSymIdx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
constexpr int kStenEntries = 123;

__global__ __launch_bounds__(256) void k_mg_to_stencil(MGLevelArgs L, float* __restrict__ sten) {
    const int64_t r = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (r >= L.n_nodes) return;
    const int64_t N = L.n_nodes;
    const uint32_t nxp = uint32_t(L.nx + 1), nyp = uint32_t(L.ny + 1), ru = uint32_t(r), rq = ru / nxp;
    const int ri = int(ru - rq * nxp), rj = int(rq % nyp), rk = int(rq / nyp);
    for (int e = 0; e < kStenEntries; e++) sten[e * N + r] = 0.f;
    const int64_t k0 = L.blk_ptr[r];
    const int deg = int(L.blk_ptr[r + 1] - k0);
    const float* __restrict__ V = L.val32 + k0 * 9;
    for (int k = 0; k < deg; k++) {
        const uint32_t cu = uint32_t(L.blk_col[k0 + k]), cq = cu / nxp;
        const int dx = int(cu - cq * nxp) - ri, dy = int(cq % nyp) - rj, dz = int(cq / nyp) - rk;
        if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) continue;   // not a lattice neighbour
        const int o = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
        if (o < 13) continue;   // kept by the neighbour
#pragma unroll
        for (int v = 0; v < 3; v++)
#pragma unroll
            for (int w = 0; w < 3; w++) {
                const float x = V[(v * deg + k) * 3 + w];
                if (o > 13) sten[((o - 14) * 9 + v * 3 + w) * N + r] = x;
                else if (v <= w) sten[(117 + kSymIdx[v][w]) * N + r] = x;
            }
    }
}

// Smoothing step of a stencil level: thread per node, coalesced FP32 stencil
// loads, FP64 products and sums, the fused Chebyshev epilogue
__global__ __launch_bounds__(256) void k_mg_stencil(MGLevelArgs L, const float* __restrict__ sten, int mode, int step,
                                                    const double* __restrict__ din, double* __restrict__ dout,
                                                    const int* __restrict__ done) {
    if (done && *done) return;
    // work-groups in XCD-aware order: the neighbouring rows whose stored
    // blocks this row reads transposed were read by work-groups of the same
    // XCD (its L2), not of another
    const int64_t r = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (r >= L.n_nodes) return;
    const int64_t N = L.n_nodes;
    const double* __restrict__ in = mode == MG_RESID_INIT ? L.x : din;
    const uint32_t nxp = uint32_t(L.nx + 1), nyp = uint32_t(L.ny + 1), ru = uint32_t(r), rq = ru / nxp;
    const int i = int(ru - rq * nxp), j = int(rq % nyp), k = int(rq / nyp);
    double y0, y1, y2;
    {   // own block (symmetric)
        const float* __restrict__ S = sten + 117 * N + r;
        const double s00 = S[0], s01 = S[N], s02 = S[2 * N], s11 = S[3 * N], s12 = S[4 * N], s22 = S[5 * N];
        const double x0 = in[r * 3], x1 = in[r * 3 + 1], x2 = in[r * 3 + 2];
        y0 = fma(s02, x2, fma(s01, x1, s00 * x0));
        y1 = fma(s12, x2, fma(s11, x1, s01 * x0));
        y2 = fma(s22, x2, fma(s12, x1, s02 * x0));
    }
    // the 26 neighbours, one z-layer of 9 offsets per pass (bounded loads in flight)
#pragma unroll 9
    for (int o = 0; o < 27; o++) {
        if (o == 13) continue;
        const int ui = i + o % 3 - 1, uj = j + (o / 3) % 3 - 1, uk = k + o / 9 - 1;
        const bool inside = ui >= 0 && ui <= L.nx && uj >= 0 && uj <= L.ny && uk >= 0 && uk <= L.nz;
        const int ni = min(max(ui, 0), L.nx), nj = min(max(uj, 0), L.ny), nk = min(max(uk, 0), L.nz);
        const int64_t m = int64_t(uint32_t(ni) + nxp * (uint32_t(nj) + nyp * uint32_t(nk)));
        // a neighbour outside the lattice is clamped into it and weighted 0
        const double wgt = inside ? 1.0 : 0.0;
        const double x0 = wgt * in[m * 3], x1 = wgt * in[m * 3 + 1], x2 = wgt * in[m * 3 + 2];
        // block (v, w) towards m: own entry (o - 14) 9 + 3 v + w at r, or the
        // neighbour's entry (12 - o) 9 + 3 w + v at m (uniform strides)
        const bool lower = o < 13;
        const float* __restrict__ S = lower ? sten + int64_t(12 - o) * 9 * N + m : sten + int64_t(o - 14) * 9 * N + r;
        const int64_t sv = lower ? N : 3 * N, sw = lower ? 3 * N : N;
        y0 = fma(double(S[0]), x0, y0); y0 = fma(double(S[sw]), x1, y0); y0 = fma(double(S[2 * sw]), x2, y0);
        y1 = fma(double(S[sv]), x0, y1); y1 = fma(double(S[sv + sw]), x1, y1); y1 = fma(double(S[sv + 2 * sw]), x2, y1);
        y2 = fma(double(S[2 * sv]), x0, y2); y2 = fma(double(S[2 * sv + sw]), x1, y2);
        y2 = fma(double(S[2 * sv + 2 * sw]), x2, y2);
    }
    const double ith = L.cheb[0];
    const double c1 = step > 0 ? L.cheb[2 * step] : 0.0, c2 = step > 0 ? L.cheb[2 * step + 1] : 0.0;
    mg_epilogue(L, mode, r * 3 + 0, y0, din, dout, c1, c2, ith);
    mg_epilogue(L, mode, r * 3 + 1, y1, din, dout, c1, c2, ith);
    mg_epilogue(L, mode, r * 3 + 2, y2, din, dout, c1, c2, ith);
}

// ---- the heat problem level as a stencil in lattice order ------------------
// Its Krylov solver keeps the caller's node-rank order; the V-cycle runs the
// problem level in lattice order, so its smoother reads a stencil formed once
// per set-up from the rank-ordered rows instead of values with a column index
// each: entry o = (dx+1) + 3 (dy+1) + 9 (dz+1) of lattice point p at o N + p
// (FP32, all 27 neighbours: the Jacobian is not symmetric; offsets outside
// the lattice hold 0).  (For the plate, 5 variables per node, the node-block
// smoother measured faster than lattice-ordered stencils.)
// A coarse level (lattice order, 1 variable) converted from its rows; the
// problem level (rank order) is gathered through precomputed row positions
// (k_mg_gather_stencil).
__global__ __launch_bounds__(256) void k_mg_to_stencil27(MGLevelArgs L, float* __restrict__ sten) {
    const int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (p >= L.n_nodes) return;
    const int64_t N = L.n_nodes;
    const uint32_t nxp = uint32_t(L.nx + 1), nyp = uint32_t(L.ny + 1), pu = uint32_t(p), pq = pu / nxp;
    const int pi = int(pu - pq * nxp), pj = int(pq % nyp), pk = int(pq / nyp);
    for (int e = 0; e < 27; e++) sten[e * N + p] = 0.f;
    const int64_t k0 = L.blk_ptr[p];
    const int deg = int(L.blk_ptr[p + 1] - k0);
    for (int k = 0; k < deg; k++) {
        const uint32_t qu = uint32_t(L.blk_col[k0 + k]), qq = qu / nxp;
        const int dx = int(qu - qq * nxp) - pi, dy = int(qq % nyp) - pj, dz = int(qq / nyp) - pk;
        if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) continue;   // not a lattice neighbour
        sten[((dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)) * N + p] = L.val32[k0 + k];
    }
}

// the same conversion with the row positions known: kpos[p 27 + o] is the
// position in lattice point p's (rank-ordered) row of its neighbour o, 255 for
// none (the mesh fixes it, so it is formed once on the host)
__global__ __launch_bounds__(256) void k_mg_gather_stencil(MGLevelArgs L, const int32_t* __restrict__ rank_of_lat,
                                                           const uint8_t* __restrict__ kpos,
                                                           float* __restrict__ sten) {
    const int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (p >= L.n_nodes) return;
    const int64_t N = L.n_nodes;
    const float* __restrict__ V = L.val32 + L.blk_ptr[rank_of_lat[p]];
#pragma unroll 9
    for (int o = 0; o < 27; o++) {
        const int kk = kpos[p * 27 + o];
        sten[o * N + p] = kk == 255 ? 0.f : V[kk];
    }
}

// The hex8 problem level (rank order, cells of several shapes) as the
// symmetric half stencil of k_mg_stencil in lattice order: lattice point p
// keeps the blocks towards its 13 forward neighbours and the upper triangle of
// its own block, read from the FP32 row values of its rank through kpos.
// Blocks symmetric to rounding become exactly symmetric (the upper ones serve
// both directions).
__global__ __launch_bounds__(256) void k_mg_gather_stencil3(MGLevelArgs L, const int32_t* __restrict__ rank_of_lat,
                                                            const uint8_t* __restrict__ kpos,
                                                            float* __restrict__ sten) {
    const int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (p >= L.n_nodes) return;
    const int64_t N = L.n_nodes;
    const int64_t r = rank_of_lat[p], k0 = L.blk_ptr[r];
    const int deg = int(L.blk_ptr[r + 1] - k0);
    const float* __restrict__ V = L.val32 + k0 * 9;
    for (int o = 13; o < 27; o++) {
        const int kk = kpos[p * 27 + o];
#pragma unroll
        for (int v = 0; v < 3; v++)
#pragma unroll
            for (int w = 0; w < 3; w++) {
                const float x = kk == 255 ? 0.f : V[(v * deg + kk) * 3 + w];
                if (o > 13) sten[((o - 14) * 9 + v * 3 + w) * N + p] = x;
                else if (v <= w) sten[(117 + kSymIdx[v][w]) * N + p] = x;
            }
    }
}

// smoothing step of the heat problem level in lattice order (FP32 stencil, FP64
// vectors and sums, fused Chebyshev epilogue)
// the heat problem level's stencil in FP16: each point's 27 entries divided
// by their largest magnitude (srow[p])
__global__ __launch_bounds__(256) void k_sten27_to_f16(int64_t N, const float* __restrict__ sten,
                                                       uint16_t* __restrict__ sten16, float* __restrict__ srow) {
    const int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (p >= N) return;
    float m = 0.f;
    for (int o = 0; o < 27; o++) m = fmaxf(m, fabsf(sten[o * N + p]));
    const float d = m > 0.f ? m : 1.f;
    srow[p] = d;
    for (int o = 0; o < 27; o++) sten16[o * N + p] = __half_as_ushort(__float2half(sten[o * N + p] / d));
}

// T = float: the stencil; T = uint16_t: its FP16 copy, rows scaled by L.rs16
template <typename T>
__global__ __launch_bounds__(256) void k_mg_stencil_heat(MGLevelArgs L, const T* __restrict__ sten, int mode,
                                                         int step, const double* __restrict__ din,
                                                         double* __restrict__ dout, const int* __restrict__ done) {
    if (done && *done) return;
    // work-groups in XCD-aware order: neighbouring rows' state through one L2
    const int64_t p = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (p >= L.n_nodes) return;
    const int64_t N = L.n_nodes;
    const double* __restrict__ in = mode == MG_RESID_INIT ? L.x : din;
    const uint32_t nxp = uint32_t(L.nx + 1), nyp = uint32_t(L.ny + 1), pu = uint32_t(p), pq = pu / nxp;
    const int i = int(pu - pq * nxp), j = int(pq % nyp), k = int(pq / nyp);
    double y = 0.0;
#pragma unroll 9
    for (int o = 0; o < 27; o++) {
        // neighbours outside the lattice are clamped into it (their entries are 0)
        const int ni = min(max(i + o % 3 - 1, 0), L.nx), nj = min(max(j + (o / 3) % 3 - 1, 0), L.ny),
                  nk = min(max(k + o / 9 - 1, 0), L.nz);
        const int64_t m = int64_t(uint32_t(ni) + nxp * (uint32_t(nj) + nyp * uint32_t(nk)));
        y = fma(mg_val(sten[o * N + p]), in[m], y);
    }
    if constexpr (sizeof(T) == 2) y *= double(L.rs16[p]);
    const double ith = L.cheb[0];
    const double c1 = step > 0 ? L.cheb[2 * step] : 0.0, c2 = step > 0 ? L.cheb[2 * step + 1] : 0.0;
    mg_epilogue(L, mode, p, y, din, dout, c1, c2, ith);
}

// power iteration for lambda_max(D^-1 A) on the heat problem level in lattice
// order with its stencil: out = D^-1 A (s in), s = *scale; the start vector
// (in == null) is the rank-ordered one, 1 + ((rank 7919) mod 1000) / 1000, read
// through rank_of_lat, so the iterates are those of the node-block form
// permuted; partial sums of out.out per work-group
__global__ __launch_bounds__(256) void k_mg_power_heat(MGLevelArgs L, const float* __restrict__ sten,
                                                       const int32_t* __restrict__ rank_of_lat,
                                                       const double* __restrict__ in, const double* __restrict__ scale,
                                                       double* __restrict__ out, double* __restrict__ part) {
    const int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x;
    double acc = 0.0;
    if (p < L.n_nodes) {
        const int64_t N = L.n_nodes;
        const double sc = in ? *scale : 1.0;
        const uint32_t nxp = uint32_t(L.nx + 1), nyp = uint32_t(L.ny + 1), pu = uint32_t(p), pq = pu / nxp;
        const int i = int(pu - pq * nxp), j = int(pq % nyp), k = int(pq / nyp);
        double y = 0.0;
#pragma unroll 9
        for (int o = 0; o < 27; o++) {
            const int ni = min(max(i + o % 3 - 1, 0), L.nx), nj = min(max(j + (o / 3) % 3 - 1, 0), L.ny),
                      nk = min(max(k + o / 9 - 1, 0), L.nz);
            const int64_t m = int64_t(uint32_t(ni) + nxp * (uint32_t(nj) + nyp * uint32_t(nk)));
            const double x = in ? sc * in[m] : 1.0 + double((int64_t(rank_of_lat[m]) * 7919) % 1000) / 1000.0;
            y = fma(double(sten[o * N + p]), x, y);
        }
        const double o = L.dinv[p] * y;
        out[p] = o;
        acc = o * o;
    }
    block_sum_store(acc, part, blockIdx.x);
}

// node-major vectors between rank order and lattice order: to_lat = 1:
// dst[p nav + v] = src[rank_of_lat[p] nav + v], else the inverse scatter
__global__ __launch_bounds__(256) void k_perm_rank_lat(int64_t n_nodes, int nav, const int32_t* __restrict__ rank_of_lat,
                                                       const double* __restrict__ src, double* __restrict__ dst,
                                                       int to_lat, const int* __restrict__ done) {
    if (done && *done) return;
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n_nodes * nav) return;
    const int64_t p = i / nav, v = i - p * nav, r = int64_t(rank_of_lat[p]) * nav + v;
    if (to_lat) dst[i] = src[r];
    else dst[r] = src[i];
}

// out (rank order) = x0 + dlast (lattice order): the V-cycle's output
__global__ __launch_bounds__(256) void k_mg_combine_rank(int64_t n_nodes, int nav,
                                                         const int32_t* __restrict__ rank_of_lat,
                                                         const double* __restrict__ x0,
                                                         const double* __restrict__ dlast, double* __restrict__ out,
                                                         const int* __restrict__ done) {
    if (done && *done) return;
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n_nodes * nav) return;
    const int64_t p = i / nav, v = i - p * nav;
    out[int64_t(rank_of_lat[p]) * nav + v] = x0[i] + dlast[i];
}

// ---- level 0 on a lattice: per-node matrix-free operator -------------------
// Thread per lattice point n, vectors component-major in lattice order (x
// fastest), so each neighbour gather of a wave is one contiguous 512-byte run;
// work-groups in XCD-aware order; for each existing cell at n
// (local corner a) y_n += s_e Khat[a][b] x_b.  Khat indices are uniform
// across the wave, so the class matrix is read with scalar-cache loads.  Each
// y_n is a fixed-order sum: results are deterministic.
__host__ __device__ constexpr int lat_cx(int b) { return (b == 1 || b == 2 || b == 5 || b == 6) ? 1 : 0; }
__host__ __device__ constexpr int lat_cy(int b) { return (b == 2 || b == 3 || b == 6 || b == 7) ? 1 : 0; }
__host__ __device__ constexpr int lat_cz(int b) { return b >= 4 ? 1 : 0; }
__host__ __device__ constexpr int lat_corner(int dx, int dy, int dz) {
    return dz * 4 + (dy ? (dx ? 2 : 3) : (dx ? 1 : 0));
}

template <typename T>
__device__ __forceinline__ T fma_t(T a, T b, T c) {
    if constexpr (sizeof(T) == 4) return fmaf(a, b, c);
    else return fma(a, b, c);
}

// Row n = (i, j, k) of the lattice operator in T (double: the CG product;
// float: the multigrid smoothing): y = sum over the existing cells at n of
// s_c Khat[a_c][b] x_b with the class matrix K (Khat or its FP32 copy), and
// cnt = the number of cells at n.  x is component-major with stride n_lat.
template <typename T>
__device__ __forceinline__ void lattice_row(const LatticeOpArgs& O, const T* __restrict__ in,
                                            const T* __restrict__ Kc, int64_t n_lat, int i, int j, int k, bool valid,
                                            T& y0, T& y1, T& y2, int& cnt) {
    const int64_t nxp = O.nx + 1, nyp = O.ny + 1;
    // neighbour-outer, one z-layer of 9 neighbours at a time (bounded live
    // ranges): each neighbour value is read once and feeds every cell shared
    // with n, y += s_c Khat[a_c][b] x_o (scaled per block product, so only y
    // is carried: fewer registers, more waves in flight to hide the gathers).
    // Neighbours outside the lattice are clamped into it; every cell holding
    // one is missing and scaled by 0.  Gathers address each component array
    // from a uniform base with a 32-bit byte offset (fewer than 2^29 lattice
    // points), and the products are explicit FMA chains: 9 FMAs per 3x3
    // block, 3 per (cell, layer) scaling.
    const char* __restrict__ b0 = reinterpret_cast<const char*>(in);
    const char* __restrict__ b1 = reinterpret_cast<const char*>(in + n_lat);
    const char* __restrict__ b2 = reinterpret_cast<const char*>(in + 2 * n_lat);
    const uint32_t sx = uint32_t(nxp), sz = uint32_t(nxp * nyp);
    auto load_layer = [&](int lz, T (&xv)[9][3]) __attribute__((always_inline)) {
        const uint32_t mk = uint32_t(min(max(k + lz - 1, 0), O.nz));
#pragma unroll
        for (int q = 0; q < 9; q++) {
            const uint32_t mi = uint32_t(min(max(i + (q % 3) - 1, 0), O.nx));
            const uint32_t mj = uint32_t(min(max(j + (q / 3) - 1, 0), O.ny));
            const uint32_t off = (mi + sx * mj + sz * mk) * uint32_t(sizeof(T));
            xv[q][0] = *reinterpret_cast<const T*>(b0 + off);
            xv[q][1] = *reinterpret_cast<const T*>(b1 + off);
            xv[q][2] = *reinterpret_cast<const T*>(b2 + off);
        }
    };
    // the first layer's gathers are issued before the scale loads (both in flight together)
    T x0v[9][3];
    load_layer(0, x0v);
    // cells at n (c bit d: the cell extends towards +d from n - 1) and their
    // scales; a missing cell gets the scale 0 (its index is clamped into the
    // lattice so every load stays in bounds)
    T sc[8];
    cnt = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
        const int ci = i - 1 + cx, cj = j - 1 + cy, ck = k - 1 + cz;
        const bool has = valid & (ci >= 0) & (ci < O.nx) & (cj >= 0) & (cj < O.ny) & (ck >= 0) & (ck < O.nz);
        const int cell = min(max(ci, 0), O.nx - 1) + O.nx * (min(max(cj, 0), O.ny - 1) + O.ny * min(max(ck, 0), O.nz - 1));
        // unconditional load (the index is clamped) weighted 0 for a missing
        // cell: used on every path, so it is not sunk into a branch
        sc[c] = T(O.sE_cell[cell]) * T(has ? 1 : 0);
        cnt += has ? 1 : 0;
    }
    auto compute_layer = [&](int lz, const T (&xv)[9][3]) __attribute__((always_inline)) {
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
            const int bz = lz - cz;                      // uniform across the wave
            if (bz < 0 || bz > 1) continue;
            const int a = lat_corner(1 - cx, 1 - cy, 1 - cz);
            T t0 = T(0), t1 = T(0), t2 = T(0);
#pragma unroll
            for (int q = 0; q < 9; q++) {
                const int bx = (q % 3) - cx, by = (q / 3) - cy;
                if (bx < 0 || bx > 1 || by < 0 || by > 1) continue;   // compile time
                const int b = bz * 4 + lat_corner(bx, by, 0);
                const T* K = Kc + (a * 8 + b) * 9;
                t0 = fma_t(K[0], xv[q][0], t0); t0 = fma_t(K[1], xv[q][1], t0); t0 = fma_t(K[2], xv[q][2], t0);
                t1 = fma_t(K[3], xv[q][0], t1); t1 = fma_t(K[4], xv[q][1], t1); t1 = fma_t(K[5], xv[q][2], t1);
                t2 = fma_t(K[6], xv[q][0], t2); t2 = fma_t(K[7], xv[q][1], t2); t2 = fma_t(K[8], xv[q][2], t2);
            }
            y0 = fma_t(sc[c], t0, y0);
            y1 = fma_t(sc[c], t1, y1);
            y2 = fma_t(sc[c], t2, y2);
        }
    };
    y0 = y1 = y2 = T(0);
    compute_layer(0, x0v);
#pragma unroll 1
    for (int lz = 1; lz < 3; lz++) {
        T xv[9][3];
        load_layer(lz, xv);
        compute_layer(lz, xv);
    }
}

__device__ __forceinline__ void lattice_point(const LatticeOpArgs& O, int64_t p, int& i, int& j, int& k) {
    // 32-bit index arithmetic (the lattice has fewer than 2^29 points)
    const uint32_t pu = uint32_t(p), nxu = uint32_t(O.nx + 1), pj = pu / nxu;
    i = int(pu - pj * nxu);
    j = int(pj % uint32_t(O.ny + 1));
    k = int(pj / uint32_t(O.ny + 1));
}

// y = A x (FP64 CG operator) and partial sums of x.y into C.part0
__global__ __launch_bounds__(256) void k_lattice_apply(CGArgs C, const double* __restrict__ in,
                                                       double* __restrict__ ycg, const int* __restrict__ done) {
    if (done && *done) return;
    const LatticeOpArgs& O = C.lat;
    const int64_t n_lat = int64_t(O.nx + 1) * (O.ny + 1) * (O.nz + 1);
    const int64_t p = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    const bool valid = p < n_lat;
    const int64_t r = valid ? p : 0;
    int i, j, k, cnt;
    lattice_point(O, r, i, j, k);
    double y0, y1, y2;
    lattice_row<double>(O, in, O.Khat, n_lat, i, j, k, valid, y0, y1, y2, cnt);
    double dsum = 0.0;
    if (valid) {
        const int m = O.nodecon[r];
        const double x0 = in[r], x1 = in[r + n_lat], x2 = in[r + 2 * n_lat];
        if (m & 1) y0 = cnt * x0;
        if (m & 2) y1 = cnt * x1;
        if (m & 4) y2 = cnt * x2;
        ycg[r] = y0;
        ycg[r + n_lat] = y1;
        ycg[r + 2 * n_lat] = y2;
        dsum = y0 * x0 + y1 * x1 + y2 * x2;
    }
    block_sum_store(dsum, C.part0, blockIdx.x);
}

// Chebyshev epilogue of one DOF of the FP32 problem-level state (see mg_epilogue)
__device__ __forceinline__ void mg_epilogue32(const MGLevelArgs& L, int mode, int64_t i, float y,
                                              const float* __restrict__ din, float* __restrict__ dout, float c1,
                                              float c2, float ith) {
    if (mode == MG_RESID_INIT) {
        const float rs = float(L.r[i]) - y;
        L.res32[i] = rs;
        dout[i] = L.dinv32[i] * rs * ith;
    } else {
        const float dv = din[i];
        L.x32[i] += dv;
        const float rs = L.res32[i] - y;
        L.res32[i] = rs;
        if (mode == MG_STEP) dout[i] = c1 * dv + c2 * L.dinv32[i] * rs;
    }
}

// Multigrid smoothing step on the hex8 problem level in FP32: the lattice
// operator on the FP32 direction (MG_RESID_INIT: on the iterate) with the
// row-local Chebyshev epilogue.  The smoother is part of the preconditioner;
// the CG operator, its residuals and the convergence test stay FP64.
__global__ __launch_bounds__(256) void k_lattice_smooth(LatticeOpArgs O, MGLevelArgs L, int mode, int step,
                                                        const float* __restrict__ din, float* __restrict__ dout,
                                                        const int* __restrict__ done) {
    if (done && *done) return;
    const int64_t n_lat = int64_t(O.nx + 1) * (O.ny + 1) * (O.nz + 1);
    const int64_t p = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (p >= n_lat) return;   // no block-level synchronisation below
    const float* __restrict__ in = mode == MG_RESID_INIT ? L.x32 : din;
    int i, j, k, cnt;
    lattice_point(O, p, i, j, k);
    float y0, y1, y2;
    lattice_row<float>(O, in, O.Khat32, n_lat, i, j, k, true, y0, y1, y2, cnt);
    const int m = O.nodecon[p];
    if (m & 1) y0 = float(cnt) * in[p];
    if (m & 2) y1 = float(cnt) * in[p + n_lat];
    if (m & 4) y2 = float(cnt) * in[p + 2 * n_lat];
    const float ith = float(L.cheb[0]);
    const float c1 = step > 0 ? float(L.cheb[2 * step]) : 0.f, c2 = step > 0 ? float(L.cheb[2 * step + 1]) : 0.f;
    mg_epilogue32(L, mode, p, y0, din, dout, c1, c2, ith);
    mg_epilogue32(L, mode, p + n_lat, y1, din, dout, c1, c2, ith);
    mg_epilogue32(L, mode, p + 2 * n_lat, y2, din, dout, c1, c2, ith);
}

// Row information of the lattice operator without an assembled matrix: for
// each row the assembled entries A_rj = sum over the cells shared by the two
// nodes of s_c Khat[a_c][b_c], so the inverse diagonal and the Gershgorin
// ratio sum_j |A_rj| / A_rr equal those of the assembled matrix (constrained
// rows and columns as assembled: zero off the diagonal, cells-at-node on it).
__global__ __launch_bounds__(256) void k_lattice_rowinfo(LatticeOpArgs O, double* __restrict__ dinv,
                                                         float* __restrict__ dinv32, double* __restrict__ part) {
    const int64_t nxp = O.nx + 1, nyp = O.ny + 1;
    const int64_t n_lat = nxp * nyp * (O.nz + 1);
    const int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x;
    double g = 0.0;
    if (p < n_lat) {
        int i, j, k;
        lattice_point(O, p, i, j, k);
        double sc[8];
        int cnt = 0;
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
            const int ci = i - 1 + cx, cj = j - 1 + cy, ck = k - 1 + cz;
            const bool has = ci >= 0 && ci < O.nx && cj >= 0 && cj < O.ny && ck >= 0 && ck < O.nz;
            const int cell = min(max(ci, 0), O.nx - 1) + O.nx * (min(max(cj, 0), O.ny - 1) + O.ny * min(max(ck, 0), O.nz - 1));
            sc[c] = has ? O.sE_cell[cell] : 0.0;
            cnt += has ? 1 : 0;
        }
        const int mn = O.nodecon[p];
        double rs[3] = {0.0, 0.0, 0.0}, dg[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int d = 0; d < 27; d++) {
            const int dx = d % 3 - 1, dy = (d / 3) % 3 - 1, dz = d / 9 - 1;
            // neighbour o = n + (dx, dy, dz), clamped into the lattice (a
            // neighbour outside lies only in missing cells: its block is 0)
            const int oi = min(max(i + dx, 0), O.nx), oj = min(max(j + dy, 0), O.ny), ok = min(max(k + dz, 0), O.nz);
            const int mo = O.nodecon[oi + nxp * (oj + nyp * int64_t(ok))];
            double B[9];
#pragma unroll
            for (int e = 0; e < 9; e++) B[e] = 0.0;
#pragma unroll
            for (int c = 0; c < 8; c++) {
                const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
                // o lies in cell c when each offset is cell_d - 1 or cell_d
                if (dx < cx - 1 || dx > cx || dy < cy - 1 || dy > cy || dz < cz - 1 || dz > cz) continue;
                const int a = lat_corner(1 - cx, 1 - cy, 1 - cz);
                const int b = lat_corner(dx - cx + 1, dy - cy + 1, dz - cz + 1);
                const double* K = O.Khat + (a * 8 + b) * 9;
#pragma unroll
                for (int e = 0; e < 9; e++) B[e] = fma(sc[c], K[e], B[e]);
            }
#pragma unroll
            for (int v = 0; v < 3; v++)
#pragma unroll
                for (int w = 0; w < 3; w++) {
                    const double val = ((mn >> v) & 1) || ((mo >> w) & 1) ? 0.0 : B[v * 3 + w];
                    rs[v] += fabs(val);
                    if (d == 13 && v == w) dg[v] = val;
                }
        }
#pragma unroll
        for (int v = 0; v < 3; v++) {
            // constrained rows hold only the diagonal (one per cell at the node)
            const bool cv = (mn >> v) & 1;
            const double diag = cv ? double(cnt) : dg[v];
            const double di = diag > 0.0 ? 1.0 / diag : 0.0;
            g = fmax(g, cv ? 1.0 : (diag > 0.0 ? rs[v] / diag : 0.0));
            dinv[v * n_lat + p] = di;
            dinv32[v * n_lat + p] = float(di);
        }
    }
    block_max_store(g, part, blockIdx.x);
}

// node-vector permutation between rank order (node-major, 3 values per node)
// and the component-major lattice order: gather dst[v N + p] = src[3 map[p] + v],
// or scatter src[v N + p] -> dst[3 map[p] + v]
__global__ __launch_bounds__(256) void k_perm_nodes(int64_t n_nodes, const int32_t* __restrict__ map,
                                                    const double* __restrict__ src, double* __restrict__ dst,
                                                    int scatter) {
    const int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (p >= n_nodes) return;
    const int64_t m = int64_t(map[p]) * 3;
    if (scatter) {
        dst[m] = src[p]; dst[m + 1] = src[n_nodes + p]; dst[m + 2] = src[2 * n_nodes + p];
    } else {
        dst[p] = src[m]; dst[n_nodes + p] = src[m + 1]; dst[2 * n_nodes + p] = src[m + 2];
    }
}

// dinv (when wanted) and the Gershgorin bound over node blocks, 8 lanes per node
template <int NAV>
__global__ __launch_bounds__(256) void k_mg_rowinfo_nb(int64_t n_nodes, const int64_t* __restrict__ blk_ptr,
                                                       const int32_t* __restrict__ blk_col,
                                                       const double* __restrict__ val, double* __restrict__ dinv,
                                                       double* __restrict__ part) {
    const int sub = threadIdx.x & 7;
    double g = 0.0;
    const int64_t nodes_per_grid = (int64_t(gridDim.x) * 256) / 8;
    for (int64_t r = (int64_t(blockIdx.x) * 256 + threadIdx.x) / 8; r < n_nodes; r += nodes_per_grid) {
        const int64_t k0 = blk_ptr[r];
        const int deg = int(blk_ptr[r + 1] - k0);
        const double* __restrict__ V = val + k0 * (NAV * NAV);
        double sv[NAV], dv[NAV];
#pragma unroll
        for (int v = 0; v < NAV; v++) sv[v] = dv[v] = 0.0;
        for (int kk = sub; kk < deg; kk += 8) {
            const bool dg = blk_col[k0 + kk] == r;
#pragma unroll
            for (int v = 0; v < NAV; v++) {
                const double* vr = V + (int64_t(v) * deg + kk) * NAV;
#pragma unroll
                for (int w = 0; w < NAV; w++) sv[v] += fabs(vr[w]);
                if (dg) dv[v] = vr[v];
            }
        }
#pragma unroll
        for (int v = 0; v < NAV; v++)
#pragma unroll
            for (int m = 4; m >= 1; m >>= 1) {
                sv[v] += __shfl_xor(sv[v], m, 8);
                dv[v] += __shfl_xor(dv[v], m, 8);
            }
        if (sub == 0)
#pragma unroll
            for (int v = 0; v < NAV; v++) {
                g = fmax(g, dv[v] > 0.0 ? sv[v] / dv[v] : 0.0);
                if (dinv) dinv[r * NAV + v] = dv[v] > 0.0 ? 1.0 / dv[v] : 0.0;
            }
    }
    block_max_store(g, part, blockIdx.x);
}

// power iteration for lambda_max(D^-1 A): out = D^-1 A (s in), s = *scale
// (1 / ||in||; the start vector is built in place when in == null), with
// partial sums of out.out
template <int NAV>
__global__ __launch_bounds__(256) void k_mg_power(MGLevelArgs L, const double* __restrict__ val,
                                                  const double* __restrict__ in, const double* __restrict__ scale,
                                                  double* __restrict__ out, double* __restrict__ part) {
    const int sub = threadIdx.x & 7;
    const double sc = in ? *scale : 1.0;
    double acc = 0.0;
    const int64_t nodes_per_grid = (int64_t(gridDim.x) * 256) / 8;
    for (int64_t r = (int64_t(blockIdx.x) * 256 + threadIdx.x) / 8; r < L.n_nodes; r += nodes_per_grid) {
        const int64_t k0 = L.blk_ptr[r];
        const int deg = int(L.blk_ptr[r + 1] - k0);
        const double* __restrict__ V = val + k0 * (NAV * NAV);
        double sv[NAV];
#pragma unroll
        for (int v = 0; v < NAV; v++) sv[v] = 0.0;
        for (int k = sub; k < deg; k += 8) {
            const int64_t cb = int64_t(L.blk_col[k0 + k]) * NAV;
            double xw[NAV];
#pragma unroll
            for (int w = 0; w < NAV; w++)   // start vector 1 + ((i * 7919) mod 1000) / 1000
                xw[w] = in ? sc * in[cb + w] : 1.0 + double(((cb + w) * 7919) % 1000) / 1000.0;
#pragma unroll
            for (int v = 0; v < NAV; v++) {
                const double* vr = V + (int64_t(v) * deg + k) * NAV;
#pragma unroll
                for (int w = 0; w < NAV; w++) sv[v] += vr[w] * xw[w];
            }
        }
#pragma unroll
        for (int v = 0; v < NAV; v++)
#pragma unroll
            for (int m = 4; m >= 1; m >>= 1) sv[v] += __shfl_xor(sv[v], m, 8);
        if (sub == 0)
#pragma unroll
            for (int v = 0; v < NAV; v++) {
                double o = L.dinv[r * NAV + v] * sv[v];
                if (NAV == 5 && L.dinvb) {   // block Jacobi: Dinv_b A
                    o = 0.0;
#pragma unroll
                    for (int w = 0; w < NAV; w++) o = fma(double(L.dinvb[r * 25 + v * 5 + w]), sv[w], o);
                }
                out[r * NAV + v] = o;
                acc += o * o;
            }
    }
    block_sum_store(acc, part, blockIdx.x);
}

// norm of the power iterate: scal[0] = 1 / ||out||, *lam = safety * ||out|| / ||start||
__global__ __launch_bounds__(256) void k_mg_power_norm(const double* __restrict__ part, int n, int first,
                                                       double safety, double* __restrict__ scal,
                                                       double* __restrict__ lam) {
    const double s = reduce_partials(part, n);
    if (threadIdx.x == 0) {
        const double nrm = sqrt(s);
        // the first step multiplies the unnormalised start vector: its norm is scal[1]
        const double rel = first ? nrm / scal[1] : nrm;
        scal[0] = nrm > 0.0 ? 1.0 / nrm : 0.0;
        *lam = safety * rel;
    }
}

// squared norm of the start vector (for the first Rayleigh-free estimate)
__global__ __launch_bounds__(256) void k_mg_power_start(int64_t n, double* __restrict__ part) {
    double acc = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        const double x = 1.0 + double((i * 7919) % 1000) / 1000.0;
        acc += x * x;
    }
    block_sum_store(acc, part, blockIdx.x);
}

__global__ __launch_bounds__(256) void k_mg_power_start_norm(const double* __restrict__ part, int n,
                                                             double* __restrict__ scal) {
    const double s = reduce_partials(part, n);
    if (threadIdx.x == 0) scal[1] = sqrt(s);
}

// Gauss-Jordan inverse of an SPD matrix (no pivoting), n <= 136, one block of
// 1024 threads.  Thread t keeps the 9 x 3 entries at rows (t >> 6) + 16 r and
// columns (t & 63) + 64 c in registers (a 144 x 192 cover of the matrix; the
// entries outside n x n stay 0).  Pivot k's row lives in wave k & 15 (register
// row k >> 4) and its column in lane k & 63 of every wave (register column
// k >> 6), so the owners are known without searching: lane k & 63 publishes
// the column, the pivot row's wave broadcasts the pivot across its lanes,
// scales the row in registers and publishes it (entry k = 1 / pivot), and
// every thread updates its entries with selects instead of branches.  The
// pivot column and row are double-buffered in LDS, so one barrier per pivot
// suffices: a buffer is rewritten two pivots later, after every thread has
// passed the barrier that follows its last read.
__global__ __launch_bounds__(1024) void k_mg_invert_lds(int n, double* __restrict__ A) {
    __shared__ double colk[2][144], rowk[2][192];
    const int t = threadIdx.x, i0 = t >> 6, j0 = t & 63;
    double a[9][3];
#pragma unroll
    for (int r = 0; r < 9; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int i = i0 + 16 * r, j = j0 + 64 * c;
            a[r][c] = (i < n && j < n) ? A[i * n + j] : 0.0;
        }
    for (int k = 0; k < n; k++) {
        const int b = k & 1, kr = k >> 4, ki = k & 15, kc = k >> 6, kj = k & 63;
        if (j0 == kj)   // column k (before this pivot's update)
#pragma unroll
            for (int r = 0; r < 9; r++) colk[b][i0 + 16 * r] = kc == 0 ? a[r][0] : (kc == 1 ? a[r][1] : a[r][2]);
        if (i0 == ki) {   // the pivot row's wave (uniform branch)
            double rv[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int r = 0; r < 9; r++)
#pragma unroll
                for (int c = 0; c < 3; c++) rv[c] = r == kr ? a[r][c] : rv[c];
            const double piv = __shfl(kc == 0 ? rv[0] : (kc == 1 ? rv[1] : rv[2]), kj);
            const double p = 1.0 / piv;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const int j = j0 + 64 * c;
                rowk[b][j] = j == k ? p : rv[c] * p;
            }
        }
        __syncthreads();
        const double p = rowk[b][k];
        double ck[9], rk[3];
#pragma unroll
        for (int r = 0; r < 9; r++) ck[r] = colk[b][i0 + 16 * r];
#pragma unroll
        for (int c = 0; c < 3; c++) rk[c] = rowk[b][j0 + 64 * c];
#pragma unroll
        for (int r = 0; r < 9; r++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const bool is_row = i0 == ki && r == kr, is_col = j0 + 64 * c == k;
                const double v = is_col ? -ck[r] * p : fma(-ck[r], rk[c], a[r][c]);
                a[r][c] = is_row ? rk[c] : v;
            }
    }
#pragma unroll
    for (int r = 0; r < 9; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int i = i0 + 16 * r, j = j0 + 64 * c;
            if (i < n && j < n) A[i * n + j] = a[r][c];
        }
}

// ---- transfers ---------------------------------------------------------------
// r_C = sum over the 3^DIM fine points around 2C of w * res_f (w = prod of 1 or
// 1/2), constrained fine DOFs weighted 0, constrained coarse DOFs zero (P^T),
// fused with the start of the coarse level's pre-smoothing.
// FL: layout of the fine vectors (FL_LAT lattice order, node-major; FL_RANK
// indexed through rank_of_lat, the assembled problem levels: plate, heat and
// hex8 with cells of several shapes; FL_PLANAR lattice order, component-major,
// the matrix-free hex8 problem level); coarse levels are in lattice order,
// node-major.
enum { FL_LAT = 0, FL_RANK = 1, FL_PLANAR = 2 };
template <int FL, int NAV>
__device__ __forceinline__ int64_t mg_fine_index(const MGLevelArgs& F, int64_t p, int v) {
    if (FL == FL_PLANAR) return int64_t(v) * F.n_nodes + p;
    return (FL == FL_RANK ? int64_t(F.rank_of_lat[p]) : p) * NAV + v;
}

template <int DIM, int NAV, int FL>
__global__ __launch_bounds__(256) void k_mg_restrict(MGLevelArgs F, MGLevelArgs Cl, const int* __restrict__ done) {
    if (done && *done) return;
    const int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (p >= Cl.n_nodes) return;
    // 32-bit index arithmetic (lattice point counts are below 2^31)
    const uint32_t cx = uint32_t(Cl.nx + 1), cy = uint32_t(Cl.ny + 1), pu = uint32_t(p), pq = pu / cx;
    const int I = int(pu - pq * cx), J = int(pq % cy), K = int(pq / cy);
    const uint32_t fx = uint32_t(F.nx + 1), fy = uint32_t(F.ny + 1);
    constexpr int NO = DIM == 3 ? 27 : 9;
    // all loads are independent (unrolled, predicated) so they are in flight together
    double sv[NAV];
#pragma unroll
    for (int v = 0; v < NAV; v++) sv[v] = 0.0;
#pragma unroll
    for (int o = 0; o < NO; o++) {
        const int oi = (o % 3) - 1, oj = ((o / 3) % 3) - 1, ok = DIM == 3 ? (o / 9) - 1 : 0;
        const int fi = 2 * I + oi, fj = 2 * J + oj, fk = 2 * K + ok;
        const bool in_lat = fi >= 0 && fi <= F.nx && fj >= 0 && fj <= F.ny && fk >= 0 && fk <= F.nz;
        // out-of-lattice points are clamped (valid address) and weighted 0
        const double w = in_lat ? (ok ? 0.5 : 1.0) * (oj ? 0.5 : 1.0) * (oi ? 0.5 : 1.0) : 0.0;
        const uint32_t fp = uint32_t(in_lat ? fi : 0) + fx * (uint32_t(in_lat ? fj : 0) + fy * uint32_t(in_lat ? fk : 0));
#pragma unroll
        for (int v = 0; v < NAV; v++) {
            const int64_t rf = mg_fine_index<FL, NAV>(F, fp, v);
            const double res = FL == FL_PLANAR ? double(F.res32[rf]) : F.res[rf];
            sv[v] += (w * (1.0 - double(F.con[rf]))) * res;
        }
    }
    // and the coarse level's pre-smoothing start (x = 0, res = r, first
    // Chebyshev direction D^-1 r / theta)
    const int64_t rc = p * NAV;
    const double ith = Cl.cheb[0];
#pragma unroll
    for (int v = 0; v < NAV; v++) {
        const double r = Cl.con[rc + v] ? 0.0 : sv[v];
        Cl.r[rc + v] = r;
        Cl.x[rc + v] = 0.0;
        Cl.res[rc + v] = r;
        Cl.da[rc + v] = Cl.dinv[rc + v] * r * ith;
    }
}

// x_f += P (x_C + dlast_C) at unconstrained fine DOFs (tri/bilinear interpolation)
template <int DIM, int NAV, int FL>
__global__ __launch_bounds__(256) void k_mg_prolong(MGLevelArgs F, MGLevelArgs Cl, const double* __restrict__ dlast,
                                                    const int* __restrict__ done) {
    if (done && *done) return;
    const int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (p >= F.n_nodes) return;
    // 32-bit index arithmetic (lattice point counts are below 2^31)
    const uint32_t fx = uint32_t(F.nx + 1), fy = uint32_t(F.ny + 1), pu = uint32_t(p), pq = pu / fx;
    const int fi = int(pu - pq * fx), fj = int(pq % fy), fk = int(pq / fy);
    const uint32_t cx = uint32_t(Cl.nx + 1), cy = uint32_t(Cl.ny + 1);
    constexpr int NQ = 1 << DIM;
    double sv[NAV];
#pragma unroll
    for (int v = 0; v < NAV; v++) sv[v] = 0.0;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int a = q & 1, b = (q >> 1) & 1, c = DIM == 3 ? (q >> 2) : 0;
        // an even fine index takes its one coarse point (a = 0), an odd one both
        // neighbours; the unused corner of an even index is clamped and weighted 0
        const bool use = !((a && !(fi & 1)) || (b && !(fj & 1)) || (c && !(fk & 1)));
        const double w = use ? ((fi & 1) ? 0.5 : 1.0) * ((fj & 1) ? 0.5 : 1.0) * ((fk & 1) ? 0.5 : 1.0) : 0.0;
        const uint32_t CI = uint32_t(min((fi >> 1) + a, Cl.nx)), CJ = uint32_t(min((fj >> 1) + b, Cl.ny)),
                       CK = uint32_t(min((fk >> 1) + c, Cl.nz));
        const int64_t rc = int64_t(CI + cx * (CJ + cy * CK)) * NAV;
        if (use)
#pragma unroll
            for (int v = 0; v < NAV; v++) sv[v] += w * (dlast ? Cl.x[rc + v] + dlast[rc + v] : Cl.x[rc + v]);
    }
#pragma unroll
    for (int v = 0; v < NAV; v++) {
        const int64_t rf = mg_fine_index<FL, NAV>(F, p, v);
        if (FL == FL_PLANAR) {
            if (!F.con[rf]) F.x32[rf] += float(sv[v]);
        } else {
            if (!F.con[rf]) F.x[rf] += sv[v];
        }
    }
}

// ---- coarsest level: dense inverse -------------------------------------------
template <int NAV>
__global__ __launch_bounds__(256) void k_mg_dense_gather(MGLevelArgs L, double* __restrict__ A) {
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    const int64_t n = L.n_nodes * NAV;
    if (i >= n) return;
    const int64_t r = i / NAV;
    const int v = int(i - r * NAV);
    const int64_t k0 = L.blk_ptr[r];
    const int64_t deg = L.blk_ptr[r + 1] - k0;
    const double* row = L.val + k0 * (NAV * NAV) + int64_t(v) * deg * NAV;
    for (int64_t k = 0; k < deg; k++) {
        const int64_t c = L.blk_col[k0 + k];
#pragma unroll
        for (int w = 0; w < NAV; w++) A[i * n + c * NAV + w] = row[k * NAV + w];
    }
}

// in-place Gauss-Jordan inverse of an SPD matrix (no pivoting), one block
__global__ __launch_bounds__(1024) void k_mg_invert(int n, double* __restrict__ A) {
    __shared__ double colk[1536];
    for (int k = 0; k < n; k++) {
        for (int i = threadIdx.x; i < n; i += 1024) colk[i] = A[int64_t(i) * n + k];
        __syncthreads();
        const double p = 1.0 / colk[k];
        for (int j = threadIdx.x; j < n; j += 1024)
            A[int64_t(k) * n + j] = (j == k) ? p : A[int64_t(k) * n + j] * p;
        __syncthreads();
        const int64_t nn = int64_t(n) * n;
        for (int64_t idx = threadIdx.x; idx < nn; idx += 1024) {
            const int64_t i = idx / n, j = idx - i * n;
            if (i == k) continue;
            A[idx] = (j == k) ? -colk[i] * p : A[idx] - colk[i] * A[int64_t(k) * n + j];
        }
        __syncthreads();
    }
}

// x = Ainv r, one wave per row
__global__ __launch_bounds__(256) void k_mg_coarse_solve(int n, const double* __restrict__ Ainv,
                                                         const double* __restrict__ r, double* __restrict__ x,
                                                         const int* __restrict__ done) {
    if (done && *done) return;
    const int64_t i = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n) return;
    double s = 0.0;
    for (int j = lane; j < n; j += 64) s += Ainv[i * n + j] * r[j];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
    if (lane == 0) x[i] = s;
}

// ---- preconditioned CG ---------------------------------------------------------
// x += alpha p ; r -= alpha q ; partials r.r
__global__ __launch_bounds__(256) void k_pcg_update(CGArgs C) {
    if (C.S->done) return;
    const double a = C.S->alpha;
    double rr = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < C.n; i += int64_t(gridDim.x) * 256) {
        C.x[i] += a * C.p[i];
        const double r = C.r[i] - a * C.q[i];
        C.r[i] = r;
        rr += r * r;
    }
    block_sum_store(rr, C.part1, blockIdx.x);
}

// the same update, with the start of the hex8 problem level's FP32
// pre-smoothing (x = 0, res = r, first direction D^-1 r / theta) written from
// the updated residual (the V-cycle's own start kernel is then skipped)
__global__ __launch_bounds__(256) void k_pcg_update_mg(CGArgs C, MGLevelArgs L) {
    if (C.S->done) return;
    const double a = C.S->alpha;
    const float ith = float(L.cheb[0]);
    double rr = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < C.n; i += int64_t(gridDim.x) * 256) {
        C.x[i] += a * C.p[i];
        const double r = C.r[i] - a * C.q[i];
        C.r[i] = r;
        rr += r * r;
        const float rf = float(r);
        L.x32[i] = 0.f;
        L.res32[i] = rf;
        L.da32[i] = L.dinv32[i] * rf * ith;
    }
    block_sum_store(rr, C.part1, blockIdx.x);
}

// convergence on the updated residual, before the V-cycle runs
__global__ __launch_bounds__(256) void k_pcg_check(CGArgs C) {
    if (C.S->done) return;
    const double rr = reduce_partials(C.part1, C.nvec);
    if (threadIdx.x == 0) {
        CGScalars* S = C.S;
        S->rr = rr;
        S->iter += 1;
        if (!isfinite(rr)) S->done = 4;
        else if (rr <= S->tol2 * S->bb) S->done = 1;
        else if (S->iter >= S->max_it) S->done = 2;
    }
}

// z = x0 + dlast (the level-0 V-cycle output) ; partials r.z
__global__ __launch_bounds__(256) void k_pcg_rz(CGArgs C, const double* __restrict__ x0,
                                                const double* __restrict__ dlast) {
    if (C.S->done) return;
    double rz = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < C.n; i += int64_t(gridDim.x) * 256) {
        const double z = dlast ? x0[i] + dlast[i] : x0[i];
        C.z[i] = z;
        rz += C.r[i] * z;
    }
    block_sum_store(rz, C.part0, blockIdx.x);
}

__global__ __launch_bounds__(256) void k_pcg_rz32(CGArgs C, const float* __restrict__ x0,
                                                  const float* __restrict__ dlast) {
    if (C.S->done) return;
    double rz = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < C.n; i += int64_t(gridDim.x) * 256) {
        const double z = dlast ? double(x0[i] + dlast[i]) : double(x0[i]);
        C.z[i] = z;
        rz += C.r[i] * z;
    }
    block_sum_store(rz, C.part0, blockIdx.x);
}

// beta = rz_new / rz ; a non-positive r.z means the preconditioner is not SPD
__global__ __launch_bounds__(256) void k_mg_combine(int64_t n, const double* __restrict__ x0,
                                                    const double* __restrict__ dlast, double* __restrict__ out,
                                                    const int* __restrict__ done) {
    if (done && *done) return;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        out[i] = dlast ? x0[i] + dlast[i] : x0[i];
}

__global__ __launch_bounds__(256) void k_pcg_beta(CGArgs C) {
    if (C.S->done) return;
    const double rz = reduce_partials(C.part0, C.nvec);
    if (threadIdx.x == 0) {
        CGScalars* S = C.S;
        if (!(rz > 0.0) || !isfinite(rz)) { S->done = 3; return; }
        S->beta = rz / S->rz;
        S->rz = rz;
    }
}

__global__ __launch_bounds__(256) void k_pcg_start(CGArgs C, const double* __restrict__ b) {
    double bb = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < C.n; i += int64_t(gridDim.x) * 256)
        bb += b[i] * b[i];
    block_sum_store(bb, C.part1, blockIdx.x);
}

__global__ __launch_bounds__(256) void k_pcg_start_scalars(CGArgs C, double tol2, int max_it) {
    const double bb = reduce_partials(C.part1, C.nvec);
    if (threadIdx.x == 0) {
        CGScalars* S = C.S;
        S->bb = bb;
        S->rr = bb;
        S->rz = 0.0;
        S->iter = 0;
        S->tol2 = tol2;
        S->max_it = max_it;
        S->done = (bb == 0.0) ? 1 : (isfinite(bb) ? 0 : 4);
    }
}

// after a restart from the true residual (partials of r.r in part1): converged?
__global__ __launch_bounds__(256) void k_pcg_recheck(CGArgs C) {
    const double rr = reduce_partials(C.part1, C.nvec);
    if (threadIdx.x == 0) {
        CGScalars* S = C.S;
        S->rr = rr;
        if (!isfinite(rr)) S->done = 4;
        else if (rr <= S->tol2 * S->bb) S->done = 1;
        else if (S->iter >= S->max_it) S->done = 2;
        else S->done = 0;
    }
}

__global__ __launch_bounds__(256) void k_pcg_start2(CGArgs C) {
    if (C.S->done) return;
    const double rz = reduce_partials(C.part0, C.nvec);
    if (threadIdx.x == 0) {
        if (!(rz > 0.0) || !isfinite(rz)) C.S->done = 3;
        C.S->rz = rz;
    }
}

// ---- launchers -------------------------------------------------------------------
static inline unsigned mg_blocks(int64_t n, int per) { return unsigned((n + per - 1) / per); }
static inline unsigned mg_grid(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return unsigned(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

void launch_mg_galerkin(const MGGalerkinArgs& g, hipStream_t s) {
    const int64_t n = g.elems ? g.n_elems_list : g.n_coarse;
    if (n <= 0) return;
    const dim3 grid(mg_blocks(n, 4));
    if (g.dim == 2)
        hipLaunchKernelGGL((k_mg_galerkin<2, 5>), grid, dim3(256), 0, s, g);
    else if (g.nav == 1)
        hipLaunchKernelGGL((k_mg_galerkin<3, 1>), grid, dim3(256), 0, s, g);
    else
        hipLaunchKernelGGL((k_mg_galerkin<3, 3>), grid, dim3(256), 0, s, g);
}
void launch_mg_galerkin_class(const MGGalerkinArgs& g, hipStream_t s) {
    if (g.n_coarse <= 0) return;
    hipLaunchKernelGGL(k_mg_galerkin_class, dim3(mg_blocks(g.n_coarse, 4)), dim3(256), 0, s, g);
}
void launch_mg_cheb(const double* lam, double* cheb, int n_levels, int k, double ratio, hipStream_t s) {
    hipLaunchKernelGGL(k_mg_cheb, dim3(1), dim3(64), 0, s, lam, cheb, n_levels, k, ratio);
}
void launch_mg_dense_gather(const MGLevelArgs& L, double* A, hipStream_t s) {
    if (L.nav == 5)
        hipLaunchKernelGGL(k_mg_dense_gather<5>, dim3(mg_blocks(L.n_nodes * 5, 256)), dim3(256), 0, s, L, A);
    else if (L.nav == 1)
        hipLaunchKernelGGL(k_mg_dense_gather<1>, dim3(mg_blocks(L.n_nodes, 256)), dim3(256), 0, s, L, A);
    else
        hipLaunchKernelGGL(k_mg_dense_gather<3>, dim3(mg_blocks(L.n_nodes * 3, 256)), dim3(256), 0, s, L, A);
}
void launch_mg_invert(int n, double* A, hipStream_t s) {
    if (n <= 136)
        hipLaunchKernelGGL(k_mg_invert_lds, dim3(1), dim3(1024), 0, s, n, A);
    else
        hipLaunchKernelGGL(k_mg_invert, dim3(1), dim3(1024), 0, s, n, A);
}
static inline int nb_grid(int64_t n_nodes) {
    const int64_t blocks = (n_nodes * 8 + 255) / 256;
    return int(blocks < 1 ? 1 : (blocks > 8192 ? 8192 : blocks));
}
void launch_mg_rowinfo_nb(int nav, int64_t n_nodes, const int64_t* blk_ptr, const int32_t* blk_col,
                          const double* val, double* dinv, double* part, double* lam, hipStream_t s) {
    const int nb = nb_grid(n_nodes);
    if (nav == 5)
        hipLaunchKernelGGL(k_mg_rowinfo_nb<5>, dim3(nb), dim3(256), 0, s, n_nodes, blk_ptr, blk_col, val, dinv, part);
    else if (nav == 1)
        hipLaunchKernelGGL(k_mg_rowinfo_nb<1>, dim3(nb), dim3(256), 0, s, n_nodes, blk_ptr, blk_col, val, dinv, part);
    else
        hipLaunchKernelGGL(k_mg_rowinfo_nb<3>, dim3(nb), dim3(256), 0, s, n_nodes, blk_ptr, blk_col, val, dinv, part);
    hipLaunchKernelGGL(k_mg_max, dim3(1), dim3(256), 0, s, part, nb, lam);
}
void launch_mg_power(const MGLevelArgs& L, const double* val, int n_it, double safety, double* w0, double* w1,
                     double* part, double* scal, double* lam, hipStream_t s) {
    const int nb = nb_grid(L.n_nodes);
    const int nbs = int(std::min<int64_t>(mg_blocks(L.n_nodes * L.nav, 256), 8192));
    hipLaunchKernelGGL(k_mg_power_start, dim3(nbs), dim3(256), 0, s, L.n_nodes * L.nav, part);
    hipLaunchKernelGGL(k_mg_power_start_norm, dim3(1), dim3(256), 0, s, part, nbs, scal);
    double* in = nullptr;
    double* out = w0;
    for (int it = 0; it < n_it; it++) {
        if (L.nav == 5)
            hipLaunchKernelGGL(k_mg_power<5>, dim3(nb), dim3(256), 0, s, L, val, in, scal, out, part);
        else if (L.nav == 1)
            hipLaunchKernelGGL(k_mg_power<1>, dim3(nb), dim3(256), 0, s, L, val, in, scal, out, part);
        else
            hipLaunchKernelGGL(k_mg_power<3>, dim3(nb), dim3(256), 0, s, L, val, in, scal, out, part);
        hipLaunchKernelGGL(k_mg_power_norm, dim3(1), dim3(256), 0, s, part, nb, it == 0 ? 1 : 0, safety, scal, lam);
        in = out;
        out = out == w0 ? w1 : w0;
    }
}
int64_t lattice_partials(const LatticeOpArgs& l) {
    return (int64_t(l.nx + 1) * (l.ny + 1) * (l.nz + 1) + 255) / 256;
}
void launch_lattice_apply(const CGArgs& c, const double* x, double* y, hipStream_t s) {
    const int* done = c.S ? &c.S->done : nullptr;
    hipLaunchKernelGGL(k_lattice_apply, dim3(unsigned(lattice_partials(c.lat))), dim3(256), 0, s, c, x, y, done);
}
void launch_lattice_rowinfo(const LatticeOpArgs& l, double* dinv, float* dinv32, double* part, double* lam,
                            hipStream_t s) {
    const int nb = int(lattice_partials(l));
    hipLaunchKernelGGL(k_lattice_rowinfo, dim3(nb), dim3(256), 0, s, l, dinv, dinv32, part);
    hipLaunchKernelGGL(k_mg_max, dim3(1), dim3(256), 0, s, part, nb, lam);
}
void launch_perm_nodes(int64_t n_nodes, const int32_t* map, const double* src, double* dst, int scatter,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_perm_nodes, dim3(mg_blocks(n_nodes, 256)), dim3(256), 0, s, n_nodes, map, src, dst, scatter);
}
void launch_mg_pre_init(const MGLevelArgs& L, const int* done, hipStream_t s) {
    if (L.planar) hipLaunchKernelGGL(k_mg_pre_init32, dim3(mg_grid(L.n_nodes * L.nav)), dim3(256), 0, s, L, done);
    else if (L.dinvb) hipLaunchKernelGGL(k_mg_pre_init_blk5, dim3(mg_grid(L.n_nodes * 5)), dim3(256), 0, s, L, done);
    else hipLaunchKernelGGL(k_mg_pre_init, dim3(mg_grid(L.n_nodes * L.nav)), dim3(256), 0, s, L, done);
}
void launch_mg_block_dinv5(const MGLevelArgs& L, float* dinvb, hipStream_t s) {
    hipLaunchKernelGGL(k_mg_block_dinv5, dim3(mg_blocks(L.n_nodes, 256)), dim3(256), 0, s, L, dinvb);
}
void launch_mg_to_stencil(const MGLevelArgs& L, float* sten, hipStream_t s) {
    hipLaunchKernelGGL(k_mg_to_stencil, dim3(mg_blocks(L.n_nodes, 256)), dim3(256), 0, s, L, sten);
}
void launch_mg_gather_stencil3(const MGLevelArgs& L, const int32_t* rank_of_lat, const uint8_t* kpos, float* sten,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_mg_gather_stencil3, dim3(mg_blocks(L.n_nodes, 256)), dim3(256), 0, s, L, rank_of_lat, kpos,
                       sten);
}
void launch_mg_stencil(const MGLevelArgs& L, const float* sten, int mode, int step, const double* din, double* dout,
                       const int* done, hipStream_t s) {
    hipLaunchKernelGGL(k_mg_stencil, dim3(mg_blocks(L.n_nodes, 256)), dim3(256), 0, s, L, sten, mode, step, din,
                       dout, done);
}
void launch_mg_to_stencil27(const MGLevelArgs& L, float* sten, hipStream_t s) {
    hipLaunchKernelGGL(k_mg_to_stencil27, dim3(mg_blocks(L.n_nodes, 256)), dim3(256), 0, s, L, sten);
}
void launch_mg_gather_stencil(const MGLevelArgs& L, const int32_t* rank_of_lat, const uint8_t* kpos, float* sten,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_mg_gather_stencil, dim3(mg_blocks(L.n_nodes, 256)), dim3(256), 0, s, L, rank_of_lat, kpos,
                       sten);
}
void launch_mg_stencil_rank(const MGLevelArgs& L, const float* sten, int mode, int step, const double* din,
                            double* dout, const int* done, hipStream_t s) {
    if (L.val16)
        hipLaunchKernelGGL(k_mg_stencil_heat<uint16_t>, dim3(mg_blocks(L.n_nodes, 256)), dim3(256), 0, s, L, L.val16,
                           mode, step, din, dout, done);
    else
        hipLaunchKernelGGL(k_mg_stencil_heat<float>, dim3(mg_blocks(L.n_nodes, 256)), dim3(256), 0, s, L, sten, mode,
                           step, din, dout, done);
}
void launch_sten27_to_f16(int64_t n, const float* sten, uint16_t* sten16, float* srow, hipStream_t s) {
    hipLaunchKernelGGL(k_sten27_to_f16, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, n, sten, sten16, srow);
}
void launch_mg_power_heat(const MGLevelArgs& L, const float* sten, const int32_t* rank_of_lat, int n_it,
                          double safety, double* w0, double* w1, double* part, double* scal, double* lam,
                          hipStream_t s) {
    const int nb = int(mg_blocks(L.n_nodes, 256));
    const int nbs = int(std::min<int64_t>(mg_blocks(L.n_nodes, 256), 8192));
    hipLaunchKernelGGL(k_mg_power_start, dim3(nbs), dim3(256), 0, s, L.n_nodes, part);
    hipLaunchKernelGGL(k_mg_power_start_norm, dim3(1), dim3(256), 0, s, part, nbs, scal);
    double* in = nullptr;
    double* out = w0;
    for (int it = 0; it < n_it; it++) {
        hipLaunchKernelGGL(k_mg_power_heat, dim3(nb), dim3(256), 0, s, L, sten, rank_of_lat, in, scal, out, part);
        hipLaunchKernelGGL(k_mg_power_norm, dim3(1), dim3(256), 0, s, part, nb, it == 0 ? 1 : 0, safety, scal, lam);
        in = out;
        out = out == w0 ? w1 : w0;
    }
}
void launch_perm_rank_lat(int64_t n_nodes, int nav, const int32_t* rank_of_lat, const double* src, double* dst,
                          int to_lat, const int* done, hipStream_t s) {
    hipLaunchKernelGGL(k_perm_rank_lat, dim3(mg_blocks(n_nodes * nav, 256)), dim3(256), 0, s, n_nodes, nav,
                       rank_of_lat, src, dst, to_lat, done);
}
void launch_mg_combine_rank(int64_t n_nodes, int nav, const int32_t* rank_of_lat, const double* x0,
                            const double* dlast, double* out, const int* done, hipStream_t s) {
    hipLaunchKernelGGL(k_mg_combine_rank, dim3(mg_blocks(n_nodes * nav, 256)), dim3(256), 0, s, n_nodes, nav,
                       rank_of_lat, x0, dlast, out, done);
}
void launch_to_f16_rows(int64_t n_nodes, int nav, const int64_t* blk_ptr, const double* val, uint16_t* val16,
                        float* rs16, hipStream_t s) {
    if (nav != 5) return;
    hipLaunchKernelGGL(k_to_f16_rows<5>, dim3(unsigned((n_nodes * 5 + 255) / 256)), dim3(256), 0, s, n_nodes, blk_ptr,
                       val, val16, rs16);
}
void launch_mg_bsr(const MGLevelArgs& L, int mode, int step, const double* din, double* dout, const int* done,
                   hipStream_t s) {
    const unsigned g = unsigned(nb_grid(L.n_nodes));
    if (L.nav == 5) {
        if (L.val16)
            hipLaunchKernelGGL((k_mg_bsr<uint16_t, 5>), dim3(g), dim3(256), 0, s, L, L.val16, mode, step, din, dout,
                               done);
        else if (L.val32)
            hipLaunchKernelGGL((k_mg_bsr<float, 5>), dim3(g), dim3(256), 0, s, L, L.val32, mode, step, din, dout, done);
        else
            hipLaunchKernelGGL((k_mg_bsr<double, 5>), dim3(g), dim3(256), 0, s, L, L.val, mode, step, din, dout, done);
    } else if (L.nav == 1) {
        if (L.val32)
            hipLaunchKernelGGL((k_mg_bsr<float, 1>), dim3(g), dim3(256), 0, s, L, L.val32, mode, step, din, dout, done);
        else
            hipLaunchKernelGGL((k_mg_bsr<double, 1>), dim3(g), dim3(256), 0, s, L, L.val, mode, step, din, dout, done);
    } else {
        if (L.val32)
            hipLaunchKernelGGL((k_mg_bsr<float, 3>), dim3(g), dim3(256), 0, s, L, L.val32, mode, step, din, dout, done);
        else
            hipLaunchKernelGGL((k_mg_bsr<double, 3>), dim3(g), dim3(256), 0, s, L, L.val, mode, step, din, dout, done);
    }
}
void launch_mg_fine(const CGArgs& c, const MGLevelArgs& L, int mode, int step, const float* din, float* dout,
                    const int* done, hipStream_t s) {
    hipLaunchKernelGGL(k_lattice_smooth, dim3(unsigned(lattice_partials(c.lat))), dim3(256), 0, s, c.lat, L, mode,
                       step, din, dout, done);
}
void launch_mg_restrict(const MGLevelArgs& F, const MGLevelArgs& C, const int* done, hipStream_t s) {
    const dim3 g(mg_blocks(C.n_nodes, 256));
    const int fl = F.planar ? FL_PLANAR : F.rank_of_lat ? FL_RANK : FL_LAT;
    if (F.dim == 2) {
        if (fl == FL_RANK) hipLaunchKernelGGL((k_mg_restrict<2, 5, FL_RANK>), g, dim3(256), 0, s, F, C, done);
        else hipLaunchKernelGGL((k_mg_restrict<2, 5, FL_LAT>), g, dim3(256), 0, s, F, C, done);
    } else if (F.nav == 1) {
        if (fl == FL_RANK) hipLaunchKernelGGL((k_mg_restrict<3, 1, FL_RANK>), g, dim3(256), 0, s, F, C, done);
        else hipLaunchKernelGGL((k_mg_restrict<3, 1, FL_LAT>), g, dim3(256), 0, s, F, C, done);
    } else {
        if (fl == FL_PLANAR) hipLaunchKernelGGL((k_mg_restrict<3, 3, FL_PLANAR>), g, dim3(256), 0, s, F, C, done);
        else if (fl == FL_RANK) hipLaunchKernelGGL((k_mg_restrict<3, 3, FL_RANK>), g, dim3(256), 0, s, F, C, done);
        else hipLaunchKernelGGL((k_mg_restrict<3, 3, FL_LAT>), g, dim3(256), 0, s, F, C, done);
    }
}
void launch_mg_prolong(const MGLevelArgs& F, const MGLevelArgs& C, const double* dlast, const int* done,
                       hipStream_t s) {
    const dim3 g(mg_blocks(F.n_nodes, 256));
    const int fl = F.planar ? FL_PLANAR : F.rank_of_lat ? FL_RANK : FL_LAT;
    if (F.dim == 2) {
        if (fl == FL_RANK) hipLaunchKernelGGL((k_mg_prolong<2, 5, FL_RANK>), g, dim3(256), 0, s, F, C, dlast, done);
        else hipLaunchKernelGGL((k_mg_prolong<2, 5, FL_LAT>), g, dim3(256), 0, s, F, C, dlast, done);
    } else if (F.nav == 1) {
        if (fl == FL_RANK) hipLaunchKernelGGL((k_mg_prolong<3, 1, FL_RANK>), g, dim3(256), 0, s, F, C, dlast, done);
        else hipLaunchKernelGGL((k_mg_prolong<3, 1, FL_LAT>), g, dim3(256), 0, s, F, C, dlast, done);
    } else {
        if (fl == FL_PLANAR)
            hipLaunchKernelGGL((k_mg_prolong<3, 3, FL_PLANAR>), g, dim3(256), 0, s, F, C, dlast, done);
        else if (fl == FL_RANK)
            hipLaunchKernelGGL((k_mg_prolong<3, 3, FL_RANK>), g, dim3(256), 0, s, F, C, dlast, done);
        else hipLaunchKernelGGL((k_mg_prolong<3, 3, FL_LAT>), g, dim3(256), 0, s, F, C, dlast, done);
    }
}
void launch_mg_coarse_solve(int n, const double* Ainv, const double* r, double* x, const int* done,
                            hipStream_t s) {
    hipLaunchKernelGGL(k_mg_coarse_solve, dim3(mg_blocks(n, 4)), dim3(256), 0, s, n, Ainv, r, x, done);
}
void launch_pcg_update(const CGArgs& c, hipStream_t s) {
    hipLaunchKernelGGL(k_pcg_update, dim3(c.nvec), dim3(256), 0, s, c);
    hipLaunchKernelGGL(k_pcg_check, dim3(1), dim3(256), 0, s, c);
}
void launch_pcg_update_mg(const CGArgs& c, const MGLevelArgs& L, hipStream_t s) {
    hipLaunchKernelGGL(k_pcg_update_mg, dim3(c.nvec), dim3(256), 0, s, c, L);
    hipLaunchKernelGGL(k_pcg_check, dim3(1), dim3(256), 0, s, c);
}
void launch_pcg_rz(const CGArgs& c, const double* x0, const double* dlast, hipStream_t s) {
    hipLaunchKernelGGL(k_pcg_rz, dim3(c.nvec), dim3(256), 0, s, c, x0, dlast);
}
void launch_pcg_rz32(const CGArgs& c, const float* x0, const float* dlast, hipStream_t s) {
    hipLaunchKernelGGL(k_pcg_rz32, dim3(c.nvec), dim3(256), 0, s, c, x0, dlast);
}
void launch_mg_combine(int64_t n, const double* x0, const double* dlast, double* out, const int* done,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_mg_combine, dim3(mg_grid(n)), dim3(256), 0, s, n, x0, dlast, out, done);
}
void launch_pcg_beta(const CGArgs& c, hipStream_t s) {
    hipLaunchKernelGGL(k_pcg_beta, dim3(1), dim3(256), 0, s, c);
}
void launch_pcg_start(const CGArgs& c, const double* b, double tol2, int max_it, hipStream_t s) {
    hipLaunchKernelGGL(k_pcg_start, dim3(c.nvec), dim3(256), 0, s, c, b);
    hipLaunchKernelGGL(k_pcg_start_scalars, dim3(1), dim3(256), 0, s, c, tol2, max_it);
}
void launch_pcg_recheck(const CGArgs& c, hipStream_t s) {
    hipLaunchKernelGGL(k_pcg_start, dim3(c.nvec), dim3(256), 0, s, c, c.r);
    hipLaunchKernelGGL(k_pcg_recheck, dim3(1), dim3(256), 0, s, c);
}
void launch_pcg_start2(const CGArgs& c, hipStream_t s) {
    hipLaunchKernelGGL(k_pcg_start2, dim3(1), dim3(256), 0, s, c);
}

}  // namespace mastg
