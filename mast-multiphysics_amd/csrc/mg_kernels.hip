// Geometric multigrid for hex8 elasticity on lattice meshes: Galerkin coarse
// element matrices, Chebyshev/Jacobi smoothing fused into the level operators,
// trilinear restriction / prolongation, dense coarsest solve, and the
// preconditioned CG kernels around the V-cycle.
//
// Every reduction is a fixed-order sum (no atomics), so a solve is bitwise
// reproducible for a given launch configuration.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"
#include "device_util.h"

namespace mastg {

// libMesh hex8 local node -> cell corner offset
__constant__ int8_t kMGCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                       {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

// trilinear weight of coarse corner A at fine local node a of child cell c
// (c bit d = the child's offset along dimension d inside the coarse cell)
__device__ __forceinline__ double mg_q(int c, int a, int A) {
    double w = 1.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int t = ((c >> d) & 1) + kMGCorner[a][d];   // 0, 1, 2 fine cells
        w *= kMGCorner[A][d] ? 0.5 * t : 1.0 - 0.5 * t;
    }
    return w;
}

// ---- Galerkin coarse element matrices: K_E = sum_c Q_c^T K_c Q_c ---------
// One wave per coarse element, 4 per block.  Fine blocks are masked (the rows
// and columns of constrained fine DOFs removed, as P has no row for them), the
// result is made exactly symmetric and gets the libMesh constraint convention
// for constrained coarse DOFs (rows / columns zero, unit diagonal).
__global__ __launch_bounds__(256) void k_mg_galerkin(MGGalerkinArgs G) {
    __shared__ double sK[4][64 * 9];
    __shared__ double sT[4][64 * 9];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int hi = lane >> 3, lo = lane & 7;
    const int64_t E = int64_t(blockIdx.x) * 4 + wave;
    const bool valid = E < G.n_coarse;
    const int64_t Ev = valid ? E : 0;
    const int64_t I = Ev % G.cnx, J = (Ev / G.cnx) % G.cny, K = Ev / (int64_t(G.cnx) * G.cny);
    double acc[9];
#pragma unroll
    for (int i = 0; i < 9; i++) acc[i] = 0.0;
    for (int c = 0; c < 8; c++) {
        const int64_t fi = 2 * I + (c & 1), fj = 2 * J + ((c >> 1) & 1), fk = 2 * K + ((c >> 2) & 1);
        const int64_t cell = fi + int64_t(G.fnx) * (fj + int64_t(G.fny) * fk);
        const int64_t fe = G.f_elem_of_cell ? int64_t(G.f_elem_of_cell[cell]) : cell;
        {   // masked fine block (a, b) = (hi, lo)
            const int32_t ra = G.f_econn[fe * 8 + hi], rb = G.f_econn[fe * 8 + lo];
            const double* kb = G.Kf + (fe * 64 + lane) * 9;
#pragma unroll
            for (int v = 0; v < 3; v++)
#pragma unroll
                for (int w = 0; w < 3; w++) {
                    double x = kb[v * 3 + w];
                    if (G.f_con[int64_t(ra) * 3 + v] || G.f_con[int64_t(rb) * 3 + w]) x = 0.0;
                    sK[wave][lane * 9 + v * 3 + w] = x;
                }
        }
        __syncthreads();
        {   // T[a][B] = sum_b K[a][b] Q_c[b][B], lane = (a, B)
            double t[9];
#pragma unroll
            for (int i = 0; i < 9; i++) t[i] = 0.0;
            for (int b = 0; b < 8; b++) {
                const double q = mg_q(c, b, lo);
                if (q != 0.0) {
                    const double* kb = &sK[wave][(hi * 8 + b) * 9];
#pragma unroll
                    for (int i = 0; i < 9; i++) t[i] += kb[i] * q;
                }
            }
#pragma unroll
            for (int i = 0; i < 9; i++) sT[wave][lane * 9 + i] = t[i];
        }
        __syncthreads();
        // K_E[A][B] += sum_a Q_c[a][A] T[a][B], lane = (A, B)
        for (int a = 0; a < 8; a++) {
            const double q = mg_q(c, a, hi);
            if (q != 0.0) {
                const double* tb = &sT[wave][(a * 8 + lo) * 9];
#pragma unroll
                for (int i = 0; i < 9; i++) acc[i] += q * tb[i];
            }
        }
        __syncthreads();
    }
    // exact symmetry: lanes below the diagonal take the transpose of their partner
#pragma unroll
    for (int i = 0; i < 9; i++) sK[wave][lane * 9 + i] = acc[i];
    __syncthreads();
    double out[9];
    if (hi > lo) {
        const double* tb = &sK[wave][(lo * 8 + hi) * 9];
#pragma unroll
        for (int v = 0; v < 3; v++)
#pragma unroll
            for (int w = 0; w < 3; w++) out[v * 3 + w] = tb[w * 3 + v];
    } else if (hi == lo) {
#pragma unroll
        for (int v = 0; v < 3; v++)
#pragma unroll
            for (int w = 0; w < 3; w++) out[v * 3 + w] = 0.5 * (acc[v * 3 + w] + acc[w * 3 + v]);
    } else {
#pragma unroll
        for (int i = 0; i < 9; i++) out[i] = acc[i];
    }
    if (!valid) return;
    const int32_t cA = G.c_econn[E * 8 + hi], cB = G.c_econn[E * 8 + lo];
    double* o = G.Kc + (E * 64 + lane) * 9;
#pragma unroll
    for (int v = 0; v < 3; v++)
#pragma unroll
        for (int w = 0; w < 3; w++) {
            double x = out[v * 3 + w];
            if (G.c_con[int64_t(cA) * 3 + v] || G.c_con[int64_t(cB) * 3 + w]) x = (hi == lo && v == w) ? 1.0 : 0.0;
            o[v * 3 + w] = x;
        }
}

// ---- row information: dinv and the Gershgorin bound of D^-1 A -------------
__global__ __launch_bounds__(256) void k_mg_rowinfo_bsr(MGLevelArgs L, double* __restrict__ dinv,
                                                        double* __restrict__ part) {
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    double g = 0.0;
    if (i < L.n_nodes * 3) {
        const int64_t r = i / 3;
        const int v = int(i - r * 3);
        const int64_t k0 = L.blk_ptr[r];
        const int64_t deg = L.blk_ptr[r + 1] - k0;
        const double* row = L.val + k0 * 9 + int64_t(v) * deg * 3;
        double sum = 0.0, diag = 0.0;
        for (int64_t k = 0; k < deg; k++) {
            const int32_t c = L.blk_col[k0 + k];
#pragma unroll
            for (int w = 0; w < 3; w++) {
                const double a = row[k * 3 + w];
                sum += fabs(a);
                if (c == r && w == v) diag = a;
            }
        }
        dinv[i] = diag > 0.0 ? 1.0 / diag : 0.0;
        g = diag > 0.0 ? sum / diag : 0.0;
    }
    block_max_store(g, part, blockIdx.x);
}

__global__ __launch_bounds__(256) void k_mg_rowinfo_csr(int64_t n, const int64_t* __restrict__ row_ptr,
                                                        const double* __restrict__ val,
                                                        const int64_t* __restrict__ diag_pos,
                                                        double* __restrict__ part) {
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    double g = 0.0;
    if (i < n) {
        double sum = 0.0;
        for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) sum += fabs(val[k]);
        const double diag = val[diag_pos[i]];
        g = diag > 0.0 ? sum / diag : 0.0;
    }
    block_max_store(g, part, blockIdx.x);
}

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
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < L.n_nodes * 3; i += int64_t(gridDim.x) * 256) {
        const double r = L.r[i];
        L.x[i] = 0.0;
        L.res[i] = r;
        L.da[i] = L.dinv[i] * r * ith;
    }
}

// node-block operator of a coarse level (values T = double or their FP32
// copy; products and sums in FP64), 8 lanes per node, fused epilogue
template <typename T>
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
        const T* __restrict__ V = vals + k0 * 9;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int k = sub; k < deg; k += 8) {
            const int64_t cb = int64_t(L.blk_col[k0 + k]) * 3;
            const double x0 = in[cb], x1 = in[cb + 1], x2 = in[cb + 2];
            const T* v0 = V + int64_t(k) * 3;
            const T* v1 = V + (int64_t(deg) + k) * 3;
            const T* v2 = V + (int64_t(2 * deg) + k) * 3;
            s0 += double(v0[0]) * x0 + double(v0[1]) * x1 + double(v0[2]) * x2;
            s1 += double(v1[0]) * x0 + double(v1[1]) * x1 + double(v1[2]) * x2;
            s2 += double(v2[0]) * x0 + double(v2[1]) * x1 + double(v2[2]) * x2;
        }
#pragma unroll
        for (int m = 4; m >= 1; m >>= 1) {
            s0 += __shfl_xor(s0, m, 8);
            s1 += __shfl_xor(s1, m, 8);
            s2 += __shfl_xor(s2, m, 8);
        }
        if (sub < 3) {
            const double y = sub == 0 ? s0 : (sub == 1 ? s1 : s2);
            mg_epilogue(L, mode, r * 3 + sub, y, din, dout, c1, c2, ith);
        }
    }
}

// ---- level 0 on a lattice: per-node matrix-free operator -------------------
// Thread per lattice point n, vectors in lattice order (x fastest), so the
// neighbour gathers of a wave are contiguous; for each existing cell at n
// (local corner a) y_n += s_e Khat[a][b] x_b.  Khat indices are uniform
// across the wave, so the class matrix is read with scalar-cache loads.  Each
// y_n is a fixed-order sum: results are deterministic.
__host__ __device__ constexpr int lat_cx(int b) { return (b == 1 || b == 2 || b == 5 || b == 6) ? 1 : 0; }
__host__ __device__ constexpr int lat_cy(int b) { return (b == 2 || b == 3 || b == 6 || b == 7) ? 1 : 0; }
__host__ __device__ constexpr int lat_cz(int b) { return b >= 4 ? 1 : 0; }
__host__ __device__ constexpr int lat_corner(int dx, int dy, int dz) {
    return dz * 4 + (dy ? (dx ? 2 : 3) : (dx ? 1 : 0));
}

// KIND 0: y = A x and partial x.y (CG operator); KIND 1: multigrid smoothing
// step with the row-local epilogue
template <int KIND>
__global__ __launch_bounds__(256) void k_lattice_apply(CGArgs C, const double* __restrict__ xcg,
                                                       double* __restrict__ ycg, MGLevelArgs L, int mode, int step,
                                                       const double* __restrict__ din, double* __restrict__ dout,
                                                       const int* __restrict__ done) {
    if (done && *done) return;
    const LatticeOpArgs& O = C.lat;
    const int64_t nxp = O.nx + 1, nyp = O.ny + 1;
    const int64_t n_lat = nxp * nyp * (O.nz + 1);
    const int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x;
    const bool valid = p < n_lat;
    const int64_t pv = valid ? p : 0;
    const int i = int(pv % nxp), j = int((pv / nxp) % nyp), k = int(pv / (nxp * nyp));
    const double* __restrict__ in = KIND == 0 ? xcg : (mode == MG_RESID_INIT ? L.x : din);
    // cells at n (c bit d: the cell extends towards +d from n - 1) and their
    // scales; a missing cell gets the scale 0 (its index is clamped into the
    // lattice so every load stays in bounds and unconditional)
    double sc[8];
    int cnt = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
        const int ci = i - 1 + cx, cj = j - 1 + cy, ck = k - 1 + cz;
        const bool has = valid && ci >= 0 && ci < O.nx && cj >= 0 && cj < O.ny && ck >= 0 && ck < O.nz;
        const int64_t cell = min(max(ci, 0), O.nx - 1) +
                             int64_t(O.nx) * (min(max(cj, 0), O.ny - 1) + int64_t(O.ny) * min(max(ck, 0), O.nz - 1));
        const double sv = O.sE[O.elem_of_cell[cell]];
        sc[c] = has ? sv : 0.0;
        cnt += has ? 1 : 0;
    }
    // neighbour-outer, one z-layer of 9 neighbours at a time (bounded live
    // ranges): each neighbour value is read once and feeds every cell shared
    // with n, y += s_c Khat[a_c][b] x_o (scaled per block product, so only y
    // is carried: fewer registers, more waves in flight to hide the gathers).
    // Neighbours outside the lattice are clamped into it; every cell holding
    // one is missing and scaled by 0.
    double y0 = 0.0, y1 = 0.0, y2 = 0.0;
#pragma unroll 1
    for (int lz = 0; lz < 3; lz++) {
        const int mk = min(max(k + lz - 1, 0), O.nz);
        double xv[9][3];
#pragma unroll
        for (int q = 0; q < 9; q++) {
            const int mi = min(max(i + (q % 3) - 1, 0), O.nx);
            const int mj = min(max(j + (q / 3) - 1, 0), O.ny);
            const int64_t rm = (mi + nxp * (mj + nyp * int64_t(mk))) * 3;
            xv[q][0] = in[rm];
            xv[q][1] = in[rm + 1];
            xv[q][2] = in[rm + 2];
        }
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
            const int bz = lz - cz;                      // uniform across the wave
            if (bz < 0 || bz > 1) continue;
            const int a = lat_corner(1 - cx, 1 - cy, 1 - cz);
            double t0 = 0.0, t1 = 0.0, t2 = 0.0;
#pragma unroll
            for (int q = 0; q < 9; q++) {
                const int bx = (q % 3) - cx, by = (q / 3) - cy;
                if (bx < 0 || bx > 1 || by < 0 || by > 1) continue;   // compile time
                const int b = bz * 4 + lat_corner(bx, by, 0);
                const double* K = O.Khat + (a * 8 + b) * 9;
                t0 += K[0] * xv[q][0] + K[1] * xv[q][1] + K[2] * xv[q][2];
                t1 += K[3] * xv[q][0] + K[4] * xv[q][1] + K[5] * xv[q][2];
                t2 += K[6] * xv[q][0] + K[7] * xv[q][1] + K[8] * xv[q][2];
            }
            y0 += sc[c] * t0;
            y1 += sc[c] * t1;
            y2 += sc[c] * t2;
        }
    }
    const int64_t r = pv;
    double dsum = 0.0;
    if (valid) {
        const int m = O.nodecon[r];
        if (m & 1) y0 = cnt * in[r * 3 + 0];
        if (m & 2) y1 = cnt * in[r * 3 + 1];
        if (m & 4) y2 = cnt * in[r * 3 + 2];
        if (KIND == 0) {
            ycg[r * 3 + 0] = y0;
            ycg[r * 3 + 1] = y1;
            ycg[r * 3 + 2] = y2;
            dsum = y0 * in[r * 3 + 0] + y1 * in[r * 3 + 1] + y2 * in[r * 3 + 2];
        } else {
            const double ith = L.cheb[0];
            const double c1 = step > 0 ? L.cheb[2 * step] : 0.0, c2 = step > 0 ? L.cheb[2 * step + 1] : 0.0;
            mg_epilogue(L, mode, r * 3 + 0, y0, din, dout, c1, c2, ith);
            mg_epilogue(L, mode, r * 3 + 1, y1, din, dout, c1, c2, ith);
            mg_epilogue(L, mode, r * 3 + 2, y2, din, dout, c1, c2, ith);
        }
    }
    if (KIND == 0) block_sum_store(dsum, C.part0, blockIdx.x);
}

// node-vector permutation between rank order and lattice order (3 values per
// node): gather dst[p] = src[map[p]], or scatter dst[map[p]] = src[p]
__global__ __launch_bounds__(256) void k_perm_nodes(int64_t n_nodes, const int32_t* __restrict__ map,
                                                    const double* __restrict__ src, double* __restrict__ dst,
                                                    int scatter) {
    const int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (p >= n_nodes) return;
    const int64_t m = int64_t(map[p]) * 3, q = p * 3;
    if (scatter) {
        dst[m] = src[q]; dst[m + 1] = src[q + 1]; dst[m + 2] = src[q + 2];
    } else {
        dst[q] = src[m]; dst[q + 1] = src[m + 1]; dst[q + 2] = src[m + 2];
    }
}

// dinv (when wanted) and the Gershgorin bound over node blocks, 8 lanes per node
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
        const double* __restrict__ V = val + k0 * 9;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, d0 = 0.0, d1 = 0.0, d2 = 0.0;
        for (int kk = sub; kk < deg; kk += 8) {
            const double* v0 = V + int64_t(kk) * 3;
            const double* v1 = V + (int64_t(deg) + kk) * 3;
            const double* v2 = V + (int64_t(2 * deg) + kk) * 3;
            s0 += fabs(v0[0]) + fabs(v0[1]) + fabs(v0[2]);
            s1 += fabs(v1[0]) + fabs(v1[1]) + fabs(v1[2]);
            s2 += fabs(v2[0]) + fabs(v2[1]) + fabs(v2[2]);
            if (blk_col[k0 + kk] == r) { d0 = v0[0]; d1 = v1[1]; d2 = v2[2]; }
        }
#pragma unroll
        for (int m = 4; m >= 1; m >>= 1) {
            s0 += __shfl_xor(s0, m, 8); s1 += __shfl_xor(s1, m, 8); s2 += __shfl_xor(s2, m, 8);
            d0 += __shfl_xor(d0, m, 8); d1 += __shfl_xor(d1, m, 8); d2 += __shfl_xor(d2, m, 8);
        }
        if (sub == 0) {
            g = fmax(g, d0 > 0.0 ? s0 / d0 : 0.0);
            g = fmax(g, d1 > 0.0 ? s1 / d1 : 0.0);
            g = fmax(g, d2 > 0.0 ? s2 / d2 : 0.0);
            if (dinv) {
                dinv[r * 3 + 0] = d0 > 0.0 ? 1.0 / d0 : 0.0;
                dinv[r * 3 + 1] = d1 > 0.0 ? 1.0 / d1 : 0.0;
                dinv[r * 3 + 2] = d2 > 0.0 ? 1.0 / d2 : 0.0;
            }
        }
    }
    block_max_store(g, part, blockIdx.x);
}

// Gauss-Jordan inverse held in LDS (n <= 136), one block
__global__ __launch_bounds__(1024) void k_mg_invert_lds(int n, double* __restrict__ A) {
    __shared__ double sA[136 * 136];
    __shared__ double colk[136];
    const int nn = n * n;
    for (int idx = threadIdx.x; idx < nn; idx += 1024) sA[idx] = A[idx];
    __syncthreads();
    for (int k = 0; k < n; k++) {
        for (int i = threadIdx.x; i < n; i += 1024) colk[i] = sA[i * n + k];
        __syncthreads();
        const double p = 1.0 / colk[k];
        for (int j = threadIdx.x; j < n; j += 1024) sA[k * n + j] = (j == k) ? p : sA[k * n + j] * p;
        __syncthreads();
        for (int idx = threadIdx.x; idx < nn; idx += 1024) {
            const int i = idx / n, j = idx - i * n;
            if (i == k) continue;
            sA[idx] = (j == k) ? -colk[i] * p : sA[idx] - colk[i] * sA[k * n + j];
        }
        __syncthreads();
    }
    for (int idx = threadIdx.x; idx < nn; idx += 1024) A[idx] = sA[idx];
}

// ---- transfers ---------------------------------------------------------------
// r_C = sum over the 27 fine points around 2C of w * res_f (w = prod (1 or 1/2)),
// constrained fine DOFs skipped, constrained coarse DOFs zero (P^T)
__global__ __launch_bounds__(256) void k_mg_restrict(MGLevelArgs F, MGLevelArgs Cl, const int* __restrict__ done) {
    if (done && *done) return;
    const int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (p >= Cl.n_nodes) return;
    const int64_t cx = Cl.nx + 1, cy = Cl.ny + 1;
    const int64_t I = p % cx, J = (p / cx) % cy, K = p / (cx * cy);
    const int64_t fx = F.nx + 1, fy = F.ny + 1;
    // all 27 loads are independent (unrolled, predicated) so they are in flight together
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int o = 0; o < 27; o++) {
        const int oi = (o % 3) - 1, oj = ((o / 3) % 3) - 1, ok = (o / 9) - 1;
        const int64_t fi = 2 * I + oi, fj = 2 * J + oj, fk = 2 * K + ok;
        const bool in_lat = fi >= 0 && fi <= F.nx && fj >= 0 && fj <= F.ny && fk >= 0 && fk <= F.nz;
        // out-of-lattice points are clamped (valid address) and weighted 0;
        // constrained fine DOFs are weighted 0 (P has no row for them)
        const double w = in_lat ? (ok ? 0.5 : 1.0) * (oj ? 0.5 : 1.0) * (oi ? 0.5 : 1.0) : 0.0;
        const int64_t fp = (in_lat ? fi : 0) + fx * ((in_lat ? fj : 0) + fy * (in_lat ? fk : 0));
        const int64_t rf = fp * 3;
        s0 += (w * (1.0 - double(F.con[rf + 0]))) * F.res[rf + 0];
        s1 += (w * (1.0 - double(F.con[rf + 1]))) * F.res[rf + 1];
        s2 += (w * (1.0 - double(F.con[rf + 2]))) * F.res[rf + 2];
    }
    const int64_t rc = p * 3;
    Cl.r[rc + 0] = Cl.con[rc + 0] ? 0.0 : s0;
    Cl.r[rc + 1] = Cl.con[rc + 1] ? 0.0 : s1;
    Cl.r[rc + 2] = Cl.con[rc + 2] ? 0.0 : s2;
}

// x_f += P (x_C + dlast_C) at unconstrained fine DOFs (trilinear interpolation)
__global__ __launch_bounds__(256) void k_mg_prolong(MGLevelArgs F, MGLevelArgs Cl, const double* __restrict__ dlast,
                                                    const int* __restrict__ done) {
    if (done && *done) return;
    const int64_t p = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (p >= F.n_nodes) return;
    const int64_t fx = F.nx + 1, fy = F.ny + 1;
    const int64_t fi = p % fx, fj = (p / fx) % fy, fk = p / (fx * fy);
    const int64_t cx = Cl.nx + 1, cy = Cl.ny + 1;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int a = q & 1, b = (q >> 1) & 1, c = q >> 2;
        // an even fine index takes its one coarse point (a = 0), an odd one both
        // neighbours; the unused corner of an even index is clamped and weighted 0
        const bool use = !((a && !(fi & 1)) || (b && !(fj & 1)) || (c && !(fk & 1)));
        const double w = use ? ((fi & 1) ? 0.5 : 1.0) * ((fj & 1) ? 0.5 : 1.0) * ((fk & 1) ? 0.5 : 1.0) : 0.0;
        const int64_t CI = min((fi >> 1) + a, int64_t(Cl.nx)), CJ = min((fj >> 1) + b, int64_t(Cl.ny)),
                      CK = min((fk >> 1) + c, int64_t(Cl.nz));
        const int64_t cp = CI + cx * (CJ + cy * CK);
        const int64_t rc = cp * 3;
        if (dlast) {
            s0 += w * (Cl.x[rc + 0] + dlast[rc + 0]);
            s1 += w * (Cl.x[rc + 1] + dlast[rc + 1]);
            s2 += w * (Cl.x[rc + 2] + dlast[rc + 2]);
        } else {
            s0 += w * Cl.x[rc + 0];
            s1 += w * Cl.x[rc + 1];
            s2 += w * Cl.x[rc + 2];
        }
    }
    const int64_t rf = p * 3;
    if (!F.con[rf + 0]) F.x[rf + 0] += s0;
    if (!F.con[rf + 1]) F.x[rf + 1] += s1;
    if (!F.con[rf + 2]) F.x[rf + 2] += s2;
}

// ---- coarsest level: dense inverse -------------------------------------------
__global__ __launch_bounds__(256) void k_mg_dense_gather(MGLevelArgs L, double* __restrict__ A) {
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    const int64_t n = L.n_nodes * 3;
    if (i >= n) return;
    const int64_t r = i / 3;
    const int v = int(i - r * 3);
    const int64_t k0 = L.blk_ptr[r];
    const int64_t deg = L.blk_ptr[r + 1] - k0;
    const double* row = L.val + k0 * 9 + int64_t(v) * deg * 3;
    for (int64_t k = 0; k < deg; k++) {
        const int64_t c = L.blk_col[k0 + k];
#pragma unroll
        for (int w = 0; w < 3; w++) A[i * n + c * 3 + w] = row[k * 3 + w];
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

// beta = rz_new / rz ; a non-positive r.z means the preconditioner is not SPD
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
    if (g.n_coarse <= 0) return;
    hipLaunchKernelGGL(k_mg_galerkin, dim3(mg_blocks(g.n_coarse, 4)), dim3(256), 0, s, g);
}
void launch_mg_rowinfo_bsr(const MGLevelArgs& L, double* dinv, double* part, double* lam, hipStream_t s) {
    const unsigned nb = mg_blocks(L.n_nodes * 3, 256);
    hipLaunchKernelGGL(k_mg_rowinfo_bsr, dim3(nb), dim3(256), 0, s, L, dinv, part);
    hipLaunchKernelGGL(k_mg_max, dim3(1), dim3(256), 0, s, part, int(nb), lam);
}
void launch_mg_rowinfo_csr(int64_t n, const int64_t* row_ptr, const double* val, const int64_t* diag_pos,
                           double* part, double* lam, hipStream_t s) {
    const unsigned nb = mg_blocks(n, 256);
    hipLaunchKernelGGL(k_mg_rowinfo_csr, dim3(nb), dim3(256), 0, s, n, row_ptr, val, diag_pos, part);
    hipLaunchKernelGGL(k_mg_max, dim3(1), dim3(256), 0, s, part, int(nb), lam);
}
void launch_mg_cheb(const double* lam, double* cheb, int n_levels, int k, double ratio, hipStream_t s) {
    hipLaunchKernelGGL(k_mg_cheb, dim3(1), dim3(64), 0, s, lam, cheb, n_levels, k, ratio);
}
void launch_mg_dense_gather(const MGLevelArgs& L, double* A, hipStream_t s) {
    hipLaunchKernelGGL(k_mg_dense_gather, dim3(mg_blocks(L.n_nodes * 3, 256)), dim3(256), 0, s, L, A);
}
void launch_mg_invert(int n, double* A, hipStream_t s) {
    if (n <= 136)
        hipLaunchKernelGGL(k_mg_invert_lds, dim3(1), dim3(1024), 0, s, n, A);
    else
        hipLaunchKernelGGL(k_mg_invert, dim3(1), dim3(1024), 0, s, n, A);
}
void launch_mg_rowinfo_nb(int64_t n_nodes, const int64_t* blk_ptr, const int32_t* blk_col, const double* val,
                          double* dinv, double* part, double* lam, hipStream_t s) {
    const int64_t blocks = (n_nodes * 8 + 255) / 256;
    const int nb = int(blocks < 1 ? 1 : (blocks > 8192 ? 8192 : blocks));
    hipLaunchKernelGGL(k_mg_rowinfo_nb, dim3(nb), dim3(256), 0, s, n_nodes, blk_ptr, blk_col, val, dinv, part);
    hipLaunchKernelGGL(k_mg_max, dim3(1), dim3(256), 0, s, part, nb, lam);
}
int64_t lattice_partials(const LatticeOpArgs& l) {
    return (int64_t(l.nx + 1) * (l.ny + 1) * (l.nz + 1) + 255) / 256;
}
void launch_lattice_apply(const CGArgs& c, const double* x, double* y, hipStream_t s) {
    MGLevelArgs none{};
    const int* done = c.S ? &c.S->done : nullptr;
    hipLaunchKernelGGL(k_lattice_apply<0>, dim3(unsigned(lattice_partials(c.lat))), dim3(256), 0, s, c, x, y, none, 0,
                       0, nullptr, nullptr, done);
}
void launch_perm_nodes(int64_t n_nodes, const int32_t* map, const double* src, double* dst, int scatter,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_perm_nodes, dim3(mg_blocks(n_nodes, 256)), dim3(256), 0, s, n_nodes, map, src, dst, scatter);
}
void launch_mg_pre_init(const MGLevelArgs& L, const int* done, hipStream_t s) {
    hipLaunchKernelGGL(k_mg_pre_init, dim3(mg_grid(L.n_nodes * 3)), dim3(256), 0, s, L, done);
}
void launch_mg_bsr(const MGLevelArgs& L, int mode, int step, const double* din, double* dout, const int* done,
                   hipStream_t s) {
    const int64_t blocks = (L.n_nodes * 8 + 255) / 256;
    const unsigned g = unsigned(blocks < 1 ? 1 : (blocks > 8192 ? 8192 : blocks));
    if (L.val32)
        hipLaunchKernelGGL(k_mg_bsr<float>, dim3(g), dim3(256), 0, s, L, L.val32, mode, step, din, dout, done);
    else
        hipLaunchKernelGGL(k_mg_bsr<double>, dim3(g), dim3(256), 0, s, L, L.val, mode, step, din, dout, done);
}
void launch_mg_fine(const CGArgs& c, const MGLevelArgs& L, int mode, int step, const double* din,
                    double* dout, const int* done, hipStream_t s) {
    hipLaunchKernelGGL(k_lattice_apply<1>, dim3(unsigned(lattice_partials(c.lat))), dim3(256), 0, s, c, nullptr,
                       nullptr, L, mode, step, din, dout, done);
}
void launch_mg_restrict(const MGLevelArgs& F, const MGLevelArgs& C, const int* done, hipStream_t s) {
    hipLaunchKernelGGL(k_mg_restrict, dim3(mg_blocks(C.n_nodes, 256)), dim3(256), 0, s, F, C, done);
}
void launch_mg_prolong(const MGLevelArgs& F, const MGLevelArgs& C, const double* dlast, const int* done,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_mg_prolong, dim3(mg_blocks(F.n_nodes, 256)), dim3(256), 0, s, F, C, dlast, done);
}
void launch_mg_coarse_solve(int n, const double* Ainv, const double* r, double* x, const int* done,
                            hipStream_t s) {
    hipLaunchKernelGGL(k_mg_coarse_solve, dim3(mg_blocks(n, 4)), dim3(256), 0, s, n, Ainv, r, x, done);
}
void launch_pcg_update(const CGArgs& c, hipStream_t s) {
    hipLaunchKernelGGL(k_pcg_update, dim3(c.nvec), dim3(256), 0, s, c);
    hipLaunchKernelGGL(k_pcg_check, dim3(1), dim3(256), 0, s, c);
}
void launch_pcg_rz(const CGArgs& c, const double* x0, const double* dlast, hipStream_t s) {
    hipLaunchKernelGGL(k_pcg_rz, dim3(c.nvec), dim3(256), 0, s, c, x0, dlast);
}
void launch_pcg_beta(const CGArgs& c, hipStream_t s) {
    hipLaunchKernelGGL(k_pcg_beta, dim3(1), dim3(256), 0, s, c);
}
void launch_pcg_start(const CGArgs& c, const double* b, double tol2, int max_it, hipStream_t s) {
    hipLaunchKernelGGL(k_pcg_start, dim3(c.nvec), dim3(256), 0, s, c, b);
    hipLaunchKernelGGL(k_pcg_start_scalars, dim3(1), dim3(256), 0, s, c, tol2, max_it);
}
void launch_pcg_start2(const CGArgs& c, hipStream_t s) {
    hipLaunchKernelGGL(k_pcg_start2, dim3(1), dim3(256), 0, s, c);
}

}  // namespace mastg
