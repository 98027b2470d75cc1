// Element-wise design space of a mesh: the density filter, the material
// volume and its gradient, and the optimality-criteria update -- the steps of
// a SIMP topology-optimisation loop after the design evaluation.
//
// Filter: MAST::FilterBase (src/level_set/filter_base.cpp:35-142, 244-420)
// over the points that carry the design variables.  MAST filters a nodal
// level-set field; the configurations here carry one density per element, so
// the points are the element centroids:
//     rho~_e = sum_{f : |c_e - c_f| <= r} (r - d_ef) rho_f / sum_f (r - d_ef).
// The neighbour lists are built once on the host (a bucket grid of cell size
// r instead of MAST's all-pairs / k-d tree search; same sets) and kept on the
// device as a CSR of unnormalised weights a_ef = r - d_ef with the row sums
// S_e; the relation is symmetric, so the transpose (the chain rule of a
// sensitivity, dJ/drho_f = sum_e a_ef / S_e dJ/drho~_e) is the same gather
// over row f with the input scaled by 1/S_e.  Both are deterministic gathers
// (no atomics).
//
// Volume: V = sum_e V_e rho_e, dV/drho_e = V_e (examples/structural/example_6/
// example_6.cpp:714-791 with element densities; V_e the element volume, or
// area for a plate whose design variable is its thickness).
//
// Optimality-criteria step (the standard SIMP update; MAST's examples drive
// external optimisers through src/optimization/*, whose algorithms are not in
// the reference): rho_new = clamp(rho B_e^(1/2) lambda^(-1/2), max(rho_min,
// rho - m), min(1, rho + m)), B_e = -dC/drho_e / dV/drho_e, with lambda such
// that V(rho_new) = vf V_total.  The multiplier is found on the device by
// multi-section: each pass evaluates the volume at 64 candidates spread
// logarithmically over the current bracket in one sweep of the elements and
// keeps the sub-interval that holds the root (7 passes narrow 60 decades to a
// relative width ~1e-10), with no host round trip between passes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "device_util.h"
#include "hip_lock.h"
#include "mastg/mastg.h"

using namespace mastg;

namespace {

constexpr int kOcCand = 64;        // candidates per multi-section pass
constexpr int kOcPasses = 7;
constexpr int kOcBlocks = 512;     // work-groups of the candidate sweep

struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
    ~Buf() {
        if (p) {
            HipLock l(hip_mutex());
            (void)hipFree(p);
        }
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

}  // namespace

struct mastg_design {
    int64_t n = 0;          // elements (design variables)
    double radius = 0;
    double v_total = 0;
    std::string err;
    hipStream_t stream = nullptr;
    Buf row_ptr, col, a, inv_s, vol, part, bracket, cvol, tmp_in, tmp_out, tmp_dc, tmp_dv, scal;
    double* h_pin = nullptr;
    int64_t nnz = 0;
};

namespace {

int dfail(mastg_design* d, int code, const std::string& m) {
    if (d) d->err = m;
    return code;
}

#define DHIPC(d, expr)                                                                   \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess)                                                            \
            return dfail((d), MASTG_ERR_DEVICE, std::string(#expr " failed: ") + hipGetErrorString(_e)); \
    } while (0)

int dalloc(mastg_design* d, Buf& b, size_t bytes) {
    if (b.p && b.bytes >= bytes) return MASTG_OK;
    HipLock l(hip_mutex());
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    if (hipMalloc(&b.p, std::max<size_t>(bytes, 8)) != hipSuccess) {
        b.p = nullptr;
        return dfail(d, MASTG_ERR_OOM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    b.bytes = bytes;
    return MASTG_OK;
}

template <class T>
int upload(mastg_design* d, Buf& b, const std::vector<T>& v) {
    int rc = dalloc(d, b, v.size() * sizeof(T));
    if (rc) return rc;
    HipLock l(hip_mutex());
    DHIPC(d, hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return MASTG_OK;
}

// element volume (hex8, 2x2x2 Gauss: exact for the trilinear map's Jacobian)
// or area (quad4, 2x2 Gauss, in the plane of its nodes' first two coordinates
// after projecting on the element's own plane: |dx/dxi x dx/deta|)
double hex8_volume(const double X[8][3]) {
    const double g = 0.57735026918962576451;
    double V = 0;
    for (int q = 0; q < 8; q++) {
        const double xi = (q & 1) ? g : -g, eta = (q & 2) ? g : -g, zeta = (q & 4) ? g : -g;
        double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int n = 0; n < 8; n++) {
            const double sx = (n == 1 || n == 2 || n == 5 || n == 6) ? 1.0 : -1.0;
            const double sy = (n == 2 || n == 3 || n == 6 || n == 7) ? 1.0 : -1.0;
            const double sz = (n >= 4) ? 1.0 : -1.0;
            const double fx = 0.5 * (1 + sx * xi), fy = 0.5 * (1 + sy * eta), fz = 0.5 * (1 + sz * zeta);
            const double d[3] = {0.5 * sx * fy * fz, fx * 0.5 * sy * fz, fx * fy * 0.5 * sz};
            for (int i = 0; i < 3; i++)
                for (int r = 0; r < 3; r++) J[i * 3 + r] += X[n][i] * d[r];
        }
        V += J[0] * (J[4] * J[8] - J[7] * J[5]) + J[3] * (J[7] * J[2] - J[1] * J[8]) + J[6] * (J[1] * J[5] - J[4] * J[2]);
    }
    return V;
}

double quad4_area(const double X[4][3]) {
    const double g = 0.57735026918962576451;
    double A = 0;
    for (int q = 0; q < 4; q++) {
        const double xi = (q & 1) ? g : -g, eta = (q & 2) ? g : -g;
        double t0[3] = {0, 0, 0}, t1[3] = {0, 0, 0};
        for (int n = 0; n < 4; n++) {
            const double sx = (n == 1 || n == 2) ? 1.0 : -1.0, sy = (n >= 2) ? 1.0 : -1.0;
            const double dxi = 0.25 * sx * (1 + sy * eta), deta = 0.25 * sy * (1 + sx * xi);
            for (int i = 0; i < 3; i++) {
                t0[i] += X[n][i] * dxi;
                t1[i] += X[n][i] * deta;
            }
        }
        const double c0 = t0[1] * t1[2] - t0[2] * t1[1], c1 = t0[2] * t1[0] - t0[0] * t1[2],
                     c2 = t0[0] * t1[1] - t0[1] * t1[0];
        A += std::sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    }
    return A;
}

// ------------------------------------------------------------------ kernels --
// out_e = s_e sum_k a_k in[col_k] over row e; transpose: the input scaled by
// 1/S of its own row (inv_s), s_e = 1.  LANES lanes per row, fixed order.
template <int LANES>
__global__ __launch_bounds__(256) void k_filter(int64_t n, const int64_t* __restrict__ rp,
                                                const int32_t* __restrict__ col, const double* __restrict__ a,
                                                const double* __restrict__ inv_s, const double* __restrict__ in,
                                                double* __restrict__ out, int transpose) {
    const int64_t row = (int64_t(blockIdx.x) * 256 + threadIdx.x) / LANES;
    const int lane = threadIdx.x % LANES;
    double s = 0.0;
    if (row < n) {
        for (int64_t k = rp[row] + lane; k < rp[row + 1]; k += LANES) {
            const int32_t f = col[k];
            s += a[k] * (transpose ? in[f] * inv_s[f] : in[f]);
        }
    }
#pragma unroll
    for (int m = LANES / 2; m >= 1; m >>= 1) s += __shfl_xor(s, m, LANES);
    if (row < n && lane == 0) out[row] = transpose ? s : s * inv_s[row];
}

__global__ __launch_bounds__(256) void k_vdot(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                                              double* __restrict__ part) {
    double s = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) s += a[i] * b[i];
    block_sum_store(s, part, blockIdx.x);
}

__global__ __launch_bounds__(256) void k_vreduce(const double* __restrict__ part, int np, double* __restrict__ out) {
    const double s = reduce_partials(part, np);
    if (threadIdx.x == 0) *out = s;
}

__device__ __forceinline__ double oc_rho(double rho, double B, double lam_m12, double move, double rho_min) {
    const double lo = fmax(rho_min, rho - move), hi = fmin(1.0, rho + move);
    return fmin(hi, fmax(lo, rho * sqrt(B) * lam_m12));
}

// candidate c of the bracket [lo, hi] (log-spaced interior points)
__device__ __forceinline__ double oc_cand(const double* br, int c) {
    const double t = double(c + 1) / double(kOcCand + 1);
    return exp(log(br[0]) + t * (log(br[1]) - log(br[0])));
}

// volume at every candidate multiplier: partial sums part[c * nb + block]
__global__ __launch_bounds__(256) void k_oc_sweep(int64_t n, const double* __restrict__ rho,
                                                  const double* __restrict__ dc, const double* __restrict__ dv,
                                                  const double* __restrict__ br, double move, double rho_min,
                                                  double* __restrict__ part) {
    __shared__ double lam[kOcCand];
    __shared__ double red[4][8];
    if (threadIdx.x < kOcCand) lam[threadIdx.x] = 1.0 / sqrt(oc_cand(br, threadIdx.x));
    __syncthreads();
    for (int c0 = 0; c0 < kOcCand; c0 += 8) {
        double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
            const double r = rho[i], v = dv[i];
            const double B = fmax(0.0, -dc[i]) / v;
#pragma unroll
            for (int k = 0; k < 8; k++) s[k] += v * oc_rho(r, B, lam[c0 + k], move, rho_min);
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
#pragma unroll
            for (int m = 32; m >= 1; m >>= 1) s[k] += __shfl_xor(s[k], m, 64);
        }
        if ((threadIdx.x & 63) == 0)
#pragma unroll
            for (int k = 0; k < 8; k++) red[threadIdx.x >> 6][k] = s[k];
        __syncthreads();
        if (threadIdx.x < 8)
            part[int64_t(c0 + threadIdx.x) * gridDim.x + blockIdx.x] =
                (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
        __syncthreads();
    }
}

// candidate volumes (fixed-order sums) and the next bracket: the root of
// V(lambda) = target lies between the last candidate above the target and the
// first one at or below it (V is non-increasing in lambda)
__global__ __launch_bounds__(64) void k_oc_select(const double* __restrict__ part, int nb, double target,
                                                  double* __restrict__ br, double* __restrict__ cvol) {
    __shared__ double vol[kOcCand];
    const int c = threadIdx.x;
    double s = 0.0;
    for (int b = 0; b < nb; b++) s += part[int64_t(c) * nb + b];
    vol[c] = s;
    __syncthreads();
    if (c == 0) {
        double lo = br[0], hi = br[1];
        int first = kOcCand;
        for (int k = 0; k < kOcCand; k++)
            if (vol[k] <= target) { first = k; break; }
        const double nlo = first == 0 ? lo : oc_cand(br, first - 1);
        const double nhi = first == kOcCand ? hi : oc_cand(br, first);
        br[0] = nlo;
        br[1] = nhi;
        cvol[0] = first == kOcCand ? vol[kOcCand - 1] : vol[first];
    }
}

__global__ __launch_bounds__(256) void k_oc_final(int64_t n, const double* __restrict__ rho,
                                                  const double* __restrict__ dc, const double* __restrict__ dv,
                                                  const double* __restrict__ br, double move, double rho_min,
                                                  double* __restrict__ out) {
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    const double B = fmax(0.0, -dc[i]) / dv[i];
    out[i] = oc_rho(rho[i], B, 1.0 / sqrt(br[1]), move, rho_min);
}

int64_t blocks_of(int64_t n, int per) { return std::max<int64_t>(1, (n + per - 1) / per); }

// host or device input -> device pointer (staged in `tmp` for host data)
int stage(mastg_design* d, const double* src, int where, Buf& tmp, const double** out) {
    if (where == MASTG_DEVICE) {
        *out = src;
        return MASTG_OK;
    }
    int rc = dalloc(d, tmp, size_t(d->n) * sizeof(double));
    if (rc) return rc;
    DHIPC(d, hipMemcpyAsync(tmp.p, src, size_t(d->n) * sizeof(double), hipMemcpyHostToDevice, d->stream));
    *out = tmp.as<double>();
    return MASTG_OK;
}

}  // namespace

extern "C" {

int mastg_design_create(const mastg_problem_desc* desc, double radius, mastg_design** out) {
    if (!desc || !out || !desc->xyz || !desc->conn || desc->n_elems <= 0) return MASTG_ERR_INVALID;
    if (!(radius > 0)) return MASTG_ERR_INVALID;
    const int npe = desc->physics == MASTG_QUAD4_MINDLIN ? 4 : 8;
    auto* d = new mastg_design;
    d->n = desc->n_elems;
    d->radius = radius;
    const int64_t n = d->n;
    // centroids and volumes
    std::vector<double> c(size_t(n) * 3), vol(n);
    double lo[3] = {1e300, 1e300, 1e300};
    for (int64_t e = 0; e < n; e++) {
        double X[8][3];
        for (int a = 0; a < npe; a++) {
            const int32_t nd = desc->conn[e * npe + a];
            if (nd < 0 || nd >= desc->n_nodes) {
                delete d;
                return MASTG_ERR_INVALID;
            }
            for (int i = 0; i < 3; i++) X[a][i] = desc->xyz[int64_t(nd) * 3 + i];
        }
        for (int i = 0; i < 3; i++) {
            double s = 0;
            for (int a = 0; a < npe; a++) s += X[a][i];
            c[e * 3 + i] = s / npe;
            lo[i] = std::min(lo[i], c[e * 3 + i]);
        }
        vol[e] = npe == 8 ? hex8_volume(X) : quad4_area(reinterpret_cast<const double(*)[3]>(X));
        d->v_total += vol[e];
    }
    // neighbours within the radius: bucket grid of cell size r
    auto key = [&](int64_t e, int k) { return int64_t(std::floor((c[e * 3 + k] - lo[k]) / radius)); };
    int64_t nb[3] = {1, 1, 1};
    for (int64_t e = 0; e < n; e++)
        for (int k = 0; k < 3; k++) nb[k] = std::max(nb[k], key(e, k) + 1);
    std::unordered_map<int64_t, std::vector<int32_t>> bucket;
    bucket.reserve(size_t(n));
    auto bid = [&](int64_t i, int64_t j, int64_t k) { return (k * nb[1] + j) * nb[0] + i; };
    for (int64_t e = 0; e < n; e++) bucket[bid(key(e, 0), key(e, 1), key(e, 2))].push_back(int32_t(e));
    std::vector<int64_t> rp(n + 1, 0);
    std::vector<int32_t> col;
    std::vector<double> a, inv_s(n);
    std::vector<std::pair<int32_t, double>> row;
    for (int64_t e = 0; e < n; e++) {
        row.clear();
        const int64_t ki = key(e, 0), kj = key(e, 1), kk = key(e, 2);
        for (int64_t z = kk - 1; z <= kk + 1; z++)
            for (int64_t y = kj - 1; y <= kj + 1; y++)
                for (int64_t x = ki - 1; x <= ki + 1; x++) {
                    if (x < 0 || y < 0 || z < 0 || x >= nb[0] || y >= nb[1] || z >= nb[2]) continue;
                    auto it = bucket.find(bid(x, y, z));
                    if (it == bucket.end()) continue;
                    for (int32_t f : it->second) {
                        const double dx = c[e * 3] - c[f * 3], dy = c[e * 3 + 1] - c[f * 3 + 1],
                                     dz = c[e * 3 + 2] - c[f * 3 + 2];
                        const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
                        if (dist <= radius) row.emplace_back(f, radius - dist);
                    }
                }
        std::sort(row.begin(), row.end());
        double s = 0;
        for (auto& p : row) {
            col.push_back(p.first);
            a.push_back(p.second);
            s += p.second;
        }
        inv_s[e] = 1.0 / s;
        rp[e + 1] = int64_t(col.size());
    }
    d->nnz = rp[n];
    {
        HipLock l(hip_mutex());
        if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess ||
            hipHostMalloc((void**)&d->h_pin, 8 * sizeof(double)) != hipSuccess) {
            mastg_design_destroy(d);
            return MASTG_ERR_DEVICE;
        }
    }
    int rc;
    if ((rc = upload(d, d->row_ptr, rp)) || (rc = upload(d, d->col, col)) || (rc = upload(d, d->a, a)) ||
        (rc = upload(d, d->inv_s, inv_s)) || (rc = upload(d, d->vol, vol)) ||
        (rc = dalloc(d, d->part, size_t(kOcCand) * kOcBlocks * sizeof(double))) ||
        (rc = dalloc(d, d->bracket, 2 * sizeof(double))) || (rc = dalloc(d, d->cvol, sizeof(double))) ||
        (rc = dalloc(d, d->scal, sizeof(double))) ||
        (rc = dalloc(d, d->tmp_out, size_t(n) * sizeof(double)))) {
        mastg_design_destroy(d);
        return rc;
    }
    *out = d;
    return MASTG_OK;
}

int mastg_design_destroy(mastg_design* d) {
    if (!d) return MASTG_OK;
    {
        HipLock l(hip_mutex());
        if (d->stream) (void)hipStreamSynchronize(d->stream);
        if (d->h_pin) (void)hipHostFree(d->h_pin);
    }
    hipStream_t s = d->stream;
    delete d;
    if (s) {
        HipLock l(hip_mutex());
        (void)hipStreamDestroy(s);
    }
    return MASTG_OK;
}

const char* mastg_design_last_error(const mastg_design* d) { return d ? d->err.c_str() : ""; }

int mastg_design_info(const mastg_design* d, int64_t* n_elems, int64_t* nnz, double* v_total) {
    if (!d) return MASTG_ERR_INVALID;
    if (n_elems) *n_elems = d->n;
    if (nnz) *nnz = d->nnz;
    if (v_total) *v_total = d->v_total;
    return MASTG_OK;
}

int mastg_design_filter(mastg_design* d, const double* in, double* out, int transpose, int where) {
    if (!d || !in || !out || (where != MASTG_HOST && where != MASTG_DEVICE)) return MASTG_ERR_INVALID;
    const double* din;
    int rc = stage(d, in, where, d->tmp_in, &din);
    if (rc) return rc;
    double* dout = where == MASTG_DEVICE ? out : d->tmp_out.as<double>();
    hipLaunchKernelGGL(k_filter<8>, dim3(blocks_of(d->n * 8, 256)), dim3(256), 0, d->stream, d->n,
                       d->row_ptr.as<int64_t>(), d->col.as<int32_t>(), d->a.as<double>(), d->inv_s.as<double>(),
                       din, dout, transpose ? 1 : 0);
    DHIPC(d, hipGetLastError());
    if (where == MASTG_HOST)
        DHIPC(d, hipMemcpyAsync(out, dout, size_t(d->n) * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    DHIPC(d, hipStreamSynchronize(d->stream));
    return MASTG_OK;
}

int mastg_design_volume(mastg_design* d, const double* rho, int where, double* volume, double* dvol) {
    if (!d || !rho || (where != MASTG_HOST && where != MASTG_DEVICE)) return MASTG_ERR_INVALID;
    const double* dr;
    int rc = stage(d, rho, where, d->tmp_in, &dr);
    if (rc) return rc;
    if (volume) {
        const int nb = int(std::min<int64_t>(kOcBlocks, blocks_of(d->n, 256)));
        hipLaunchKernelGGL(k_vdot, dim3(nb), dim3(256), 0, d->stream, d->n, d->vol.as<double>(), dr,
                           d->part.as<double>());
        hipLaunchKernelGGL(k_vreduce, dim3(1), dim3(256), 0, d->stream, d->part.as<double>(), nb,
                           d->scal.as<double>());
        DHIPC(d, hipGetLastError());
        DHIPC(d, hipMemcpyAsync(d->h_pin, d->scal.p, sizeof(double), hipMemcpyDeviceToHost, d->stream));
    }
    if (dvol)
        DHIPC(d, hipMemcpyAsync(dvol, d->vol.p, size_t(d->n) * sizeof(double),
                                where == MASTG_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, d->stream));
    DHIPC(d, hipStreamSynchronize(d->stream));
    if (volume) *volume = d->h_pin[0];
    return MASTG_OK;
}

int mastg_design_oc_update(mastg_design* d, const double* rho, const double* dc, const double* dv,
                           double volfrac, double move, double rho_min, double* rho_new, int where,
                           double* lambda) {
    if (!d || !rho || !dc || !dv || !rho_new || (where != MASTG_HOST && where != MASTG_DEVICE))
        return MASTG_ERR_INVALID;
    if (!(volfrac > 0 && volfrac <= 1) || !(move > 0) || !(rho_min >= 0 && rho_min < 1))
        return dfail(d, MASTG_ERR_INVALID, "volume fraction in (0, 1], move > 0, 0 <= rho_min < 1");
    const double* dr;
    const double* ddc;
    const double* ddv;
    int rc;
    if ((rc = stage(d, rho, where, d->tmp_in, &dr)) || (rc = stage(d, dc, where, d->tmp_dc, &ddc)) ||
        (rc = stage(d, dv, where, d->tmp_dv, &ddv)))
        return rc;
    const double target = volfrac * d->v_total;
    d->h_pin[0] = 1e-30;
    d->h_pin[1] = 1e30;
    DHIPC(d, hipMemcpyAsync(d->bracket.p, d->h_pin, 2 * sizeof(double), hipMemcpyHostToDevice, d->stream));
    const int nb = int(std::min<int64_t>(kOcBlocks, blocks_of(d->n, 256)));
    for (int p = 0; p < kOcPasses; p++) {
        hipLaunchKernelGGL(k_oc_sweep, dim3(nb), dim3(256), 0, d->stream, d->n, dr, ddc, ddv,
                           d->bracket.as<double>(), move, rho_min, d->part.as<double>());
        hipLaunchKernelGGL(k_oc_select, dim3(1), dim3(kOcCand), 0, d->stream, d->part.as<double>(), nb, target,
                           d->bracket.as<double>(), d->cvol.as<double>());
    }
    double* dout = where == MASTG_DEVICE ? rho_new : d->tmp_out.as<double>();
    hipLaunchKernelGGL(k_oc_final, dim3(blocks_of(d->n, 256)), dim3(256), 0, d->stream, d->n, dr, ddc, ddv,
                       d->bracket.as<double>(), move, rho_min, dout);
    DHIPC(d, hipGetLastError());
    if (where == MASTG_HOST)
        DHIPC(d, hipMemcpyAsync(rho_new, dout, size_t(d->n) * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    DHIPC(d, hipMemcpyAsync(d->h_pin, d->bracket.p, 2 * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    DHIPC(d, hipStreamSynchronize(d->stream));
    if (lambda) *lambda = d->h_pin[1];
    return MASTG_OK;
}

}  // extern "C"
