// Block-level reductions shared by the kernel files (blockDim 256, 4 waves).
#pragma once
#include <hip/hip_runtime.h>

namespace mastg {

// XCD-aware work-group order.  The dispatcher deals work-groups to the 8 XCDs
// (each with its own L2) round-robin: block b runs on XCD b % 8.  Renumbering
// block b to this index gives XCD x the contiguous run [x q + min(x, r),
// (x + 1) q + min(x + 1, r)) of an nb-block grid (nb = 8 q + r), so stencil
// neighbours gathered by one work-group were mostly fetched into the same L2
// by the work-groups before it.  A bijection of [0, nb) for every nb >= 1.
__host__ __device__ constexpr int64_t xcd_block(int64_t b, int64_t nb) {
    return (b & 7) * (nb >> 3) + ((b & 7) < (nb & 7) ? (b & 7) : (nb & 7)) + (b >> 3);
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

// deterministic final reduction of `n` partials (one block of 256)
__device__ __forceinline__ double reduce_partials(const double* __restrict__ p, int n) {
    __shared__ double red[4];
    // four independent accumulators keep four loads in flight per thread
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int i = threadIdx.x;
    for (; i + 768 < n; i += 1024) {
        s0 += p[i];
        s1 += p[i + 256];
        s2 += p[i + 512];
        s3 += p[i + 768];
    }
    for (; i < n; i += 256) s0 += p[i];
    double s = (s0 + s1) + (s2 + s3);
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    const double t = (red[0] + red[1]) + (red[2] + red[3]);
    __syncthreads();
    return t;
}

// block-level max of non-negative values into partials[slot]
__device__ __forceinline__ void block_max_store(double v, double* partials, int slot) {
    __shared__ double red[4];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) partials[slot] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

}  // namespace mastg
