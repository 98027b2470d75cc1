// The MAST-exact structural matrix: all 6 variables of every node in the
// caller's numbering, from the active-variable assembly plus the small
// diagonal values MAST gives the variables the element does not use.
//
// StructuralElement3D::internal_residual adds 1e-20 x (largest diagonal entry
// of the element Jacobian) to the diagonal of the three rotations
// (src/elasticity/solid_element_3d.cpp:413-417); StructuralElement2D sets the
// theta_z diagonal to 1e-8 per element (src/elasticity/structural_element_2d.
// cpp:1215-1219).  libMesh's constraint convention replaces an element's
// contribution on a constrained DOF by 1.  The inactive variables couple to
// nothing else, so every other entry they add to the full-coupling pattern is 0.
#include <hip/hip_runtime.h>

#include "device_util.h"
#include "kernels.h"

namespace mastg {

namespace {

__device__ __forceinline__ double ex_gp(int i) { return i ? 0.57735026918962576451 : -0.57735026918962576451; }

// largest diagonal entry of each hex8 element matrix: K_aa[i][i] =
// E_e sum_q JxW ((lam + mu) g_a,i^2 + mu |g_a|^2) (the isotropic B^T D B
// diagonal, D = lam 1 1^T + mu diag(2, 2, 2, 1, 1, 1)); thread per element
__global__ __launch_bounds__(128) void k_hex8_max_diag(int64_t n_elems, const double* __restrict__ xyz,
                                                       const int32_t* __restrict__ econn,
                                                       const double* __restrict__ sE, double lam, double mu,
                                                       double* __restrict__ out) {
    const int64_t e = int64_t(blockIdx.x) * 128 + threadIdx.x;
    if (e >= n_elems) return;
    double x[8][3];
#pragma unroll
    for (int n = 0; n < 8; n++) {
        const int64_t rn = econn[e * 8 + n];
#pragma unroll
        for (int d = 0; d < 3; d++) x[n][d] = xyz[rn * 3 + d];
    }
    double dg[8][3];
#pragma unroll
    for (int a = 0; a < 8; a++) dg[a][0] = dg[a][1] = dg[a][2] = 0.0;
    for (int q = 0; q < 8; q++) {
        const double xi = ex_gp(q & 1), eta = ex_gp((q >> 1) & 1), zeta = ex_gp(q >> 2);
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
        const double i00 = (J[4] * J[8] - J[5] * J[7]) * id, i01 = (J[2] * J[7] - J[1] * J[8]) * id,
                     i02 = (J[1] * J[5] - J[2] * J[4]) * id, i10 = (J[5] * J[6] - J[3] * J[8]) * id,
                     i11 = (J[0] * J[8] - J[2] * J[6]) * id, i12 = (J[2] * J[3] - J[0] * J[5]) * id,
                     i20 = (J[3] * J[7] - J[4] * J[6]) * id, i21 = (J[1] * J[6] - J[0] * J[7]) * id,
                     i22 = (J[0] * J[4] - J[1] * J[3]) * id;
#pragma unroll
        for (int a = 0; a < 8; a++) {
            const double g0 = dN[a][0] * i00 + dN[a][1] * i10 + dN[a][2] * i20;
            const double g1 = dN[a][0] * i01 + dN[a][1] * i11 + dN[a][2] * i21;
            const double g2 = dN[a][0] * i02 + dN[a][1] * i12 + dN[a][2] * i22;
            const double gg = mu * (g0 * g0 + g1 * g1 + g2 * g2);
            dg[a][0] += det * ((lam + mu) * g0 * g0 + gg);
            dg[a][1] += det * ((lam + mu) * g1 * g1 + gg);
            dg[a][2] += det * ((lam + mu) * g2 * g2 + gg);
        }
    }
    double mx = 0.0;
#pragma unroll
    for (int a = 0; a < 8; a++)
#pragma unroll
        for (int i = 0; i < 3; i++) mx = fmax(mx, sE[e] * dg[a][i]);
    out[e] = mx;
}

// diagonal of the inactive variables of each node: the sum over the node's
// incidences (incidence order) of 1 when (node, variable) is constrained,
// else coef x (per_elem ? per_elem[element] : 1); thread per node
__global__ __launch_bounds__(256) void k_inactive_diag(int64_t n_nodes, int nv, int npe, uint32_t inactive,
                                                       const int32_t* __restrict__ ninc_ptr,
                                                       const int32_t* __restrict__ ninc,
                                                       const double* __restrict__ per_elem, double coef,
                                                       const uint8_t* __restrict__ con_var,
                                                       double* __restrict__ diag) {
    const int64_t r = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (r >= n_nodes) return;
    for (int v = 0; v < nv; v++) {
        double s = 0.0;
        if ((inactive >> v) & 1) {
            const bool con = con_var[r * nv + v] != 0;
            for (int32_t i = ninc_ptr[r]; i < ninc_ptr[r + 1]; i++) {
                const int64_t e = ninc[i] / npe;
                s += con ? 1.0 : coef * (per_elem ? per_elem[e] : 1.0);
            }
        }
        diag[r * nv + v] = s;
    }
}

// row (rank r, caller variable v) of the full-variable matrix, one thread
// per row: columns base(neighbour) + w for every neighbour block and w; the
// active (v, w) entries from the active CSR chunk of node r ((v, block, w)
// order), the inactive diagonal on the node's own block, 0 elsewhere
__global__ __launch_bounds__(256) void k_export_full(FullExportArgs A) {
    const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (t >= A.n_nodes * A.nv) return;
    const int64_t r = t / A.nv;
    const int v = int(t - r * A.nv);
    const int64_t k0 = A.blk_ptr[r];
    const int deg = int(A.blk_ptr[r + 1] - k0);
    const int va = A.act[v];
    const int64_t out0 = A.row_ptr_f[A.base_r[r] + v];
    const int64_t a0 = va >= 0 ? A.row_ptr_a[r * A.nav + va] : 0;
    for (int k = 0; k < deg; k++) {
        const int32_t cb = A.blk_col[k0 + k];
        const int64_t cbase = A.base_r[cb];
        for (int w = 0; w < A.nv; w++) {
            const int64_t o = out0 + int64_t(k) * A.nv + w;
            const int wa = A.act[w];
            double x = 0.0;
            if (va >= 0 && wa >= 0) x = A.val_a[a0 + int64_t(k) * A.nav + wa];
            else if (v == w && cb == int32_t(r)) x = A.diag[r * A.nv + v];
            A.col_f[o] = int32_t(cbase + w);
            A.val_f[o] = x;
        }
    }
}

}  // namespace

void launch_hex8_max_diag(int64_t n_elems, const double* xyz, const int32_t* econn, const double* sE, double lam,
                          double mu, double* out, hipStream_t s) {
    if (n_elems <= 0) return;
    hipLaunchKernelGGL(k_hex8_max_diag, dim3(unsigned((n_elems + 127) / 128)), dim3(128), 0, s, n_elems, xyz, econn,
                       sE, lam, mu, out);
}

void launch_inactive_diag(int64_t n_nodes, int nv, int npe, uint32_t inactive, const int32_t* ninc_ptr,
                          const int32_t* ninc, const double* per_elem, double coef, const uint8_t* con_var,
                          double* diag, hipStream_t s) {
    if (n_nodes <= 0) return;
    hipLaunchKernelGGL(k_inactive_diag, dim3(unsigned((n_nodes + 255) / 256)), dim3(256), 0, s, n_nodes, nv, npe,
                       inactive, ninc_ptr, ninc, per_elem, coef, con_var, diag);
}

void launch_export_full(const FullExportArgs& a, hipStream_t s) {
    const int64_t n = a.n_nodes * a.nv;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_export_full, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, a);
}

}  // namespace mastg
