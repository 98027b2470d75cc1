// C ABI of the mastg library (include/mastg/mastg.h): context set-up, the
// device-resident assembly / solve / sensitivity pipeline and its timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "kernels.h"
#include "plate_frame.h"
#include "mastg/mastg.h"
#include "hip_lock.h"
#include "multigrid.h"
#include "structure.h"

using namespace mastg;

namespace mastg {
std::recursive_mutex& hip_mutex();
}

namespace {

thread_local std::string g_create_error;

// Stream capture and the implicitly synchronising HIP calls (allocation, free,
// synchronous copies, stream / event creation) are serialised process-wide:
// contexts driven from several threads (the request server) must never
// allocate while another thread captures a solver graph.
std::recursive_mutex g_hip_mu;
inline hipError_t ev_create(hipEvent_t* e) { HipLock l(g_hip_mu); return hipEventCreate(e); }
inline hipError_t ev_destroy(hipEvent_t e) { HipLock l(g_hip_mu); return hipEventDestroy(e); }

// Device scratch of the allocation-free output entry point
// (mastg_output_noalloc): the partial sums of its reduction.  One copy per
// GPU, held until the process ends; every call on a GPU shares it, so calls
// that overlap on one GPU race.  Host code never passes `scratch` itself to a
// kernel (in host code the name is a stand-in, not a device address): it
// uses the address hipGetSymbolAddress returns for the current device.
constexpr int kScratchPartials = 4096;
static __device__ __align__(256) double scratch[kScratchPartials];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// the scratch array's address on the current device (looked up once per device)
hipError_t scratch_address(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> l(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace

std::recursive_mutex& mastg::hip_mutex() { return g_hip_mu; }

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { release(); }
    void release() {
        if (p) {
            HipLock l(g_hip_mu);
            (void)hipFree(p);
        }
        p = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

struct KTimer {   // accumulated device time of one kernel kind
    double ms = 0;
    int64_t launches = 0;
};

// one multigrid level on the device (level 0 borrows the context's operator)
struct MGLevelDev {
    MGLevelArgs a{};
    int64_t n_elems = 0;
    DevBuf rank_of_lat, elem_of_cell, con_lat, nodecon_lat, dinv_lat, b_lat, sE_cell;   // level 0
    DevBuf x32, res32, da32, db32, dinv32, y_lat;   // hex8 level 0: FP32 smoothing state, FP64 scratch
    DevBuf blk_ptr, blk_col, cptr, cidx, econn, con;   // levels >= 1
    DevBuf gal_dirty, gal_list;   // hex8 level 1: coarse cells touching constraints
    int64_t n_dirty = 0;
    DevBuf sten;                  // hex8 level 1: the operator as a symmetric 27-point FP32 stencil;
                                  // heat levels 0 and 1: 27-point stencils in lattice order
    DevBuf kpos;                  // heat level 0: row position of each stencil entry (255: none)
    DevBuf dinvb;                 // plate level 0: inverses of the 5 x 5 diagonal node blocks
    DevBuf sten16, srow16;        // heat level 0: the stencil in FP16 and its row scales; plate
                                  // level 1: its node blocks in FP16 and their row scales
    bool lat0 = false;            // heat / hex8 level 0 smooths in lattice order
    DevBuf Kel, val, val32, dinv, r, x, res, da, db;
};

}  // namespace

struct mastg_ctx {
    mastg_problem_desc d{};          // scalars of the description (pointers cleared)
    Structure S;
    std::string err;
    hipStream_t stream = nullptr;      // where the context's work is enqueued
    hipStream_t own_stream = nullptr;  // the stream the context created
    int nparts = 1024;
    int64_t device_bytes = 0;
    // device structure
    DevBuf xyz, econn, blk_ptr, blk_col, row_ptr, col, cptr, cidx, diag_pos, con, dof_of_active;
    DevBuf ninc_ptr, ninc;
    // matrix-free operator data (hex8 congruence classes)
    DevBuf Khat, sE, eclass, nodecon, Fel;
    DevBuf heat_geo;            // heat: per congruence class gradients and JxW
    DevBuf Khat32;              // FP32 copy of Khat (lattice smoothing)
    DevBuf Gc;                  // [8][576] Q_c^T Khat Q_c (lattice Galerkin products)
    // heat: convective side masks, mean-T weights, BiCGStab work space, A^T values
    DevBuf sidemask, radmask, wmean, tpos, valT, vr0, vv, vs, vt, vph, vsh, bscal, vT;
    bool has_rad = false;
    bool row_fused = true;           // fused row-owner assembly allowed (MASTG_OPT_ROW_ASSEMBLY)
    bool row_plan_ok = true;         // ... over the row plan where it applies (MASTG_OPT_ROW_ASSEMBLY 1)
    DevBuf row_pos;                  // the class assembly's row plan + node constraint bits (mesh-static)
    bool row_plan_tried = false;     // whether the mesh qualifies for the plan has been decided
    double* scratch_ptr = nullptr;   // this device's `scratch` (allocation-free output)
    double ks_k = 10.0, ks_sref = 1.0;   // KS stress output: aggregation constant, reference stress
    // p-norm and smooth-ramp stress outputs: p, weight exponent, rho, reference stress
    // (StressStrainOutputBase's defaults 2, 1, 1; the reference stress must be set)
    double agg_p = 2.0, agg_pw = 1.0, agg_rho = 1.0, agg_s0 = 0.0;
    DevBuf ks_part, ks_IV;
    double volume = 0;
    hipGraphExec_t bicg_exec = nullptr;
    int bicg_exec_len = 0;
    bool bicg_exec_T = false;
    // device data
    DevBuf param, Kel, Kel32, Rel, val, dinv, F;
    DevBuf Fp;                        // plate, hex8 pressure sides: load vector of a unit pressure
    bool has_pressure = false;        // hex8 with pressure sides
    DevBuf vx, vr, vz, vp, vq, vb, vtmp, vcaller, vg, part0, part1, scal, dscalar;
    CGScalars* h_scal = nullptr;      // pinned mirror of the CG scalars
    double* h_pin = nullptr;          // pinned scratch for scalar read-back
    double Dhat[36];
    double Dps[9], Gmod = 0;          // plate: plane-stress D and shear modulus
    bool assembled = false;
    bool params_set = false;
    // captured CG iterations, one graph per operator (CSR, matrix-free)
    hipGraphExec_t cg_exec[3] = {nullptr, nullptr, nullptr};   // CSR, matrix-free, FP32 node-block
    int cg_exec_len[3] = {0, 0, 0};
    DevBuf val32;              // FP32 copy of val for the mixed-precision solver
    DevBuf val16, rs16;        // plate problem level: row-scaled FP16 copy of val and its row scales
    bool val32_valid = false;
    int cg_replacements = 0;
    bool kel64_valid = false;  // FP64 element blocks of the last assembly are resident
    // geometric multigrid (hex8 lattice meshes), built on first use
    std::vector<std::unique_ptr<MGLevelDev>> mg;
    bool mg_tried = false;
    std::string mg_why;
    bool mg_valid = false;     // coarse operators match the last assembly
    int mg_k0 = 3, mg_kc = 3;  // Chebyshev steps per smoothing: problem level, coarse levels
    bool mg_k0_set = false;    // problem-level steps chosen by the caller (MASTG_OPT_MG_STEPS_FINE)
    int mg_ratio = 0;          // Chebyshev interval [lam / ratio, lam] (0: the physics' default)
    int64_t mg_n_coarse = 0;   // DOFs of the dense coarsest level
    DevBuf mg_cheb, mg_lam, mg_part, mg_dense, mg_scal;
    bool mg_power = false;     // Chebyshev bounds from power iteration (plates)
    hipGraphExec_t mg_exec = nullptr;
    int mg_exec_len = 0;
    hipGraphExec_t mg_bicg_exec = nullptr;
    int mg_bicg_exec_len = 0;
    bool mg_bicg_exec_T = false;
    // launch shapes (autotuned on the first solve)
    int spmv_sw = 16;
    bool use_graphs = true;
    bool tuned = false;
    bool sE_valid = false;
    bool sEc_valid = false;     // the lattice operator's cell-ordered copy of sE
    // timing
    bool timing = false;
    KTimer kt[8];
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace {

int fail(mastg_ctx* c, int code, const std::string& m) {
    if (c) c->err = m;
    else g_create_error = m;
    return code;
}

#define HIPC(c, expr)                                                                     \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess)                                                             \
            return fail((c), MASTG_ERR_DEVICE,                                            \
                        std::string(#expr " failed: ") + hipGetErrorString(_e));          \
    } while (0)

int dalloc(mastg_ctx* c, DevBuf& b, size_t bytes) {
    if (b.p && b.bytes >= bytes) return MASTG_OK;
    c->device_bytes -= int64_t(b.bytes);
    b.release();
    if (bytes == 0) bytes = 8;
    hipError_t e;
    {
        HipLock l(g_hip_mu);
        e = hipMalloc(&b.p, bytes);
    }
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(c, MASTG_ERR_OOM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    b.bytes = bytes;
    c->device_bytes += int64_t(bytes);
    return MASTG_OK;
}

template <class T>
int upload(mastg_ctx* c, DevBuf& b, const std::vector<T>& v) {
    int rc = dalloc(c, b, v.size() * sizeof(T));
    if (rc) return rc;
    if (!v.empty()) {
        HipLock l(g_hip_mu);
        HIPC(c, hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    }
    return MASTG_OK;
}

#define TRY(x)                  \
    do {                        \
        int _rc = (x);          \
        if (_rc) return _rc;    \
    } while (0)

void material_D3_unitE(double nu, double D[36]) {
    // IsotropicMaterialProperty::StiffnessMatrix3D at E = 1
    // (src/property_cards/isotropic_material_property_card.cpp:512-528)
    std::memset(D, 0, 36 * sizeof(double));
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++)
            D[i * 6 + j] = (i == j) ? (1. - nu) / (1. - nu - 2. * nu * nu) : nu / (1. - nu - 2. * nu * nu);
        D[(i + 3) * 6 + i + 3] = 1. / 2. / (1. + nu);
    }
}

int global_dRdp(mastg_ctx* c, int param, const double* Xa, double* out);

Hex8Args hex8_args(mastg_ctx* c, const double* X, bool want_K) {
    Hex8Args a{};
    a.n_elems = c->S.n_elems;
    a.xyz = c->xyz.as<double>();
    a.econn = c->econn.as<int32_t>();
    a.param = c->param.as<double>();
    a.E0 = c->d.E;
    a.penalty = c->d.penalty;
    std::memcpy(a.Dhat, c->Dhat, sizeof(a.Dhat));
    a.con = c->con.as<uint8_t>();
    a.X = X;
    a.Rel = c->Rel.as<double>();
    a.Kel = want_K ? c->Kel.as<double>() : nullptr;
    return a;
}

Quad4Args quad4_args(mastg_ctx* c, const double* X, bool want_K) {
    Quad4Args a{};
    a.n_elems = c->S.n_elems;
    a.nav = c->S.nav;
    a.xyz = c->xyz.as<double>();
    a.econn = c->econn.as<int32_t>();
    a.param = c->param.as<double>();
    std::memcpy(a.Dps, c->Dps, sizeof(a.Dps));
    a.G = c->Gmod;
    a.kappa = c->d.kappa;
    a.off = c->d.offset;
    a.con = c->con.as<uint8_t>();
    a.X = X;
    a.Rel = c->Rel.as<double>();
    a.Kel = want_K ? c->Kel.as<double>() : nullptr;
    return a;
}

HeatArgs heat_args(mastg_ctx* c, const double* X, bool want_R, bool want_K) {
    HeatArgs a{};
    a.n_elems = c->S.n_elems;
    a.xyz = c->xyz.as<double>();
    a.econn = c->econn.as<int32_t>();
    a.param = c->param.as<double>();
    a.k0 = c->d.k0;
    a.alpha = c->d.alpha;
    a.h = c->d.conv_h;
    a.Tinf = c->d.conv_tinf;
    a.sides = c->sidemask.as<uint8_t>();
    a.rsides = c->has_rad ? c->radmask.as<uint8_t>() : nullptr;
    a.rad_se = c->d.rad_sigma * c->d.rad_emissivity;
    a.rad_Tamb = c->d.rad_tamb;
    a.rad_T0 = c->d.rad_tzero;
    a.src_scale = 1.0;
    a.geo = c->heat_geo.as<double>();
    a.eclass = c->heat_geo.p ? c->eclass.as<uint8_t>() : nullptr;
    a.con = c->con.as<uint8_t>();
    a.X = X;
    a.Rel = want_R ? c->Rel.as<double>() : nullptr;
    a.Kel = want_K ? c->Kel.as<double>() : nullptr;
    return a;
}

// element kernel of the context's physics
void launch_elements(mastg_ctx* c, const double* X, bool want_K) {
    if (c->d.physics == MASTG_HEX8_ELASTIC) launch_hex8_K(hex8_args(c, X, want_K), c->stream);
    else if (c->d.physics == MASTG_QUAD4_MINDLIN) launch_quad4_K(quad4_args(c, X, want_K), c->stream);
    else launch_heat_K(heat_args(c, X ? X : c->vT.as<double>(), X != nullptr, want_K), c->stream);
}

// g_e = lambda . dR_e/dp_e for the context's physics
void launch_sens(mastg_ctx* c, const double* X, const double* lam, double* g) {
    if (c->d.physics == MASTG_HEX8_ELASTIC) launch_hex8_sens(hex8_args(c, X, false), lam, g, c->stream);
    else if (c->d.physics == MASTG_QUAD4_MINDLIN) launch_quad4_sens(quad4_args(c, X, false), lam, g, c->stream);
    else launch_heat_source_dot(heat_args(c, X, false, false), lam, -1.0, g, c->stream);
}

CGArgs cg_args(mastg_ctx* c) {
    CGArgs a{};
    a.n = c->S.n_active;
    a.row_ptr = c->row_ptr.as<int64_t>();
    a.col = c->col.as<int32_t>();
    a.val = c->val.as<double>();
    a.dinv = c->dinv.as<double>();
    a.x = c->vx.as<double>();
    a.r = c->vr.as<double>();
    a.z = c->vz.as<double>();
    a.p = c->vp.as<double>();
    a.q = c->vq.as<double>();
    a.part0 = c->part0.as<double>();
    a.part1 = c->part1.as<double>();
    a.nparts = c->nparts;
    a.S = c->scal.as<CGScalars>();
    a.use_mf = 0;
    a.spmv_sw = c->spmv_sw;
    a.mf.n_nodes = c->S.n_nodes;
    a.mf.n_classes = c->S.n_classes;
    a.mf.Khat = c->Khat.as<double>();
    a.mf.sE = c->sE.as<double>();
    a.mf.eclass = c->eclass.as<uint8_t>();
    a.mf.econn = c->econn.as<int32_t>();
    a.mf.ninc_ptr = c->ninc_ptr.as<int32_t>();
    a.mf.ninc = c->ninc.as<int32_t>();
    a.mf.nodecon = c->nodecon.as<uint8_t>();
    a.mf.n_elems = c->S.n_elems;
    a.mf.Fel = c->Fel.as<double>();
    a.nparts_op = c->nparts;
    a.nvec = std::min<int64_t>(2048, std::max<int64_t>(1, (c->S.n_active + 255) / 256));
    a.blk_ptr = c->blk_ptr.as<int64_t>();
    a.blk_col = c->blk_col.as<int32_t>();
    a.val32 = c->val32.as<float>();
    a.nav = c->S.nav;
    return a;
}

// timing helpers: bracket a launch with events when timing is on
struct TimedRegion {
    mastg_ctx* c;
    int kind;
    TimedRegion(mastg_ctx* c_, int k) : c(c_), kind(k) {
        if (c->timing) (void)hipEventRecord(c->ev0, c->stream);
    }
    ~TimedRegion() {
        if (!c->timing) return;
        (void)hipEventRecord(c->ev1, c->stream);
        (void)hipEventSynchronize(c->ev1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
        c->kt[kind].ms += ms;
        c->kt[kind].launches += 1;
    }
};

// caller vector -> active vector (device), or zero
int to_active(mastg_ctx* c, const double* src, int where, double* dst_active) {
    const int64_t n = c->S.n_active;
    if (!src) {
        launch_fill(n, 0.0, dst_active, c->stream);
        return MASTG_OK;
    }
    const double* s = src;
    if (where == MASTG_HOST) {
        TRY(dalloc(c, c->vcaller, c->S.n_dofs * sizeof(double)));
        HIPC(c, hipMemcpyAsync(c->vcaller.p, src, c->S.n_dofs * sizeof(double), hipMemcpyHostToDevice,
                               c->stream));
        s = c->vcaller.as<double>();
    }
    launch_gather(n, c->dof_of_active.as<int64_t>(), s, dst_active, c->stream);
    return MASTG_OK;
}

// active vector (device) -> caller vector (inactive entries 0)
int from_active(mastg_ctx* c, const double* src_active, double* dst, int where) {
    TRY(dalloc(c, c->vcaller, c->S.n_dofs * sizeof(double)));
    double* t = where == MASTG_HOST ? c->vcaller.as<double>() : dst;
    launch_fill(c->S.n_dofs, 0.0, t, c->stream);
    launch_scatter_vec(c->S.n_active, c->dof_of_active.as<int64_t>(), src_active, t, c->stream);
    if (where == MASTG_HOST) {
        HIPC(c, hipMemcpyAsync(dst, t, c->S.n_dofs * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPC(c, hipStreamSynchronize(c->stream));
    }
    return MASTG_OK;
}

int ensure_vectors(mastg_ctx* c) {
    const size_t nb = size_t(c->S.n_active) * sizeof(double);
    for (DevBuf* b : {&c->vx, &c->vr, &c->vz, &c->vp, &c->vq, &c->vb, &c->vtmp}) TRY(dalloc(c, *b, nb));
    // partial sums: enough for the largest grid any kernel writes them from
    // (autotuned SpMV grid up to 8192, matrix-free node phase n_nodes/256;
    // BiCGStab keeps the SpMV partials and a vector grid of up to 2048 after them)
    const size_t np = size_t(std::max<int64_t>(8192 + 2048, mf_node_blocks(c->S.n_nodes)));
    TRY(dalloc(c, c->part0, np * sizeof(double)));
    TRY(dalloc(c, c->part1, np * sizeof(double)));
    TRY(dalloc(c, c->scal, sizeof(CGScalars)));
    TRY(dalloc(c, c->dscalar, 4 * sizeof(double)));
    TRY(dalloc(c, c->vg, size_t(c->S.n_elems) * sizeof(double)));
    return MASTG_OK;
}

// device dot of two active vectors -> host double
int dot_active(mastg_ctx* c, const double* a, const double* b, double* out) {
    launch_dot(c->S.n_active, a, b, c->part0.as<double>(), c->nparts, c->dscalar.as<double>(), c->stream);
    HIPC(c, hipMemcpyAsync(c->h_pin, c->dscalar.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    *out = c->h_pin[0];
    return MASTG_OK;
}

void ensure_scales(mastg_ctx* c);

// the fused row-owner assembly applies: hex8 elasticity in FP64, rows of at
// most the accumulator's block count, and no consumer of the element blocks
// (the multigrid on an assembled problem level forms its Galerkin products
// from them)
bool row_assembly_ok(mastg_ctx* c, bool fp32) {
    if (c->d.physics != MASTG_HEX8_ELASTIC || fp32 || !c->row_fused) return false;
    if (!c->mg.empty() && c->mg_power) return false;
    return c->S.max_blocks_per_row <= hex8_row_assembly_max_blocks();
}

// ---- assembly: element kernel -> segmented scatter -> Jacobi diagonal, or
// the fused row-owner kernel (hex8 elasticity)
int assemble_active(mastg_ctx* c, const double* X_active, bool want_R, bool want_K, bool fp32 = false) {
    if (!c->params_set) return fail(c, MASTG_ERR_INVALID, "design parameters not set");
    const Structure& S = c->S;
    if (want_K && row_assembly_ok(c, fp32)) {
        TRY(dalloc(c, c->val, size_t(S.nnz) * sizeof(double)));
        TRY(dalloc(c, c->dinv, size_t(S.n_active) * sizeof(double)));
        if (!c->sE.p) TRY(dalloc(c, c->sE, size_t(S.n_elems) * sizeof(double)));   // meshes without classes
        ensure_scales(c);
        Hex8RowArgs a{};
        a.n_nodes = S.n_nodes;
        a.xyz = c->xyz.as<double>();
        a.econn = c->econn.as<int32_t>();
        a.ninc_ptr = c->ninc_ptr.as<int32_t>();
        a.ninc = c->ninc.as<int32_t>();
        a.blk_ptr = c->blk_ptr.as<int64_t>();
        a.blk_col = c->blk_col.as<int32_t>();
        a.row_ptr = c->row_ptr.as<int64_t>();
        a.sE = c->sE.as<double>();
        a.lam = c->Dhat[1];        // D(E = 1) = lam 1 1^T + mu diag(2, 2, 2, 1, 1, 1)
        a.mu = c->Dhat[21];
        // congruence classes: the blocks are scaled class matrices
        a.Khat = S.n_classes > 0 ? c->Khat.as<double>() : nullptr;
        a.eclass = S.n_classes > 0 ? c->eclass.as<uint8_t>() : nullptr;
        a.n_classes = S.n_classes;
        a.con = c->con.as<uint8_t>();
        a.X = X_active;
        a.F = c->F.as<double>();
        a.val = c->val.as<double>();
        a.R = want_R ? c->vtmp.as<double>() : nullptr;
        a.dinv = c->dinv.as<double>();
        // the class kernel's row plan: at most 8 elements per node
        if (a.Khat && c->row_plan_ok && !c->row_pos.p && !c->row_plan_tried) {
            c->row_plan_tried = true;   // the mesh does not change: decided once
            int32_t max_inc = 0;
            for (int64_t r = 0; r < S.n_nodes; r++)
                max_inc = std::max(max_inc, S.ninc_ptr[r + 1] - S.ninc_ptr[r]);
            if (max_inc <= 8) {
                TRY(dalloc(c, c->row_pos, size_t(S.n_nodes) * 65));
                launch_hex8_row_plan(a, c->row_pos.as<int8_t>(), c->row_pos.as<uint8_t>() + S.n_nodes * 64,
                                     c->stream);
            }
        }
        a.row_pos = c->row_plan_ok ? c->row_pos.as<int8_t>() : nullptr;
        a.cmask = a.row_pos ? c->row_pos.as<uint8_t>() + S.n_nodes * 64 : nullptr;
        {
            TimedRegion t(c, 0);
            launch_hex8_row_assemble(a, int(S.max_blocks_per_row), c->stream);
        }
        c->assembled = true;
        c->val32_valid = false;
        c->kel64_valid = false;
        c->mg_valid = false;
        HIPC(c, hipGetLastError());
        return MASTG_OK;
    }
    if (want_K) {
        const size_t nkel = size_t(S.n_elems) * S.npe * S.npe * S.nav * S.nav;
        if (fp32) TRY(dalloc(c, c->Kel32, nkel * sizeof(float)));
        else TRY(dalloc(c, c->Kel, nkel * sizeof(double)));
        TRY(dalloc(c, c->val, size_t(S.nnz) * sizeof(double)));
        TRY(dalloc(c, c->dinv, size_t(S.n_active) * sizeof(double)));
    }
    if (want_R) TRY(dalloc(c, c->Rel, size_t(S.n_elems) * S.npe * S.nav * sizeof(double)));
    if (c->d.physics == MASTG_HEX8_HEAT) {
        // the conduction residual depends on the state (k(T), convection,
        // source): the element kernel always runs, at T = 0 when no state is given
        if (!X_active) launch_fill(S.n_active, 0.0, c->vT.as<double>(), c->stream);
        TimedRegion t(c, 0);
        launch_heat_K(heat_args(c, X_active ? X_active : c->vT.as<double>(), want_R, want_K), c->stream);
    } else if (fp32) {
        // FP32 element blocks on MFMA (hex8 only), FP64 residual when a state is given
        TimedRegion t(c, 0);
        launch_hex8_K_f32(hex8_args(c, nullptr, false), c->Kel32.as<float>(), c->stream);
        if (want_R && X_active) launch_elements(c, X_active, false);
    } else {
        TimedRegion t(c, 0);
        if (want_K || (want_R && X_active)) launch_elements(c, want_R ? X_active : nullptr, want_K);
    }
    if (want_K) {
        ScatterArgs sa{S.n_nodes, S.nav, c->blk_ptr.as<int64_t>(), c->cptr.as<int32_t>(),
                       c->cidx.as<int32_t>(), c->Kel.as<double>(), c->val.as<double>()};
        {
            TimedRegion t(c, 1);
            if (fp32) launch_scatter_f32(sa, c->Kel32.as<float>(), c->stream);
            else launch_scatter(sa, c->stream);
        }
        launch_jacobi_inv(S.n_active, c->diag_pos.as<int64_t>(), c->val.as<double>(), c->dinv.as<double>(),
                          c->stream);
        c->assembled = true;
        c->val32_valid = false;
        c->kel64_valid = !fp32;
        c->mg_valid = false;
    }
    if (want_R) {
        const bool have_rel = X_active || c->d.physics == MASTG_HEX8_HEAT;
        launch_node_residual(S.n_nodes, S.nav, S.npe, c->ninc_ptr.as<int32_t>(), c->ninc.as<int32_t>(),
                             have_rel ? c->Rel.as<double>() : nullptr, c->F.as<double>(),
                             c->con.as<uint8_t>(), c->vtmp.as<double>(), c->stream);
    }
    HIPC(c, hipGetLastError());
    return MASTG_OK;
}

// ---- FP64 CSR CG + Jacobi on the resident matrix: b (active, device) -> c->vx
int autotune(mastg_ctx* c);
void ensure_scales(mastg_ctx* c);
bool lattice_args(mastg_ctx* c, CGArgs& a);
int cg_finish(mastg_ctx* c, const CGArgs& a, const double* b, hipEvent_t e0, hipEvent_t e1,
              mastg_solve_stats* st, int done);
int solve_stats(mastg_ctx* c, const double* r, const double* b, hipEvent_t e0, hipEvent_t e1,
                mastg_solve_stats* st, int done);

int cg_loop(mastg_ctx* c, const CGArgs& a, const double* b, const mastg_solve_opts* o, int op, int chunk,
            hipEvent_t e0, hipEvent_t e1, mastg_solve_stats* st);

// op: 0 FP64 CSR SpMV, 1 matrix-free FP64, 2 mixed precision (FP32 node-block
// SpMV inside CG whose residual is replaced by the FP64 true residual b - A x
// each time it has dropped by reliable_delta, and at convergence)
int cg_solve(mastg_ctx* c, const double* b, const mastg_solve_opts* o, mastg_solve_stats* st, int op) {
    const bool matfree = op == 1;
    const Structure& S = c->S;
    const double rtol = (o && o->rtol > 0) ? o->rtol : 1e-10;
    const int max_it = (o && o->max_iters > 0) ? o->max_iters : 100000;
    const int chunk = (o && o->check_every > 0) ? o->check_every : 32;
    if (!c->tuned) TRY(autotune(c));
    if (matfree) ensure_scales(c);
    if (op == 2 && !c->val32_valid) {
        TRY(dalloc(c, c->val32, size_t(S.nnz) * sizeof(float)));
        launch_to_f32(S.nnz, c->val.as<double>(), c->val32.as<float>(), c->stream);
        c->val32_valid = true;
    }
    CGArgs a = cg_args(c);
    a.use_mf = op;
    if (matfree) a.nparts_op = mf_node_blocks(S.n_nodes);
    c->cg_replacements = 0;
    hipEvent_t e0, e1;
    HIPC(c, ev_create(&e0));
    HIPC(c, ev_create(&e1));
    HIPC(c, hipEventRecord(e0, c->stream));
    launch_fill(S.n_active, 0.0, a.x, c->stream);
    HIPC(c, hipMemcpyAsync(a.r, b, S.n_active * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    launch_cg_start(a, b, rtol * rtol, max_it, c->stream);
    return cg_loop(c, a, b, o, op, chunk, e0, e1, st);
}

// Captures `body` (kernel launches on c->stream) into *ex.  The capture is
// always ended, also when a step fails; *ex is null unless the graph was
// instantiated.  Capture and the exec's destruction hold the process-wide HIP
// lock (other threads' contexts may capture or allocate at the same time).
template <class Body>
int capture_graph(mastg_ctx* c, hipGraphExec_t* ex, Body&& body) {
    HipLock capture_lock(g_hip_mu);
    if (*ex) (void)hipGraphExecDestroy(*ex);
    *ex = nullptr;
    HIPC(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    body();
    hipGraph_t g = nullptr;
    const hipError_t e_end = hipStreamEndCapture(c->stream, &g);
    if (e_end != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return fail(c, MASTG_ERR_DEVICE, std::string("hipStreamEndCapture failed: ") + hipGetErrorString(e_end));
    }
    hipGraphExec_t out = nullptr;
    const hipError_t e_inst = hipGraphInstantiate(&out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e_inst != hipSuccess) {
        (void)hipGetLastError();
        return fail(c, MASTG_ERR_DEVICE, std::string("hipGraphInstantiate failed: ") + hipGetErrorString(e_inst));
    }
    *ex = out;
    return MASTG_OK;
}

// every captured solver graph of the context (launch shapes changed)
void drop_graphs(mastg_ctx* c) {
    HipLock l(g_hip_mu);
    for (hipGraphExec_t* ex : {&c->cg_exec[0], &c->cg_exec[1], &c->cg_exec[2], &c->bicg_exec, &c->mg_exec,
                               &c->mg_bicg_exec}) {
        if (*ex) (void)hipGraphExecDestroy(*ex);
        *ex = nullptr;
    }
}

// Iterations run in chunks (hipGraph replays or direct launches) between host
// polls of the device scalars.  The recursive residual drifts from b - A x on
// ill-conditioned systems (and with the FP32 operator), so convergence is
// confirmed on the FP64 true residual: when it misses the tolerance it
// replaces the recursive residual and CG restarts from there.  The
// mixed-precision operator also replaces whenever ||r|| has dropped by
// reliable_delta (search direction kept: reliable updates).  Eight replacements in a row that fail to halve the best
// true residual end the solve as not converged (stagnation).
int cg_loop(mastg_ctx* c, const CGArgs& a, const double* b, const mastg_solve_opts* o, int op, int chunk,
            hipEvent_t e0, hipEvent_t e1, mastg_solve_stats* st) {
    const Structure& S = c->S;
    const double delta = (o && o->reliable_delta > 0) ? o->reliable_delta : 0.1;
    if (c->use_graphs && (!c->cg_exec[op] || c->cg_exec_len[op] != chunk)) {
        TRY(capture_graph(c, &c->cg_exec[op], [&]() {
            for (int i = 0; i < chunk; i++) launch_cg_iteration(a, c->stream);
        }));
        c->cg_exec_len[op] = chunk;
    }
    HIPC(c, hipMemcpyAsync(c->h_scal, a.S, sizeof(CGScalars), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    int done = c->h_scal->done;
    double rref2 = c->h_scal->rr;   // ||r||^2 at the last replacement (start: ||b||^2)
    double best2 = rref2;
    int stalls = 0;
    while (!done) {
        if (c->use_graphs) HIPC(c, hipGraphLaunch(c->cg_exec[op], c->stream));
        else
            for (int i = 0; i < chunk; i++) launch_cg_iteration(a, c->stream);
        HIPC(c, hipMemcpyAsync(c->h_scal, a.S, sizeof(CGScalars), hipMemcpyDeviceToHost, c->stream));
        HIPC(c, hipStreamSynchronize(c->stream));
        done = c->h_scal->done;
        if (done >= 2) break;
        if (done == 1 || (op == 2 && c->h_scal->rr <= delta * delta * rref2)) {
            launch_spmv_csr(a, a.x, c->vtmp.as<double>(), 0, c->stream);
            launch_axpby(S.n_active, 1.0, b, -1.0, c->vtmp.as<double>(), c->stream);
            // a confirmation replacement restarts CG from the true residual
            // (p = z): keeping a direction that is not conjugate to the new
            // residual diverges on ill-conditioned plates
            launch_cg_replace(a, c->vtmp.as<double>(), done == 1 ? 1 : 0, c->stream);
            HIPC(c, hipMemcpyAsync(c->h_scal, a.S, sizeof(CGScalars), hipMemcpyDeviceToHost, c->stream));
            HIPC(c, hipStreamSynchronize(c->stream));
            c->cg_replacements++;
            done = c->h_scal->done;
            rref2 = c->h_scal->rr;
            if (!done) {
                if (rref2 < 0.25 * best2) stalls = 0;
                else if (++stalls >= 8) { done = 5; break; }
                best2 = std::min(best2, rref2);
            }
        }
    }
    return cg_finish(c, a, b, e0, e1, st, done);
}

// Element scales E(rho) of the matrix-free operators for the current
// parameters; once the lattice operator is built, also its copy in
// lattice-cell order (one independent load per cell in the operator).
void ensure_scales(mastg_ctx* c) {
    if (!c->sE_valid) {
        launch_elem_scale(c->S.n_elems, c->param.as<double>(), c->d.E, c->d.penalty, c->sE.as<double>(), c->stream);
        c->sE_valid = true;
        c->sEc_valid = false;
    }
    if (!c->sEc_valid && !c->mg.empty() && !c->mg_power) {
        const MGLevelDev& L0 = *c->mg[0];
        launch_gather_values(L0.n_elems, L0.elem_of_cell.as<int32_t>(), c->sE.as<double>(),
                             const_cast<double*>(L0.sE_cell.as<double>()), c->stream);
        c->sEc_valid = true;
    }
}

// ---- geometric multigrid preconditioned CG (lattice meshes) ----------------
// power iteration for the plates' Chebyshev bounds: steps and safety factor
constexpr int kMGPowerIters = 20;
constexpr double kMGPowerSafety = 1.1;
// Chebyshev smoothing interval [lam / ratio, lam], measured per physics:
// hex8 elasticity (Gershgorin bound) 80 (PCG iterations on the scipy
// restatement 20-21 vs 22 with 30 at 32x16x16 and 48x24x24; C2 22 vs 24;
// 100 and 120 the same); the plate (power-iteration bound) 150 (512^2: 290 ms
// per design, the same up to 600, 297 with 80, 322 with 30); heat
// (Gershgorin bound) 5 (96^3 with 3/2 steps: 92.8 ms, 68 iterations; 96.2
// with 7, 93.8 with 10, 100.4 with 15, 94.3 with 3)
constexpr double kMGChebRatioHex8 = 80.0;
constexpr double kMGChebRatioPlate = 150.0;
constexpr double kMGChebRatioHeat = 5.0;

// The hierarchy (lattice detection, coarse meshes, their node-block structure)
// is built on the first multigrid solve; coarse operators are re-formed after
// every assembly (Galerkin products of the element matrices, so they follow
// the densities), all on the device.
int mg_build(mastg_ctx* c) {
    if (c->mg_tried) return c->mg.empty() ? fail(c, MASTG_ERR_INVALID, "multigrid unavailable: " + c->mg_why)
                                          : MASTG_OK;
    c->mg_tried = true;
    // the plate and heat problem levels, and hex8 elasticity on a lattice of
    // cells that are not all congruent (displaced nodes), smooth with their
    // assembled node blocks in node-rank order; hex8 elasticity on one element
    // shape uses the per-node lattice operator (no assembled matrix)
    const bool plate = c->d.physics != MASTG_HEX8_ELASTIC || c->S.n_classes != 1;
    MGHierarchy H = build_mg_hierarchy(c->S);
    if (!H.ok) {
        c->mg_why = H.why;
        return fail(c, MASTG_ERR_INVALID, "multigrid unavailable: " + H.why);
    }
    const int nl = int(H.levels.size());
    std::vector<std::unique_ptr<MGLevelDev>> lv;
    for (int l = 0; l < nl; l++) {
        const MGLevelHost& h = H.levels[l];
        auto L = std::make_unique<MGLevelDev>();
        MGLevelArgs& a = L->a;
        a.n_nodes = h.n_nodes;
        a.nav = h.nav;
        a.dim = h.dim;
        a.nx = h.nx; a.ny = h.ny; a.nz = h.nz;
        L->n_elems = h.n_elems;
        const int nav = h.nav;
        const size_t nv = size_t(h.n_nodes) * nav * sizeof(double);
        if (l == 0 && plate) {
            // plate / heat: the problem level smooths with the assembled node
            // blocks (FP32 copy) in the caller's node-rank order; the transfers
            // map lattice points to ranks
            TRY(upload(c, L->rank_of_lat, h.rank_of_lat));
            TRY(upload(c, L->elem_of_cell, h.elem_of_cell));
            TRY(dalloc(c, c->val32, size_t(c->S.nnz) * sizeof(float)));
            // the level's arguments hold the context's CSR values and inverse
            // diagonal: allocate them here at the assembly's sizes (the
            // assembly then reuses them), since a design evaluation builds the
            // hierarchy before its first assembly
            TRY(dalloc(c, c->val, size_t(c->S.nnz) * sizeof(double)));
            TRY(dalloc(c, c->dinv, size_t(c->S.n_active) * sizeof(double)));
            // heat and hex8 elasticity: the V-cycle runs this level in lattice
            // order (node-major) with a stencil smoother: the stencil's row positions, constraint
            // flags, inverse diagonal and right side in lattice order, the stencil
            // (the plate's node-block smoother measured faster than
            // lattice-ordered stencils: 324 vs 352 and 387 ms per 512^2 design
            // with full stencils, 257 vs 284 ms with the symmetric half one)
            if (nav == 1 || nav == 3) {
                std::vector<int32_t> lat_of_rank(size_t(h.n_nodes));
                for (int64_t p = 0; p < h.n_nodes; p++) lat_of_rank[size_t(h.rank_of_lat[p])] = int32_t(p);
                std::vector<uint8_t> con_lat(size_t(h.n_nodes) * nav);
                for (int64_t p = 0; p < h.n_nodes; p++)
                    for (int v = 0; v < nav; v++)
                        con_lat[size_t(p * nav + v)] = c->S.constrained[size_t(h.rank_of_lat[p]) * nav + v];
                // position of neighbour o of lattice point p in p's row (the
                // stencil is gathered through it at every set-up)
                std::vector<uint8_t> kpos(size_t(h.n_nodes) * 27, uint8_t(255));
                const int64_t nxp = h.nx + 1, nyp = h.ny + 1;
                for (int64_t p = 0; p < h.n_nodes; p++) {
                    const int64_t r = h.rank_of_lat[p], k0 = c->S.blk_ptr[r];
                    const int64_t pi = p % nxp, pj = (p / nxp) % nyp, pk = p / (nxp * nyp);
                    for (int64_t k = k0; k < c->S.blk_ptr[r + 1]; k++) {
                        const int64_t q = lat_of_rank[size_t(c->S.blk_col[k])];
                        const int64_t dx = q % nxp - pi, dy = (q / nxp) % nyp - pj, dz = q / (nxp * nyp) - pk;
                        if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1 || k - k0 > 254) continue;
                        kpos[size_t(p * 27 + (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1))] = uint8_t(k - k0);
                    }
                }
                TRY(upload(c, L->kpos, kpos));
                TRY(upload(c, L->con_lat, con_lat));
                TRY(dalloc(c, L->dinv_lat, nv));
                TRY(dalloc(c, L->b_lat, nv));
                // heat: all 27 entries (the Jacobian is not symmetric); hex8
                // elasticity: the symmetric half of the 27 3 x 3 blocks
                TRY(dalloc(c, L->sten, size_t(h.n_nodes) * (nav == 1 ? 27 : kMGHalfStencil) * sizeof(float)));
                L->lat0 = true;
                // hex8 cells of several shapes: 3 problem-level steps were
                // fastest per design (distorted 128x64x64: 3/2 55 ms, 33
                // iterations; 4/2 56-60; 6/2 64.5, 25 iterations)
                if (nav == 3 && !c->mg_k0_set) c->mg_k0 = 3;
            }
            if (nav == 5) {   // the plate's problem level smooths with block Jacobi
                TRY(dalloc(c, L->dinvb, size_t(h.n_nodes) * 25 * sizeof(float)));
                a.dinvb = L->dinvb.as<float>();
            }
            a.rank_of_lat = L->rank_of_lat.as<int32_t>();
            a.con = c->con.as<uint8_t>();
            a.blk_ptr = c->blk_ptr.as<int64_t>();
            a.blk_col = c->blk_col.as<int32_t>();
            a.val = c->val.as<double>();
            a.val32 = c->val32.as<float>();
            a.dinv = c->dinv.as<double>();
            a.r = c->vr.as<double>();
        } else if (l == 0) {
            // the problem level runs in lattice order (x fastest): its vectors
            // are permuted from / to the caller's node-rank order at the solve
            // (component-major: DOF v of lattice point p at v n_nodes + p)
            std::vector<uint8_t> con_lat(size_t(h.n_nodes) * 3), nodecon_lat(size_t(h.n_nodes));
            for (int64_t p = 0; p < h.n_nodes; p++) {
                uint8_t bits = 0;
                for (int v = 0; v < 3; v++) {
                    const uint8_t cv = c->S.constrained[size_t(h.rank_of_lat[p]) * 3 + v];
                    con_lat[size_t(v * h.n_nodes + p)] = cv;
                    if (cv) bits |= uint8_t(1 << v);
                }
                nodecon_lat[p] = bits;
            }
            TRY(upload(c, L->rank_of_lat, h.rank_of_lat));
            TRY(upload(c, L->elem_of_cell, h.elem_of_cell));
            TRY(upload(c, L->con_lat, con_lat));
            TRY(upload(c, L->nodecon_lat, nodecon_lat));
            TRY(dalloc(c, L->sE_cell, size_t(h.n_elems) * sizeof(double)));
            TRY(dalloc(c, L->dinv_lat, nv));
            TRY(dalloc(c, L->b_lat, nv));
            a.con = L->con_lat.as<uint8_t>();
            a.dinv = L->dinv_lat.as<double>();
            a.r = c->vr.as<double>();
            a.planar = 1;
            // FP32 smoothing state; y_lat: FP64 lattice-ordered scratch for
            // operator products outside the solve
            TRY(dalloc(c, L->x32, nv / 2));
            TRY(dalloc(c, L->res32, nv / 2));
            TRY(dalloc(c, L->da32, nv / 2));
            TRY(dalloc(c, L->db32, nv / 2));
            TRY(dalloc(c, L->dinv32, nv / 2));
            TRY(dalloc(c, L->y_lat, nv));
            a.x32 = L->x32.as<float>();
            a.res32 = L->res32.as<float>();
            a.da32 = L->da32.as<float>();
            a.db32 = L->db32.as<float>();
            a.dinv32 = L->dinv32.as<float>();
            TRY(dalloc(c, c->Khat32, 64 * 9 * sizeof(float)));
            launch_to_f32(64 * 9, c->Khat.as<double>(), c->Khat32.as<float>(), c->stream);
        } else {
            const Structure& S = h.S;
            TRY(upload(c, L->blk_ptr, S.blk_ptr));
            TRY(upload(c, L->blk_col, S.blk_col));
            TRY(upload(c, L->cptr, S.cptr));
            TRY(upload(c, L->cidx, S.cidx));
            TRY(upload(c, L->econn, S.econn));
            TRY(upload(c, L->con, S.constrained));
            const size_t nb2 = size_t(nav) * nav;
            TRY(dalloc(c, L->Kel, size_t(S.n_elems) * S.npe * S.npe * nb2 * sizeof(double)));
            TRY(dalloc(c, L->val, size_t(S.n_blocks) * nb2 * sizeof(double)));
            TRY(dalloc(c, L->val32, size_t(S.n_blocks) * nb2 * sizeof(float)));
            TRY(dalloc(c, L->dinv, nv));
            TRY(dalloc(c, L->r, nv));
            a.con = L->con.as<uint8_t>();
            a.blk_ptr = L->blk_ptr.as<int64_t>();
            a.blk_col = L->blk_col.as<int32_t>();
            a.val = L->val.as<double>();
            // the smoother reads FP32 values (an exactly symmetric FP64 matrix
            // rounds to an exactly symmetric FP32 one); the coarsest level's
            // dense inverse and the bounds use the FP64 values
            a.val32 = L->val32.as<float>();
            a.dinv = L->dinv.as<double>();
            a.r = L->r.as<double>();
        }
        if (!a.planar) {
            TRY(dalloc(c, L->x, nv));
            TRY(dalloc(c, L->res, nv));
            TRY(dalloc(c, L->da, nv));
            TRY(dalloc(c, L->db, nv));
            a.x = L->x.as<double>();
            a.res = L->res.as<double>();
            a.da = L->da.as<double>();
            a.db = L->db.as<double>();
        }
        lv.push_back(std::move(L));
    }
    if (c->d.physics == MASTG_HEX8_HEAT && nl > 2) {
        // heat: the first coarse level (lattice order) smooths with a 27-point
        // stencil too when it is not the coarsest
        TRY(dalloc(c, lv[1]->sten, size_t(H.levels[1].n_nodes) * 27 * sizeof(float)));
    }
    if (!plate) {
        // one-class lattice: the children's Galerkin products of the class
        // matrix, and the first coarse level's cells that touch constraints
        std::vector<double> Kh(64 * 9), G(8 * 576);
        {
            HipLock l(g_hip_mu);
            HIPC(c, hipMemcpy(Kh.data(), c->Khat.p, Kh.size() * sizeof(double), hipMemcpyDeviceToHost));
        }
        mg_class_child_products(Kh.data(), G.data());
        TRY(upload(c, c->Gc, G));
        std::vector<uint8_t> dirty;
        std::vector<int32_t> list;
        mg_dirty_coarse(H.levels[0], H.levels[1], c->S.constrained, dirty, list);
        TRY(upload(c, lv[1]->gal_dirty, dirty));
        if (!list.empty()) TRY(upload(c, lv[1]->gal_list, list));
        lv[1]->n_dirty = int64_t(list.size());
        // the first coarse level smooths with its operator as a stencil
        // (coalesced values, no column indices) when it is not the coarsest
        if (nl > 2) TRY(dalloc(c, lv[1]->sten, size_t(H.levels[1].n_nodes) * 123 * sizeof(float)));
    }
    TRY(dalloc(c, c->mg_cheb, size_t(nl) * kMGChebStride * sizeof(double)));
    TRY(dalloc(c, c->mg_lam, size_t(nl) * sizeof(double)));
    TRY(dalloc(c, c->mg_part, size_t(std::max<int64_t>(8192, (c->S.n_active + 255) / 256 + 1)) * sizeof(double)));
    c->mg_n_coarse = H.coarsest_dofs;
    TRY(dalloc(c, c->mg_dense, size_t(H.coarsest_dofs * H.coarsest_dofs) * sizeof(double)));
    TRY(dalloc(c, c->mg_scal, 2 * sizeof(double)));
    // plates: the Gershgorin bound of D^-1 A is loose by ~50x (thin-plate
    // bending against membrane rows); their Chebyshev intervals come from
    // power iteration instead (heat too: its Jacobian is not symmetric)
    c->mg_power = plate;
    for (int l = 0; l < nl; l++) lv[l]->a.cheb = c->mg_cheb.as<double>() + l * kMGChebStride;
    c->mg = std::move(lv);
    return MASTG_OK;
}

// coarse operators, smoothing bounds and the coarsest inverse for the current
// assembly (needs the FP64 element blocks and CSR values of that assembly)
int mg_setup(mastg_ctx* c) {
    if (c->mg_valid) return MASTG_OK;
    // hex8 on a lattice (one element class): the problem level is matrix-free,
    // its row information and the first Galerkin product come from the class
    // matrix and the cell scales; plates and heat read their FP64 assembly
    const bool lat = !c->mg_power;
    if (!lat && !c->kel64_valid) return fail(c, MASTG_ERR_INVALID, "multigrid needs an FP64 assembly");
    const int nl = int(c->mg.size());
    const int nav = c->S.nav;
    hipStream_t s = c->stream;
    MGLevelDev& L0 = *c->mg[0];
    if (lat) {
        ensure_scales(c);
        CGArgs la = cg_args(c);
        if (!lattice_args(c, la)) return fail(c, MASTG_ERR_INVALID, "no lattice operator for this mesh");
        launch_lattice_rowinfo(la.lat, L0.dinv_lat.as<double>(), L0.dinv32.as<float>(), c->mg_part.as<double>(),
                               c->mg_lam.as<double>(), s);
    } else {
        // the problem level reads the context's assembly
        L0.a.val = c->val.as<double>();
        L0.a.val32 = c->val32.as<float>();
        L0.a.dinv = c->dinv.as<double>();
        if (!L0.a.val || !L0.a.val32 || !L0.a.dinv) return fail(c, MASTG_ERR_INVALID, "multigrid needs an FP64 assembly");
        launch_mg_rowinfo_nb(nav, c->S.n_nodes, c->blk_ptr.as<int64_t>(), c->blk_col.as<int32_t>(),
                             c->val.as<double>(), nullptr, c->mg_part.as<double>(), c->mg_lam.as<double>(), s);
        launch_to_f32(c->S.nnz, c->val.as<double>(), c->val32.as<float>(), s);
        if (nav == 5 && L0.a.dinvb) {
            // the plate's problem-level smoother reads the values in FP16,
            // each block row scaled by its largest magnitude
            TRY(dalloc(c, c->val16, size_t(c->S.nnz) * sizeof(uint16_t)));
            TRY(dalloc(c, c->rs16, size_t(c->S.n_active) * sizeof(float)));
            launch_to_f16_rows(c->S.n_nodes, nav, c->blk_ptr.as<int64_t>(), c->val.as<double>(),
                               c->val16.as<uint16_t>(), c->rs16.as<float>(), s);
            L0.a.val16 = c->val16.as<uint16_t>();
            L0.a.rs16 = c->rs16.as<float>();
        }
        if (L0.lat0 && nav == 3) {
            // hex8 elasticity on cells of several shapes: the symmetric half
            // stencil and the inverse diagonal of this assembly in lattice
            // order; the Chebyshev bound is the Gershgorin bound above (1
            // PCG iteration more at 128 x 64 x 64 than 1.1 x a 20-step power
            // iteration, 3 ms less per design)
            launch_mg_gather_stencil3(L0.a, L0.a.rank_of_lat, L0.kpos.as<uint8_t>(), L0.sten.as<float>(), s);
            launch_perm_rank_lat(L0.a.n_nodes, nav, L0.a.rank_of_lat, c->dinv.as<double>(), L0.dinv_lat.as<double>(),
                                 1, nullptr, s);
        } else if (L0.lat0) {
            // heat: the lattice-ordered stencil and inverse diagonal of this
            // assembly; the Chebyshev bound is the Gershgorin bound above
            launch_mg_gather_stencil(L0.a, L0.a.rank_of_lat, L0.kpos.as<uint8_t>(), L0.sten.as<float>(), s);
            launch_perm_rank_lat(L0.a.n_nodes, nav, L0.a.rank_of_lat, c->dinv.as<double>(), L0.dinv_lat.as<double>(),
                                 1, nullptr, s);
            // the smoother reads it in FP16, each point's row scaled by its
            // largest magnitude
            TRY(dalloc(c, L0.sten16, size_t(L0.a.n_nodes) * 27 * sizeof(uint16_t)));
            TRY(dalloc(c, L0.srow16, size_t(L0.a.n_nodes) * sizeof(float)));
            launch_sten27_to_f16(L0.a.n_nodes, L0.sten.as<float>(), L0.sten16.as<uint16_t>(), L0.srow16.as<float>(), s);
            L0.a.val16 = L0.sten16.as<uint16_t>();
            L0.a.rs16 = L0.srow16.as<float>();
        } else {
            if (L0.a.dinvb) launch_mg_block_dinv5(L0.a, L0.dinvb.as<float>(), s);
            launch_mg_power(L0.a, c->val.as<double>(), kMGPowerIters, kMGPowerSafety, L0.a.da, L0.a.db,
                            c->mg_part.as<double>(), c->mg_scal.as<double>(), c->mg_lam.as<double>(), s);
        }
    }
    for (int l = 0; l + 1 < nl; l++) {
        MGLevelDev& F = *c->mg[l];
        MGLevelDev& C = *c->mg[l + 1];
        MGGalerkinArgs g{};
        g.dim = C.a.dim;
        g.nav = nav;
        g.n_coarse = C.n_elems;
        g.cnx = C.a.nx; g.cny = C.a.ny;
        g.fnx = F.a.nx; g.fny = F.a.ny;
        g.f_elem_of_cell = l == 0 ? F.elem_of_cell.as<int32_t>() : nullptr;
        g.f_econn = l == 0 ? c->econn.as<int32_t>() : F.econn.as<int32_t>();
        g.f_con = l == 0 ? c->con.as<uint8_t>() : F.a.con;   // by node rank, as f_econn
        g.c_econn = C.econn.as<int32_t>();
        g.c_con = C.a.con;
        g.Kf = l == 0 ? (lat ? nullptr : c->Kel.as<double>()) : F.Kel.as<double>();
        g.Khat = l == 0 && lat ? c->Khat.as<double>() : nullptr;
        g.sE_cell = l == 0 && lat ? L0.sE_cell.as<double>() : nullptr;
        g.Kc = C.Kel.as<double>();
        if (l == 0 && lat) {
            // cells clear of constraints from the class products, the rest masked
            g.Gc = c->Gc.as<double>();
            g.dirty = C.gal_dirty.as<uint8_t>();
            launch_mg_galerkin_class(g, s);
            g.elems = C.gal_list.as<int32_t>();
            g.n_elems_list = C.n_dirty;
            if (C.n_dirty > 0) launch_mg_galerkin(g, s);
        } else {
            launch_mg_galerkin(g, s);
        }
        ScatterArgs sa{C.a.n_nodes, nav, C.a.blk_ptr, C.cptr.as<int32_t>(), C.cidx.as<int32_t>(),
                       C.Kel.as<double>(), C.val.as<double>()};
        launch_scatter(sa, s);
        launch_mg_rowinfo_nb(nav, C.a.n_nodes, C.a.blk_ptr, C.a.blk_col, C.a.val, C.dinv.as<double>(),
                             c->mg_part.as<double>(), c->mg_lam.as<double>() + l + 1, s);
        // heat: the Gershgorin bound above (C4 96^3: 96 ms per design against
        // 104 with 20-step power iterations on every level of every Newton
        // step's hierarchy); plates (a loose Gershgorin bound) and hex8 on
        // cells of several shapes (2 PCG iterations fewer than with the
        // Gershgorin bound at 128 x 64 x 64): power iteration
        if (c->d.physics != MASTG_HEX8_HEAT)
            launch_mg_power(C.a, C.a.val, kMGPowerIters, kMGPowerSafety, C.a.da, C.a.db, c->mg_part.as<double>(),
                            c->mg_scal.as<double>(), c->mg_lam.as<double>() + l + 1, s);
        launch_to_f32(int64_t(C.val32.bytes / sizeof(float)), C.a.val, C.val32.as<float>(), s);
        if (nav == 5 && l == 0) {
            // the plate's first coarse level smooths in FP16 too (row scales)
            TRY(dalloc(c, C.sten16, size_t(C.val32.bytes / sizeof(float)) * sizeof(uint16_t)));
            TRY(dalloc(c, C.srow16, size_t(C.a.n_nodes) * 5 * sizeof(float)));
            launch_to_f16_rows(C.a.n_nodes, nav, C.a.blk_ptr, C.a.val, C.sten16.as<uint16_t>(), C.srow16.as<float>(), s);
            C.a.val16 = C.sten16.as<uint16_t>();
            C.a.rs16 = C.srow16.as<float>();
        }
        if (C.sten.p) {
            if (nav == 1) launch_mg_to_stencil27(C.a, C.sten.as<float>(), s);
            else launch_mg_to_stencil(C.a, C.sten.as<float>(), s);
        }
    }
    // hex8 cells of several shapes (assembled problem level): [lam/30, lam]
    // (33 iterations at 3 steps on the distorted 128x64x64 against 35 at the
    // lattice's 80; 20 and 45 the same as 30)
    const bool hex8_assembled = c->d.physics == MASTG_HEX8_ELASTIC && c->mg[0]->lat0;
    const double ratio = c->mg_ratio > 0                        ? double(c->mg_ratio)
                         : hex8_assembled                         ? 30.0
                         : c->d.physics == MASTG_HEX8_ELASTIC    ? kMGChebRatioHex8
                         : c->d.physics == MASTG_QUAD4_MINDLIN ? kMGChebRatioPlate
                                                               : kMGChebRatioHeat;
    launch_mg_cheb(c->mg_lam.as<double>(), c->mg_cheb.as<double>(), nl, std::max(c->mg_k0, c->mg_kc), ratio, s);
    const MGLevelDev& Lc = *c->mg[nl - 1];
    const int64_t n = c->mg_n_coarse;
    launch_fill(n * n, 0.0, c->mg_dense.as<double>(), s);
    launch_mg_dense_gather(Lc.a, c->mg_dense.as<double>(), s);
    launch_mg_invert(int(n), c->mg_dense.as<double>(), s);
    HIPC(c, hipGetLastError());
    c->mg_valid = true;
    return MASTG_OK;
}

// the per-node lattice operator's arguments (built multigrid hierarchy, one
// element class); false when the mesh does not qualify
bool lattice_args(mastg_ctx* c, CGArgs& a) {
    if (c->mg.empty() || c->mg_power) return false;   // hex8 only
    const MGLevelDev& L0 = *c->mg[0];
    if (L0.a.n_nodes >= (int64_t(1) << 29)) return false;   // 32-bit byte offsets in the operator
    a.lat.nx = L0.a.nx; a.lat.ny = L0.a.ny; a.lat.nz = L0.a.nz;
    a.lat.sE_cell = L0.sE_cell.as<double>();
    a.lat.Khat = c->Khat.as<double>();
    a.lat.Khat32 = c->Khat32.as<float>();
    a.lat.nodecon = L0.nodecon_lat.as<uint8_t>();
    return lattice_partials(a.lat) <= int64_t(c->part0.bytes / sizeof(double));
}

// one V-cycle on the problem-level vector r0 (null: the level-0 args' r); the
// level-0 output is x0 + dlast0 (formed by the caller): the returned dlast0
// is FP32 on the hex8 problem level (its smoothing state), else FP64
// (started: the caller has written the problem level's pre-smoothing start)
const void* mg_vcycle(mastg_ctx* c, const CGArgs& a, const int* done, hipStream_t s, const double* r0 = nullptr,
                      bool started = false) {
    const int nl = int(c->mg.size());
    const MGLevelDev& D0 = *c->mg[0];
    MGLevelArgs L0 = D0.a;
    if (r0) L0.r = const_cast<double*>(r0);
    const bool f32 = L0.planar != 0;
    if (D0.lat0) {
        // heat and hex8 on cells of several shapes: the problem level runs in
        // lattice order (its output is mapped back to rank order by the
        // caller: mg_rz / mg_bicg_iteration)
        launch_perm_rank_lat(L0.n_nodes, L0.nav, L0.rank_of_lat, L0.r, D0.b_lat.as<double>(), 1, done, s);
        L0.r = D0.b_lat.as<double>();
        L0.dinv = D0.dinv_lat.as<double>();
        L0.con = D0.con_lat.as<uint8_t>();
        L0.rank_of_lat = nullptr;
    }
    auto lv = [&](int l) -> const MGLevelArgs& { return l == 0 ? L0 : c->mg[l]->a; };
    auto dbuf = [&](int l, int i) { return (i % 2 == 0) ? lv(l).da : lv(l).db; };
    auto dbuf32 = [&](int i) { return (i % 2 == 0) ? L0.da32 : L0.db32; };
    // smoothing step from direction buffer iin to iout (-1: none)
    auto apply = [&](int l, int mode, int step, int iin, int iout) {
        if (l == 0 && f32)
            launch_mg_fine(a, L0, mode, step, iin < 0 ? nullptr : dbuf32(iin), iout < 0 ? nullptr : dbuf32(iout),
                           done, s);
        else if (l == 0 && D0.lat0 && L0.nav == 3)
            launch_mg_stencil(L0, D0.sten.as<float>(), mode, step, iin < 0 ? nullptr : dbuf(l, iin),
                              iout < 0 ? nullptr : dbuf(l, iout), done, s);
        else if (l == 0 && D0.lat0)
            launch_mg_stencil_rank(L0, D0.sten.as<float>(), mode, step, iin < 0 ? nullptr : dbuf(l, iin),
                                   iout < 0 ? nullptr : dbuf(l, iout), done, s);
        else if (c->mg[l]->sten.p && lv(l).nav == 1)
            launch_mg_stencil_rank(lv(l), c->mg[l]->sten.as<float>(), mode, step, iin < 0 ? nullptr : dbuf(l, iin),
                                   iout < 0 ? nullptr : dbuf(l, iout), done, s);
        else if (c->mg[l]->sten.p)
            launch_mg_stencil(lv(l), c->mg[l]->sten.as<float>(), mode, step, iin < 0 ? nullptr : dbuf(l, iin),
                              iout < 0 ? nullptr : dbuf(l, iout), done, s);
        else
            launch_mg_bsr(lv(l), mode, step, iin < 0 ? nullptr : dbuf(l, iin), iout < 0 ? nullptr : dbuf(l, iout),
                          done, s);
    };
    auto steps = [&](int l) { return l == 0 ? c->mg_k0 : c->mg_kc; };
    for (int l = 0; l + 1 < nl; l++) {
        const int k = steps(l);
        // coarse levels start their pre-smoothing inside the restriction
        if (l == 0 && !started) launch_mg_pre_init(lv(l), done, s);
        for (int i = 1; i < k; i++) apply(l, MG_STEP, i, i - 1, i);
        apply(l, MG_FINAL, 0, k - 1, -1);
        launch_mg_restrict(lv(l), lv(l + 1), done, s);
    }
    const MGLevelArgs& Lc = lv(nl - 1);
    launch_mg_coarse_solve(int(c->mg_n_coarse), c->mg_dense.as<double>(), Lc.r, Lc.x, done, s);
    const double* dlast = nullptr;   // last direction of level l + 1 (coarse levels are FP64)
    for (int l = nl - 2; l >= 0; l--) {
        const int k = steps(l);
        launch_mg_prolong(lv(l), lv(l + 1), dlast, done, s);
        apply(l, MG_RESID_INIT, 0, -1, 0);
        for (int i = 1; i < k; i++) apply(l, MG_STEP, i, i - 1, i);
        if (l > 0) dlast = dbuf(l, k - 1);
    }
    const int k0 = steps(0);
    return f32 ? static_cast<const void*>(dbuf32(k0 - 1)) : static_cast<const void*>(dbuf(0, k0 - 1));
}

// z = x0 + dlast0 on the problem level and the partial sums of r.z
void mg_rz(mastg_ctx* c, const CGArgs& a, const void* dl, hipStream_t s) {
    const MGLevelDev& D0 = *c->mg[0];
    const MGLevelArgs& L0 = D0.a;
    if (L0.planar) {
        launch_pcg_rz32(a, L0.x32, static_cast<const float*>(dl), s);
    } else if (D0.lat0) {   // z = x0 + dlast back in rank order, then r.z
        launch_mg_combine_rank(L0.n_nodes, L0.nav, L0.rank_of_lat, L0.x, static_cast<const double*>(dl), a.z,
                               &a.S->done, s);
        launch_pcg_rz(a, a.z, nullptr, s);
    } else {
        launch_pcg_rz(a, L0.x, static_cast<const double*>(dl), s);
    }
}

void mg_pcg_iteration(mastg_ctx* c, const CGArgs& a, hipStream_t s) {
    const int* done = &a.S->done;
    launch_cg_op_alpha(a, s);
    const MGLevelArgs& L0 = c->mg[0]->a;
    if (L0.planar) launch_pcg_update_mg(a, L0, s);   // also starts the level-0 smoothing
    else launch_pcg_update(a, s);
    mg_rz(c, a, mg_vcycle(c, a, done, s, nullptr, L0.planar != 0), s);
    launch_pcg_beta(a, s);
    launch_cg_p(a, s);
}

int mg_solve(mastg_ctx* c, const double* b, const mastg_solve_opts* o, mastg_solve_stats* st) {
    const Structure& S = c->S;
    const double rtol = (o && o->rtol > 0) ? o->rtol : 1e-10;
    const int max_it = (o && o->max_iters > 0) ? o->max_iters : 100000;
    const int chunk = (o && o->check_every > 0) ? o->check_every : 4;
    TRY(mg_build(c));
    if (c->mg_power && !c->tuned) TRY(autotune(c));   // launch shape of the CSR products
    if (!c->mg_power) ensure_scales(c);
    hipEvent_t e0, e1;
    HIPC(c, ev_create(&e0));
    HIPC(c, ev_create(&e1));
    HIPC(c, hipEventRecord(e0, c->stream));
    TRY(mg_setup(c));
    CGArgs a = cg_args(c);
    const int32_t* lat_map = c->mg[0]->rank_of_lat.as<int32_t>();
    const double* bl = b;
    if (!c->mg_power) {
        // hex8: the whole PCG runs in lattice order with the per-node lattice
        // operator; b is permuted in, x out
        if (!lattice_args(c, a)) return fail(c, MASTG_ERR_INVALID, "partial-sum buffer too small for the lattice");
        a.use_mf = 3;
        a.nparts_op = int(lattice_partials(a.lat));
        double* blat = c->mg[0]->b_lat.as<double>();
        launch_perm_nodes(S.n_nodes, lat_map, b, blat, 0, c->stream);
        bl = blat;
    } else {
        // plate: node-rank order, the FP64 product over the CSR values in
        // node blocks (one column index per 5 x 5 block instead of per value)
        a.use_mf = 4;
    }
    c->cg_replacements = 0;
    launch_fill(S.n_active, 0.0, a.x, c->stream);
    HIPC(c, hipMemcpyAsync(a.r, bl, S.n_active * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    launch_pcg_start(a, bl, rtol * rtol, max_it, c->stream);
    mg_rz(c, a, mg_vcycle(c, a, &a.S->done, c->stream), c->stream);
    launch_pcg_start2(a, c->stream);
    HIPC(c, hipMemcpyAsync(a.p, a.z, S.n_active * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPC(c, hipMemcpyAsync(c->h_scal, a.S, sizeof(CGScalars), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    int done = c->h_scal->done;
    if (c->use_graphs && !done && (!c->mg_exec || c->mg_exec_len != chunk)) {
        TRY(capture_graph(c, &c->mg_exec, [&]() {
            for (int i = 0; i < chunk; i++) mg_pcg_iteration(c, a, c->stream);
        }));
        c->mg_exec_len = chunk;
    }
    double best2 = 1e300;
    int stalls = 0;
    for (int launched = 0; !done && launched < max_it + chunk; launched += chunk) {
        if (c->use_graphs) HIPC(c, hipGraphLaunch(c->mg_exec, c->stream));
        else
            for (int i = 0; i < chunk; i++) mg_pcg_iteration(c, a, c->stream);
        HIPC(c, hipMemcpyAsync(c->h_scal, a.S, sizeof(CGScalars), hipMemcpyDeviceToHost, c->stream));
        HIPC(c, hipStreamSynchronize(c->stream));
        done = c->h_scal->done;
        if (done == 1) {
            // confirm on the true residual b - A x (the recursion drifts on
            // ill-conditioned plates); when it misses, PCG restarts from it
            CGArgs an = a;
            an.S = nullptr;
            if (a.use_mf == 3) launch_lattice_apply(an, a.x, c->vtmp.as<double>(), c->stream);
            else launch_spmv_csr(a, a.x, c->vtmp.as<double>(), 0, c->stream);
            launch_axpby(S.n_active, 1.0, bl, -1.0, c->vtmp.as<double>(), c->stream);
            HIPC(c, hipMemcpyAsync(a.r, c->vtmp.p, S.n_active * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            launch_pcg_recheck(a, c->stream);
            HIPC(c, hipMemcpyAsync(c->h_scal, a.S, sizeof(CGScalars), hipMemcpyDeviceToHost, c->stream));
            HIPC(c, hipStreamSynchronize(c->stream));
            done = c->h_scal->done;
            if (done == 0) {
                c->cg_replacements++;
                const double rr = c->h_scal->rr;
                if (rr < 0.25 * best2) stalls = 0;
                else if (++stalls >= 8) { done = 5; break; }
                best2 = std::min(best2, rr);
                mg_rz(c, a, mg_vcycle(c, a, &a.S->done, c->stream), c->stream);
                launch_pcg_start2(a, c->stream);
                HIPC(c, hipMemcpyAsync(a.p, a.z, S.n_active * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            }
        }
    }
    if (!c->mg_power) {
        // the true residual in lattice order with the FP64 lattice operator,
        // then x back to rank order (x lives in c->vx for the callers)
        HIPC(c, hipEventRecord(e1, c->stream));
        CGArgs an = a;
        an.S = nullptr;
        launch_lattice_apply(an, a.x, c->vtmp.as<double>(), c->stream);
        launch_axpby(S.n_active, 1.0, bl, -1.0, c->vtmp.as<double>(), c->stream);
        launch_perm_nodes(S.n_nodes, lat_map, a.x, a.z, 1, c->stream);
        HIPC(c, hipMemcpyAsync(a.x, a.z, S.n_active * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        HIPC(c, hipGetLastError());
        return solve_stats(c, c->vtmp.as<double>(), bl, e0, e1, st, done);
    }
    HIPC(c, hipGetLastError());
    return cg_finish(c, a, b, e0, e1, st, done);
}

int cg_finish(mastg_ctx* c, const CGArgs& a, const double* b, hipEvent_t e0, hipEvent_t e1,
              mastg_solve_stats* st, int done) {
    const Structure& S = c->S;
    HIPC(c, hipEventRecord(e1, c->stream));
    // true residual ||b - A x|| / ||b|| with the assembled CSR matrix (for the
    // matrix-free operator this also checks its result against the CSR)
    launch_spmv_csr(a, a.x, c->vtmp.as<double>(), 0, c->stream);
    launch_axpby(S.n_active, 1.0, b, -1.0, c->vtmp.as<double>(), c->stream);
    return solve_stats(c, c->vtmp.as<double>(), b, e0, e1, st, done);
}

// ||r|| / ||b|| of the true residual r, the solve's device time (e0 .. e1) and
// the outcome of `done`
int solve_stats(mastg_ctx* c, const double* r, const double* b, hipEvent_t e0, hipEvent_t e1,
                mastg_solve_stats* st, int done) {
    double rr = 0, bb = 0;
    TRY(dot_active(c, r, r, &rr));
    TRY(dot_active(c, b, b, &bb));
    float ms = 0;
    HIPC(c, hipEventElapsedTime(&ms, e0, e1));
    (void)ev_destroy(e0);
    (void)ev_destroy(e1);
    if (st) {
        st->iters = c->h_scal->iter;
        st->replacements = c->cg_replacements;
        st->rel_res = bb > 0 ? std::sqrt(rr / bb) : 0.0;
        st->seconds = ms * 1e-3;
    }
    if (done == 3 || done == 4) return fail(c, MASTG_ERR_NAN, "CG breakdown or non-finite iterate");
    if (done == 5) return fail(c, MASTG_ERR_NOT_CONVERGED, "CG stagnated above the tolerance");
    if (done != 1) return fail(c, MASTG_ERR_NOT_CONVERGED, "CG did not reach the tolerance");
    return MASTG_OK;
}



BiCGArgs bicg_args(mastg_ctx* c, bool transpose) {
    BiCGArgs b{};
    b.n = c->S.n_active;
    b.row_ptr = c->row_ptr.as<int64_t>();
    b.col = c->col.as<int32_t>();
    b.val = transpose ? c->valT.as<double>() : c->val.as<double>();
    b.dinv = c->dinv.as<double>();
    b.x = c->vx.as<double>(); b.r = c->vr.as<double>(); b.r0 = c->vr0.as<double>();
    b.p = c->vp.as<double>(); b.v = c->vv.as<double>(); b.s = c->vs.as<double>();
    b.t = c->vt.as<double>(); b.ph = c->vph.as<double>(); b.sh = c->vsh.as<double>();
    b.part0 = c->part0.as<double>();
    b.part1 = c->part1.as<double>();
    b.nparts = c->nparts;
    b.nvec = std::min<int64_t>(2048, std::max<int64_t>(1, (c->S.n_active + 255) / 256));
    b.spmv_sw = c->spmv_sw;
    b.S = c->bscal.as<BiCGScalars>();
    return b;
}

int ensure_bicg(mastg_ctx* c) {
    const size_t nb = size_t(c->S.n_active) * sizeof(double);
    for (DevBuf* b : {&c->vr0, &c->vv, &c->vs, &c->vt, &c->vph, &c->vsh}) TRY(dalloc(c, *b, nb));
    TRY(dalloc(c, c->bscal, sizeof(BiCGScalars)));
    return MASTG_OK;
}

// BiCGStab + Jacobi on the resident matrix (or its transpose): b -> c->vx.
// The true residual ||b - A x|| / ||b|| is recomputed at the end.
// BiCGStab iteration right-preconditioned by the multigrid V-cycle: ph = M p,
// sh = M s replace the Jacobi scalings of the plain iteration
void mg_bicg_iteration(mastg_ctx* c, const BiCGArgs& a, hipStream_t s) {
    const int* done = &a.S->done;
    const CGArgs none{};
    const int64_t n = c->S.n_active;
    launch_bicg_part(a, 0, s);
    // (plate and heat: FP64 problem-level state)
    const MGLevelDev& D0 = *c->mg[0];
    auto combine = [&](const double* dl, double* out) {
        if (D0.lat0) launch_mg_combine_rank(D0.a.n_nodes, D0.a.nav, D0.a.rank_of_lat, D0.a.x, dl, out, done, s);
        else launch_mg_combine(n, D0.a.x, dl, out, done, s);
    };
    combine(static_cast<const double*>(mg_vcycle(c, none, done, s, a.p)), a.ph);
    launch_bicg_part(a, 1, s);
    combine(static_cast<const double*>(mg_vcycle(c, none, done, s, a.s)), a.sh);
    launch_bicg_part(a, 2, s);
}

int bicg_solve(mastg_ctx* c, const double* b, const mastg_solve_opts* o, mastg_solve_stats* st,
               bool transpose, bool mg = false) {
    const Structure& S = c->S;
    const double rtol = (o && o->rtol > 0) ? o->rtol : 1e-10;
    const int max_it = (o && o->max_iters > 0) ? o->max_iters : 100000;
    const int chunk = (o && o->check_every > 0) ? o->check_every : 16;
    if (!c->tuned) TRY(autotune(c));
    TRY(ensure_bicg(c));
    if (transpose) {
        if (c->S.nav != 1) return fail(c, MASTG_ERR_INVALID, "transpose solve needs a scalar field");
        TRY(dalloc(c, c->valT, size_t(S.nnz) * sizeof(double)));
        launch_gather_values(S.nnz, c->tpos.as<int32_t>(), c->val.as<double>(), c->valT.as<double>(),
                             c->stream);
    }
    BiCGArgs a = bicg_args(c, transpose);
    hipEvent_t e0, e1;
    HIPC(c, ev_create(&e0));
    HIPC(c, ev_create(&e1));
    HIPC(c, hipEventRecord(e0, c->stream));
    if (mg) {
        TRY(mg_build(c));
        TRY(mg_setup(c));
    }
    launch_bicg_start(a, b, rtol * rtol, max_it, c->stream);
    int done = 0;
    // graphs: one per (preconditioner, operator) pair
    hipGraphExec_t& ex = mg ? c->mg_bicg_exec : c->bicg_exec;
    int& ex_len = mg ? c->mg_bicg_exec_len : c->bicg_exec_len;
    bool& ex_T = mg ? c->mg_bicg_exec_T : c->bicg_exec_T;
    auto iteration = [&]() {
        if (mg) mg_bicg_iteration(c, a, c->stream);
        else launch_bicg_iteration(a, c->stream);
    };
    if (c->use_graphs) {
        if (!ex || ex_len != chunk || ex_T != transpose) {
            TRY(capture_graph(c, &ex, [&]() {
                for (int i = 0; i < chunk; i++) iteration();
            }));
            ex_len = chunk;
            ex_T = transpose;
        }
    }
    BiCGScalars hs{};
    for (int launched = 0; launched < max_it + chunk; launched += chunk) {
        if (c->use_graphs) HIPC(c, hipGraphLaunch(ex, c->stream));
        else
            for (int i = 0; i < chunk; i++) iteration();
        HIPC(c, hipMemcpyAsync(c->h_pin, a.S, sizeof(BiCGScalars), hipMemcpyDeviceToHost, c->stream));
        HIPC(c, hipStreamSynchronize(c->stream));
        std::memcpy(&hs, c->h_pin, sizeof(BiCGScalars));
        done = hs.done;
        if (done) break;
    }
    HIPC(c, hipEventRecord(e1, c->stream));
    // true residual with the operator that was solved
    CGArgs ca = cg_args(c);
    ca.val = a.val;
    launch_spmv_csr(ca, c->vx.as<double>(), c->vtmp.as<double>(), 0, c->stream);
    launch_axpby(S.n_active, 1.0, b, -1.0, c->vtmp.as<double>(), c->stream);
    double rr = 0, bb = 0;
    TRY(dot_active(c, c->vtmp.as<double>(), c->vtmp.as<double>(), &rr));
    TRY(dot_active(c, b, b, &bb));
    float ms = 0;
    HIPC(c, hipEventElapsedTime(&ms, e0, e1));
    (void)ev_destroy(e0);
    (void)ev_destroy(e1);
    if (st) {
        st->iters = hs.iter;
        st->replacements = 0;
        st->rel_res = bb > 0 ? std::sqrt(rr / bb) : 0.0;
        st->seconds = ms * 1e-3;
    }
    if (done == 3 || done == 4) return fail(c, MASTG_ERR_NAN, "BiCGStab breakdown or non-finite iterate");
    if (done != 1) return fail(c, MASTG_ERR_NOT_CONVERGED, "BiCGStab did not reach the tolerance");
    return MASTG_OK;
}

// Launch-shape autotuner: times the CSR SpMV over lanes-per-row x grid size on
// the resident matrix (HIP events, 3 launches each) and keeps the fastest.
// The grid size also sets the CG vector kernels and the partial-sum count.
int autotune(mastg_ctx* c) {
    c->tuned = true;
    if (!c->assembled) return MASTG_OK;
    const int sws[] = {4, 8, 16, 32, 64};
    const int grids[] = {1024, 2048, 4096, 8192};
    double best = 1e30;
    int bsw = c->spmv_sw, bgrid = c->nparts;
    for (int g : grids)
        for (int sw : sws) {
            c->nparts = g;
            c->spmv_sw = sw;
            CGArgs a = cg_args(c);
            launch_spmv_csr(a, a.p, a.q, 0, c->stream);
            HIPC(c, hipEventRecord(c->ev0, c->stream));
            for (int i = 0; i < 3; i++) launch_spmv_csr(a, a.p, a.q, 0, c->stream);
            HIPC(c, hipEventRecord(c->ev1, c->stream));
            HIPC(c, hipEventSynchronize(c->ev1));
            float ms = 0;
            HIPC(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
            if (ms < best) { best = ms; bsw = sw; bgrid = g; }
        }
    c->nparts = bgrid;
    c->spmv_sw = bsw;
    drop_graphs(c);   // every captured solver reads the launch shapes
    return MASTG_OK;
}


// Steady nonlinear conduction: Newton to ||R|| <= rtol_newton ||R_0|| (one fresh
// residual + Jacobian assembly per step, BiCGStab + Jacobi inner solves with
// an Eisenstat-Walker style forcing term), then the mean temperature and its
// adjoint sensitivity: J^T lambda = dJ/dT, dJ/dq_e = lambda_e . int N_e.
int heat_design_eval(mastg_ctx* c, double* dobj, int where, const mastg_solve_opts* o,
                     mastg_design_result* out) {
    const Structure& S = c->S;
    const double newton_rtol = (o && o->rtol > 0) ? o->rtol : 1e-10;
    // the Newton and adjoint BiCGStab solves are preconditioned by the
    // multigrid V-cycle (rebuilt from each Newton Jacobian) or by Jacobi
    const bool heat_mg = o && o->solver == MASTG_SOLVER_MG;
    hipEvent_t e[4];
    for (auto& x : e) HIPC(c, ev_create(&x));
    double t_asm = 0, t_solve = 0;
    float ms = 0;
    launch_fill(S.n_active, 0.0, c->vT.as<double>(), c->stream);
    double r0 = -1, rn = 0;
    int newton = 0, lin_iters = 0, rc = MASTG_OK;
    double last_rel = 0;
    for (int it = 0; it < 50; it++) {
        HIPC(c, hipEventRecord(e[0], c->stream));
        TRY(assemble_active(c, c->vT.as<double>(), true, true));
        HIPC(c, hipEventRecord(e[1], c->stream));
        double rr = 0;
        TRY(dot_active(c, c->vtmp.as<double>(), c->vtmp.as<double>(), &rr));
        HIPC(c, hipEventElapsedTime(&ms, e[0], e[1]));
        t_asm += ms * 1e-3;
        rn = std::sqrt(rr);
        if (r0 < 0) r0 = rn;
        if (!std::isfinite(rn)) return fail(c, MASTG_ERR_NAN, "non-finite residual in Newton");
        if (rn <= newton_rtol * r0 || rn == 0.0) break;
        // b = -R
        HIPC(c, hipMemcpyAsync(c->vb.p, c->vtmp.p, S.n_active * sizeof(double), hipMemcpyDeviceToDevice,
                               c->stream));
        launch_axpby(S.n_active, 0.0, c->vb.as<double>(), -1.0, c->vb.as<double>(), c->stream);
        mastg_solve_opts lo{};
        lo.solver = MASTG_SOLVER_BICGSTAB_JACOBI;
        lo.rtol = std::max(1e-13, std::min(1e-4, 1e-2 * rn / r0));
        lo.max_iters = (o && o->max_iters > 0) ? o->max_iters : 100000;
        mastg_solve_stats st{};
        rc = bicg_solve(c, c->vb.as<double>(), &lo, &st, false, heat_mg);
        t_solve += st.seconds;
        lin_iters += st.iters;
        last_rel = st.rel_res;
        if (rc) return rc;
        launch_axpby(S.n_active, 1.0, c->vx.as<double>(), 1.0, c->vT.as<double>(), c->stream);
        newton = it + 1;
    }
    if (!(rn <= newton_rtol * r0 || rn == 0.0))
        return fail(c, MASTG_ERR_NOT_CONVERGED, "Newton did not reach the tolerance");
    // output and adjoint at the converged state (the resident matrix is J(T*))
    double J = 0;
    TRY(dot_active(c, c->wmean.as<double>(), c->vT.as<double>(), &J));
    HIPC(c, hipEventRecord(e[2], c->stream));
    HIPC(c, hipMemcpyAsync(c->vb.p, c->wmean.p, S.n_active * sizeof(double), hipMemcpyDeviceToDevice,
                           c->stream));
    launch_mask_constrained(S.n_active, c->con.as<uint8_t>(), c->vb.as<double>(), c->stream);
    mastg_solve_opts ao{};
    ao.solver = MASTG_SOLVER_BICGSTAB_JACOBI;
    ao.rtol = 1e-12;
    ao.max_iters = (o && o->max_iters > 0) ? o->max_iters : 100000;
    mastg_solve_stats ast{};
    rc = bicg_solve(c, c->vb.as<double>(), &ao, &ast, true, heat_mg);
    if (rc) return rc;
    lin_iters += ast.iters;
    // dJ/dq_e = -lambda . dR/dq_e = lambda_e . int N_e
    double* gd = (where == MASTG_HOST || !dobj) ? c->vg.as<double>() : dobj;
    {
        TimedRegion t(c, 4);
        launch_heat_source_dot(heat_args(c, c->vT.as<double>(), false, false), c->vx.as<double>(), 1.0, gd,
                               c->stream);
    }
    HIPC(c, hipEventRecord(e[3], c->stream));
    if (dobj && where == MASTG_HOST)
        HIPC(c, hipMemcpyAsync(dobj, gd, S.n_elems * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    float t23 = 0;
    (void)hipEventElapsedTime(&t23, e[2], e[3]);
    for (auto& x : e) (void)ev_destroy(x);
    if (out) {
        out->objective = J;
        out->iters = lin_iters;
        out->newton_iters = newton;
        out->rel_res = last_rel;
        out->t_assemble = t_asm;
        out->t_solve = t_solve;
        out->t_sens = t23 * 1e-3;
    }
    return MASTG_OK;
}

}  // namespace

// ========================================================================
extern "C" {

const char* mastg_last_error(const mastg_ctx* ctx) {
    return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

int mastg_ctx_create(const mastg_problem_desc* desc, mastg_ctx** out) {
    if (!desc || !out) return fail(nullptr, MASTG_ERR_INVALID, "null argument");
    *out = nullptr;
    mastg_ctx* c = nullptr;
    try {
        c = new mastg_ctx;
        c->d = *desc;
        c->d.xyz = nullptr; c->d.conn = nullptr; c->d.node_dof_base = nullptr;
        c->d.constrained_dofs = nullptr; c->d.side_elem = nullptr; c->d.side_index = nullptr;
        c->d.side_id = nullptr; c->d.point_loads = nullptr;
        if (c->d.penalty == 0) c->d.penalty = 3.0;
        // smoothing steps measured fastest per design (128x64x64 cantilever:
        // 6/2 14.8 ms, 5/2 15.7, 7/2 15.8, 5/3 15.1; 512^2 plate: 4/2 290 ms,
        // 5/2 304; 96^3 heat at its interval: 3/2 110 ms, 4/2 122, 2/2 108
        // with more iterations): the coarse levels are launch-latency bound,
        // so extra coarse steps do not pay
        c->mg_k0 = c->d.physics == MASTG_HEX8_ELASTIC ? 6 : c->d.physics == MASTG_HEX8_HEAT ? 3 : 4;
        c->mg_kc = 2;
        std::string m = build_structure(desc->physics, desc->n_nodes, desc->n_elems, desc->xyz, desc->conn,
                                        desc->n_vars, desc->node_dof_base, desc->n_constrained,
                                        desc->constrained_dofs, c->S);
        if (!m.empty()) {
            delete c;
            return fail(nullptr, MASTG_ERR_INVALID, m);
        }
    } catch (const std::bad_alloc&) {
        delete c;
        return fail(nullptr, MASTG_ERR_OOM, "host allocation failed while building the structure");
    }
    const Structure& S = c->S;
    int rc = MASTG_OK;
    auto bail = [&](int r) {
        g_create_error = c->err;
        mastg_ctx_destroy(c);
        return r;
    };
    std::unique_lock<std::recursive_mutex> create_lock(g_hip_mu);
    if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
        g_create_error = "hipStreamCreate failed";
        delete c;
        return MASTG_ERR_DEVICE;
    }
    c->stream = c->own_stream;
    if (scratch_address(&c->scratch_ptr) != hipSuccess) {
        c->err = "hipGetSymbolAddress of the scratch array failed";
        return bail(MASTG_ERR_DEVICE);
    }
    if (hipHostMalloc((void**)&c->h_scal, sizeof(CGScalars)) != hipSuccess ||
        hipHostMalloc((void**)&c->h_pin, 64 * sizeof(double)) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        c->err = "pinned host allocation / event creation failed";
        return bail(MASTG_ERR_DEVICE);
    }
    create_lock.unlock();
    material_D3_unitE(c->d.nu, c->Dhat);
    // loads: nodal point loads mapped onto the active layout
    std::vector<double> F(S.n_active, 0.0);
    if (desc->point_loads)
        for (int64_t i = 0; i < S.n_active; i++) F[i] = desc->point_loads[S.caller_dof_of_active[i]];
    if (desc->physics == MASTG_QUAD4_MINDLIN) {
        // plane-stress D and G (IsotropicMaterialProperty StiffnessMatrix2D / TransverseShear,
        // src/property_cards/isotropic_material_property_card.cpp:334-426)
        const double E = c->d.E, nu = c->d.nu;
        std::memset(c->Dps, 0, sizeof(c->Dps));
        c->Dps[0] = c->Dps[4] = E / (1. - nu * nu);
        c->Dps[1] = c->Dps[3] = E * nu / (1. - nu * nu);
        c->Dps[8] = E / 2. / (1. + nu);
        c->Gmod = E / 2. / (1. + nu);
        // uniform pressure along the element's local -z: R_w += p int N
        // (surface_pressure_residual, src/elasticity/structural_element_base.cpp:
        // 1167-1235), i.e. F = -p int N v3 with v3 the frame normal ((0, 0, 1)
        // for an element in a z = const plane); the load of a unit pressure is
        // kept for dR/dp
        const int nav = S.nav;
        std::vector<double> Fp(S.n_active, 0.0);
        {
            const double gq = 0.57735026918962576451;
            for (int64_t e = 0; e < S.n_elems; e++) {
                double XG[4][3], XL[4][3], T[9];
                for (int n = 0; n < 4; n++)
                    for (int i = 0; i < 3; i++) XG[n][i] = S.xyz_r[int64_t(S.econn[e * 4 + n]) * 3 + i];
                quad4_frame(XG, T, XL);
                for (int qj = 0; qj < 2; qj++)
                    for (int qi = 0; qi < 2; qi++) {
                        const double xi = qi ? gq : -gq, eta = qj ? gq : -gq;
                        double N[4], g[4][2], jxw;
                        if (nav == 5) {
                            double dN[4][2];
                            quad4_shape_all(xi, eta, N, dN);
                            double j00 = 0, j01 = 0, j10 = 0, j11 = 0;
                            for (int n = 0; n < 4; n++) {
                                j00 += XG[n][0] * dN[n][0]; j01 += XG[n][0] * dN[n][1];
                                j10 += XG[n][1] * dN[n][0]; j11 += XG[n][1] * dN[n][1];
                            }
                            jxw = std::fabs(j00 * j11 - j01 * j10);
                        } else {
                            jxw = quad4_grad_surface(XL, xi, eta, g, N);
                        }
                        for (int n = 0; n < 4; n++)
                            for (int i = 0; i < 3; i++)
                                Fp[int64_t(S.econn[e * 4 + n]) * nav + i] -= jxw * N[n] * T[i * 3 + 2];
                    }
            }
        }
        for (int64_t i = 0; i < S.n_active; i++) F[i] += c->d.pressure * Fp[i];
        if ((rc = upload(c, c->Fp, Fp))) return bail(rc);
    }
    if (desc->physics == MASTG_HEX8_ELASTIC && desc->pres_bid_mask) {
        // uniform pressure on the marked sides: R -= int N p n dA, n the
        // outward normal (surface_pressure_residual, src/elasticity/
        // solid_element_3d.cpp:815-876), i.e. F += p int N n dA; the load of a
        // unit pressure is kept for dR/dp.  2x2 Gauss on the side; n dA =
        // dx/ds x dx/dt, turned to point away from the element's centroid.
        if (desc->n_sides > 0 && (!desc->side_elem || !desc->side_index || !desc->side_id)) {
            c->err = "pressure sides need side_elem, side_index and side_id";
            return bail(MASTG_ERR_INVALID);
        }
        std::vector<double> Fp(S.n_active, 0.0);
        const double gq = 0.57735026918962576451;
        // libMesh hex8 side -> (fixed reference axis, its value)
        static const int kAxis[6] = {2, 1, 0, 1, 0, 2};
        static const double kVal[6] = {-1.0, -1.0, 1.0, 1.0, -1.0, 1.0};
        for (int64_t sd = 0; sd < desc->n_sides; sd++) {
            if (!((desc->pres_bid_mask >> desc->side_id[sd]) & 1u)) continue;
            const int64_t e = desc->side_elem[sd];
            const int side = desc->side_index[sd];
            if (e < 0 || e >= S.n_elems || side < 0 || side > 5) {
                c->err = "pressure side out of range";
                return bail(MASTG_ERR_INVALID);
            }
            double X[8][3], cen[3] = {0, 0, 0};
            for (int n = 0; n < 8; n++)
                for (int d = 0; d < 3; d++) {
                    X[n][d] = S.xyz_r[int64_t(S.econn[e * 8 + n]) * 3 + d];
                    cen[d] += 0.125 * X[n][d];
                }
            const int ax = kAxis[side], u0 = ax == 0 ? 1 : 0, u1 = ax == 2 ? 1 : 2;
            for (int qj = 0; qj < 2; qj++)
                for (int qi = 0; qi < 2; qi++) {
                    double xi[3];
                    xi[ax] = kVal[side];
                    xi[u0] = qi ? gq : -gq;
                    xi[u1] = qj ? gq : -gq;
                    double N[8], dN[8][3];
                    for (int n = 0; n < 8; n++) {
                        const double sx = (n == 1 || n == 2 || n == 5 || n == 6) ? 1.0 : -1.0;
                        const double sy = (n == 2 || n == 3 || n == 6 || n == 7) ? 1.0 : -1.0;
                        const double sz = (n >= 4) ? 1.0 : -1.0;
                        const double fx = 0.5 * (1 + sx * xi[0]), fy = 0.5 * (1 + sy * xi[1]),
                                     fz = 0.5 * (1 + sz * xi[2]);
                        N[n] = fx * fy * fz;
                        dN[n][0] = 0.5 * sx * fy * fz;
                        dN[n][1] = fx * 0.5 * sy * fz;
                        dN[n][2] = fx * fy * 0.5 * sz;
                    }
                    double a[3] = {0, 0, 0}, b[3] = {0, 0, 0}, xq[3] = {0, 0, 0};
                    for (int n = 0; n < 8; n++)
                        for (int d = 0; d < 3; d++) {
                            a[d] += X[n][d] * dN[n][u0];
                            b[d] += X[n][d] * dN[n][u1];
                            xq[d] += X[n][d] * N[n];
                        }
                    double nda[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
                    if (nda[0] * (xq[0] - cen[0]) + nda[1] * (xq[1] - cen[1]) + nda[2] * (xq[2] - cen[2]) < 0)
                        for (double& v : nda) v = -v;
                    for (int n = 0; n < 8; n++)
                        for (int d = 0; d < 3; d++) Fp[int64_t(S.econn[e * 8 + n]) * 3 + d] += N[n] * nda[d];
                }
        }
        for (int64_t i = 0; i < S.n_active; i++) F[i] += c->d.pressure * Fp[i];
        if ((rc = upload(c, c->Fp, Fp))) return bail(rc);
        c->has_pressure = true;
    }
    std::vector<int32_t> conv(S.constrained.begin(), S.constrained.end());
    if ((rc = upload(c, c->xyz, S.xyz_r)) || (rc = upload(c, c->econn, S.econn)) ||
        (rc = upload(c, c->blk_ptr, S.blk_ptr)) || (rc = upload(c, c->blk_col, S.blk_col)) ||
        (rc = upload(c, c->row_ptr, S.row_ptr)) || (rc = upload(c, c->col, S.col)) ||
        (rc = upload(c, c->cptr, S.cptr)) || (rc = upload(c, c->cidx, S.cidx)) ||
        (rc = upload(c, c->diag_pos, S.diag_pos)) || (rc = upload(c, c->con, S.constrained)) ||
        (rc = upload(c, c->dof_of_active, S.caller_dof_of_active)) ||
        (rc = upload(c, c->ninc_ptr, S.ninc_ptr)) || (rc = upload(c, c->ninc, S.ninc)) ||
        (rc = upload(c, c->F, F)) || (rc = dalloc(c, c->param, S.n_elems * sizeof(double))) ||
        (rc = ensure_vectors(c)))
        return bail(rc);
    if (desc->physics == MASTG_HEX8_HEAT) {
        // convective sides per element and the mean-temperature weights
        std::vector<uint8_t> sm(S.n_elems, 0), rm(S.n_elems, 0);
        if (c->d.rad_bid_mask && !(c->d.rad_emissivity >= 0 && c->d.rad_sigma >= 0)) {
            c->err = "radiation needs a non-negative emissivity and Stefan-Boltzmann constant";
            return bail(MASTG_ERR_INVALID);
        }
        for (int64_t k = 0; k < desc->n_sides; k++) {
            const int32_t e = desc->side_elem[k], sd = desc->side_index[k], id = desc->side_id[k];
            if (e < 0 || e >= S.n_elems || sd < 0 || sd > 5 || id < 0 || id > 31) {
                c->err = "boundary side out of range";
                return bail(MASTG_ERR_INVALID);
            }
            if ((c->d.conv_bid_mask >> id) & 1u) sm[e] |= uint8_t(1u << sd);
            if ((c->d.rad_bid_mask >> id) & 1u) {
                rm[e] |= uint8_t(1u << sd);
                c->has_rad = true;
            }
        }
        if (c->has_rad && (rc = upload(c, c->radmask, rm))) return bail(rc);
        std::vector<double> w(S.n_active, 0.0);
        const double gq = 0.57735026918962576451;
        double V = 0;
        for (int64_t e = 0; e < S.n_elems; e++) {
            double X[8][3];
            for (int n = 0; n < 8; n++)
                for (int i = 0; i < 3; i++) X[n][i] = S.xyz_r[int64_t(S.econn[e * 8 + n]) * 3 + i];
            for (int q = 0; q < 8; q++) {
                const double xi = (q & 1) ? gq : -gq, eta = (q & 2) ? gq : -gq, zeta = (q & 4) ? gq : -gq;
                double N[8], d[8][3], J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                for (int n = 0; n < 8; n++) {
                    const double sx = (n == 1 || n == 2 || n == 5 || n == 6) ? 1.0 : -1.0;
                    const double sy = (n == 2 || n == 3 || n == 6 || n == 7) ? 1.0 : -1.0;
                    const double sz = (n >= 4) ? 1.0 : -1.0;
                    const double fx = 0.5 * (1 + sx * xi), fy = 0.5 * (1 + sy * eta), fz = 0.5 * (1 + sz * zeta);
                    N[n] = fx * fy * fz;
                    d[n][0] = 0.5 * sx * fy * fz; d[n][1] = fx * 0.5 * sy * fz; d[n][2] = fx * fy * 0.5 * sz;
                    for (int i = 0; i < 3; i++)
                        for (int r = 0; r < 3; r++) J[i * 3 + r] += X[n][i] * d[n][r];
                }
                const double det = J[0] * (J[4] * J[8] - J[7] * J[5]) + J[3] * (J[7] * J[2] - J[1] * J[8]) +
                                   J[6] * (J[1] * J[5] - J[4] * J[2]);
                for (int n = 0; n < 8; n++) w[S.econn[e * 8 + n]] += det * N[n];
                V += det;
            }
        }
        for (auto& x : w) x /= V;
        c->volume = V;
        std::vector<int32_t> tp(S.blk_transpose.begin(), S.blk_transpose.end());
        if ((rc = upload(c, c->sidemask, sm)) || (rc = upload(c, c->wmean, w)) || (rc = upload(c, c->tpos, tp)) ||
            (rc = dalloc(c, c->vT, S.n_active * sizeof(double))) || (rc = ensure_bicg(c)))
            return bail(rc);
    }
    if (desc->physics == MASTG_HEX8_HEAT && S.n_classes > 0) {
        // the geometry of each congruence class once (gradients and JxW at the
        // 8 points of its representative element): the element kernel then
        // forms no Jacobian per element
        DevBuf rep_d;
        if ((rc = upload(c, rep_d, S.class_rep)) || (rc = upload(c, c->eclass, S.elem_class)) ||
            (rc = dalloc(c, c->heat_geo, size_t(S.n_classes) * kHeatGeo * sizeof(double))))
            return bail(rc);
        launch_heat_geometry(S.n_classes, c->xyz.as<double>(), c->econn.as<int32_t>(), rep_d.as<int32_t>(),
                             c->heat_geo.as<double>(), c->stream);
        if (hipStreamSynchronize(c->stream) != hipSuccess) {
            c->err = "class geometry kernel failed";
            return bail(MASTG_ERR_DEVICE);
        }
        c->device_bytes -= int64_t(rep_d.bytes);
    }
    if (desc->physics == MASTG_HEX8_ELASTIC && S.n_classes > 0) {
        // class reference matrices at E = 1 from the element kernel itself
        std::vector<int32_t> rep_conn(S.n_classes * 8);
        for (int k = 0; k < S.n_classes; k++)
            for (int a = 0; a < 8; a++) rep_conn[k * 8 + a] = S.econn[int64_t(S.class_rep[k]) * 8 + a];
        std::vector<double> ones(S.n_classes, 1.0);
        std::vector<uint8_t> nodecon(S.n_nodes, 0);
        for (int64_t r = 0; r < S.n_nodes; r++)
            for (int v = 0; v < 3; v++)
                if (S.constrained[r * 3 + v]) nodecon[r] |= uint8_t(1 << v);
        DevBuf rc_d, ones_d;
        if ((rc = upload(c, rc_d, rep_conn)) || (rc = upload(c, ones_d, ones)) ||
            (rc = dalloc(c, c->Khat, size_t(S.n_classes) * 64 * 9 * sizeof(double))) ||
            (rc = upload(c, c->eclass, S.elem_class)) || (rc = upload(c, c->nodecon, nodecon)) ||
            (rc = dalloc(c, c->sE, S.n_elems * sizeof(double))) ||
            (rc = dalloc(c, c->Fel, size_t(S.n_elems) * 24 * sizeof(double))))
            return bail(rc);
        Hex8Args a{};
        a.n_elems = S.n_classes;
        a.xyz = c->xyz.as<double>();
        a.econn = rc_d.as<int32_t>();
        a.param = ones_d.as<double>();
        a.E0 = 1.0;
        a.penalty = 1.0;
        std::memcpy(a.Dhat, c->Dhat, sizeof(a.Dhat));
        a.con = nullptr;
        a.X = nullptr;
        a.Rel = nullptr;
        a.Kel = c->Khat.as<double>();
        launch_hex8_K(a, c->stream);
        if (hipStreamSynchronize(c->stream) != hipSuccess) {
            c->err = "class matrix kernel failed";
            return bail(MASTG_ERR_DEVICE);
        }
        c->device_bytes -= int64_t(rc_d.bytes + ones_d.bytes);
    }
    *out = c;
    return MASTG_OK;
}

int mastg_ctx_destroy(mastg_ctx* c) {
    if (!c) return MASTG_OK;
    HipLock destroy_lock(g_hip_mu);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int i = 0; i < 3; i++)
        if (c->cg_exec[i]) (void)hipGraphExecDestroy(c->cg_exec[i]);
    if (c->bicg_exec) (void)hipGraphExecDestroy(c->bicg_exec);
    if (c->mg_exec) (void)hipGraphExecDestroy(c->mg_exec);
    if (c->mg_bicg_exec) (void)hipGraphExecDestroy(c->mg_bicg_exec);
    if (c->h_scal) (void)hipHostFree(c->h_scal);
    if (c->h_pin) (void)hipHostFree(c->h_pin);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream && c->stream != c->own_stream) (void)hipStreamSynchronize(c->own_stream);
    hipStream_t s = c->own_stream;
    delete c;   // DevBuf destructors free device memory
    if (s) (void)hipStreamDestroy(s);
    return MASTG_OK;
}

int mastg_get_stream(const mastg_ctx* c, void** stream) {
    if (!c || !stream) return MASTG_ERR_INVALID;
    *stream = static_cast<void*>(c->stream);
    return MASTG_OK;
}

int mastg_set_stream(mastg_ctx* c, void* stream) {
    if (!c) return MASTG_ERR_INVALID;
    hipStream_t next = stream ? static_cast<hipStream_t>(stream) : c->own_stream;
    if (next == c->stream) return MASTG_OK;
    // the work already enqueued finishes first; captured solver graphs are not
    // bound to a stream and replay on the new one
    HIPC(c, hipStreamSynchronize(c->stream));
    c->stream = next;
    return MASTG_OK;
}

int mastg_get_sizes(const mastg_ctx* c, mastg_sizes* o) {
    if (!c || !o) return MASTG_ERR_INVALID;
    o->n_dofs = c->S.n_dofs;
    o->n_active = c->S.n_active;
    o->nnz = c->S.nnz;
    o->n_blocks = c->S.n_blocks;
    o->n_active_vars = c->S.nav;
    o->nodes_per_elem = c->S.npe;
    o->device_bytes = c->device_bytes;
    return MASTG_OK;
}

int mastg_get_csr_pattern(const mastg_ctx* c, int64_t* row_ptr, int32_t* col_idx) {
    if (!c) return MASTG_ERR_INVALID;
    // rows / columns in caller order restricted to the active variables: the
    // active index is monotone in the caller index, so the pattern is the
    // same sequence of positions
    if (row_ptr) std::memcpy(row_ptr, c->S.row_ptr.data(), c->S.row_ptr.size() * sizeof(int64_t));
    if (col_idx) std::memcpy(col_idx, c->S.col.data(), c->S.col.size() * sizeof(int32_t));
    return MASTG_OK;
}

int mastg_get_active_map(const mastg_ctx* c, int64_t* a) {
    if (!c || !a) return MASTG_ERR_INVALID;
    for (int64_t i = 0; i < c->S.n_dofs; i++) a[i] = -1;
    for (int64_t i = 0; i < c->S.n_active; i++) a[c->S.caller_dof_of_active[i]] = i;
    return MASTG_OK;
}

int mastg_set_params(mastg_ctx* c, const double* p, int where) {
    if (!c || !p) return MASTG_ERR_INVALID;
    const size_t nb = c->S.n_elems * sizeof(double);
    if (where == MASTG_HOST) {
        // densities and thicknesses must be positive; a heat source may take any
        // finite value (zero: no source, negative: a sink)
        const bool heat = c->d.physics == MASTG_HEX8_HEAT;
        for (int64_t e = 0; e < c->S.n_elems; e++)
            if (!std::isfinite(p[e]) || (!heat && p[e] <= 0.0))
                return fail(c, MASTG_ERR_NAN, "design parameter " + std::to_string(e) +
                                                  (heat ? " is not finite" : " is not a positive finite number"));
        HIPC(c, hipMemcpyAsync(c->param.p, p, nb, hipMemcpyHostToDevice, c->stream));
    } else {
        HIPC(c, hipMemcpyAsync(c->param.p, p, nb, hipMemcpyDeviceToDevice, c->stream));
    }
    c->params_set = true;
    c->assembled = false;
    c->sE_valid = false;
    c->sEc_valid = false;
    c->mg_valid = false;   // the coarse operators follow the parameters
    return MASTG_OK;
}

int mastg_assemble(mastg_ctx* c, const double* X, double* R, double* J, int where, int prec) {
    if (!c) return MASTG_ERR_INVALID;
    if (prec != MASTG_ELEM_FP64 && prec != MASTG_ELEM_FP32)
        return fail(c, MASTG_ERR_INVALID, "unknown element precision");
    if (prec == MASTG_ELEM_FP32 && c->d.physics != MASTG_HEX8_ELASTIC)
        return fail(c, MASTG_ERR_INVALID, "FP32 element matrices are built for hex8 elasticity only");
    const Structure& S = c->S;
    double* Xa = c->vb.as<double>();
    TRY(to_active(c, X, where, Xa));
    TRY(assemble_active(c, X ? Xa : nullptr, R != nullptr, true, prec == MASTG_ELEM_FP32));
    if (R) TRY(from_active(c, c->vtmp.as<double>(), R, where));
    if (J) {
        HIPC(c, hipMemcpyAsync(J, c->val.p, S.nnz * sizeof(double),
                               where == MASTG_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice,
                               c->stream));
    }
    HIPC(c, hipStreamSynchronize(c->stream));
    return MASTG_OK;
}

static int cg_op(int solver) {
    return solver == MASTG_SOLVER_MATFREE ? 1 : solver == MASTG_SOLVER_MIXED ? 2 : 0;
}
static int check_cg_solver(mastg_ctx* c, int solver) {
    if (solver != MASTG_SOLVER_CG_JACOBI && solver != MASTG_SOLVER_MATFREE && solver != MASTG_SOLVER_MIXED &&
        solver != MASTG_SOLVER_MG)
        return fail(c, MASTG_ERR_INVALID, "unknown solver");
    if (solver == MASTG_SOLVER_MG) return mg_build(c);
    if (solver == MASTG_SOLVER_MATFREE && (c->S.n_classes == 0 || c->d.physics != MASTG_HEX8_ELASTIC))
        return fail(c, MASTG_ERR_INVALID, "matrix-free solver needs hex8 elasticity on a mesh of at most 8 element shapes");
    if (solver == MASTG_SOLVER_MIXED && c->d.physics != MASTG_HEX8_ELASTIC)
        return fail(c, MASTG_ERR_INVALID,
                    "mixed-precision CG is for hex8 elasticity (the plate Jacobian's h^3 bending vs h membrane "
                    "conditioning is beyond what an FP32 operator resolves)");
    return MASTG_OK;
}

// J x = b (b active, in c->vb) with the assembled matrix, x -> c->vx
static int solve_active(mastg_ctx* c, const mastg_solve_opts* o, mastg_solve_stats* st) {
    if (!c->assembled) return fail(c, MASTG_ERR_INVALID, "no assembled matrix");
    if (o && o->transpose) { /* the elastic Jacobian is symmetric */ }
    int solver = o ? o->solver : MASTG_SOLVER_CG_JACOBI;
    if (c->d.physics == MASTG_HEX8_HEAT && solver == MASTG_SOLVER_CG_JACOBI)
        solver = MASTG_SOLVER_BICGSTAB_JACOBI;   // the k(T) Jacobian is not symmetric
    const bool heat_mg = c->d.physics == MASTG_HEX8_HEAT && solver == MASTG_SOLVER_MG;
    if (solver == MASTG_SOLVER_BICGSTAB_JACOBI || heat_mg)
        return bicg_solve(c, c->vb.as<double>(), o, st, o && o->transpose, heat_mg);
    TRY(check_cg_solver(c, solver));
    return solver == MASTG_SOLVER_MG ? mg_solve(c, c->vb.as<double>(), o, st)
                                     : cg_solve(c, c->vb.as<double>(), o, st, cg_op(solver));
}

int mastg_solve(mastg_ctx* c, const double* rhs, double* x, int where, const mastg_solve_opts* o,
                mastg_solve_stats* st) {
    if (!c || !rhs || !x) return MASTG_ERR_INVALID;
    if (!c->assembled) return fail(c, MASTG_ERR_INVALID, "no assembled matrix");
    TRY(to_active(c, rhs, where, c->vb.as<double>()));
    int rc = solve_active(c, o, st);
    int rc2 = from_active(c, c->vx.as<double>(), x, where);
    return rc ? rc : rc2;
}

int mastg_sensitivity_solve(mastg_ctx* c, int param, const double* X, double* dXdp, int where,
                            const mastg_solve_opts* o, mastg_solve_stats* st) {
    if (!c || !X || !dXdp) return MASTG_ERR_INVALID;
    if (!c->assembled) return fail(c, MASTG_ERR_INVALID, "no assembled matrix (assemble the Jacobian at X first)");
    if (o && o->transpose) return fail(c, MASTG_ERR_INVALID, "the forward sensitivity solve is not transposed");
    TRY(to_active(c, X, where, c->vb.as<double>()));
    TRY(global_dRdp(c, param, c->vb.as<double>(), c->vtmp.as<double>()));
    // b = -dR/dp
    launch_axpby(c->S.n_active, -1.0, c->vtmp.as<double>(), 0.0, c->vb.as<double>(), c->stream);
    int rc = solve_active(c, o, st);
    int rc2 = from_active(c, c->vx.as<double>(), dXdp, where);
    return rc ? rc : rc2;
}

int mastg_jacobian_product(mastg_ctx* c, const double* dX, double* JdX, int where, int op) {
    if (!c || !dX || !JdX) return MASTG_ERR_INVALID;
    if (!c->assembled) return fail(c, MASTG_ERR_INVALID, "no assembled matrix");
    if (op == 1 && (c->S.n_classes == 0 || c->d.physics != MASTG_HEX8_ELASTIC))
        return fail(c, MASTG_ERR_INVALID, "no matrix-free operator");
    if (op < 0 || op > 4) return fail(c, MASTG_ERR_INVALID, "unknown operator");
    if (op == 3) TRY(mg_build(c));
    if ((op == 2 || op == 4) && c->d.physics == MASTG_HEX8_HEAT)
        return fail(c, MASTG_ERR_INVALID, "the node-block operators are for the structural Jacobians");
    TRY(to_active(c, dX, where, c->vb.as<double>()));
    if (op == 2 && !c->val32_valid) {
        TRY(dalloc(c, c->val32, size_t(c->S.nnz) * sizeof(float)));
        launch_to_f32(c->S.nnz, c->val.as<double>(), c->val32.as<float>(), c->stream);
        c->val32_valid = true;
    }
    CGArgs a = cg_args(c);
    if (op == 3) {
        if (!lattice_args(c, a)) return fail(c, MASTG_ERR_INVALID, "no lattice operator for this mesh");
        a.S = nullptr;
        ensure_scales(c);
        const MGLevelDev& L0 = *c->mg[0];
        const int32_t* lat_map = L0.rank_of_lat.as<int32_t>();
        launch_mask_constrained(c->S.n_active, c->con.as<uint8_t>(), c->vb.as<double>(), c->stream);
        double* xl = L0.b_lat.as<double>();
        double* yl = L0.y_lat.as<double>();
        launch_perm_nodes(c->S.n_nodes, lat_map, c->vb.as<double>(), xl, 0, c->stream);
        launch_lattice_apply(a, xl, yl, c->stream);
        launch_perm_nodes(c->S.n_nodes, lat_map, yl, c->vtmp.as<double>(), 1, c->stream);
    } else if (op == 2) {
        launch_spmv_nb32(a, c->vb.as<double>(), c->vtmp.as<double>(), 0, c->stream);
    } else if (op == 4) {
        launch_spmv_nb64(a, c->vb.as<double>(), c->vtmp.as<double>(), 0, c->stream);
    } else if (op == 1) {
        ensure_scales(c);
        // element phase reads a copy zeroed on constrained DOFs; the node phase
        // applies the constrained rows (diagonal = elements at the node) to the input
        HIPC(c, hipMemcpyAsync(c->vz.p, c->vb.p, c->S.n_active * sizeof(double), hipMemcpyDeviceToDevice,
                               c->stream));
        launch_mask_constrained(c->S.n_active, c->con.as<uint8_t>(), c->vz.as<double>(), c->stream);
        launch_mf_apply(a, c->vz.as<double>(), c->vb.as<double>(), c->vtmp.as<double>(), 0, c->stream);
    } else {
        launch_spmv_csr(a, c->vb.as<double>(), c->vtmp.as<double>(), 0, c->stream);
    }
    HIPC(c, hipGetLastError());
    return from_active(c, c->vtmp.as<double>(), JdX, where);
}

int mastg_sensitivity(mastg_ctx* c, const double* X, const double* lam, double* g, int where) {
    if (!c || !X || !lam || !g) return MASTG_ERR_INVALID;
    if (!c->params_set) return fail(c, MASTG_ERR_INVALID, "design parameters not set");
    TRY(to_active(c, X, where, c->vb.as<double>()));
    TRY(to_active(c, lam, where, c->vtmp.as<double>()));
    double* gd = where == MASTG_HOST ? c->vg.as<double>() : g;
    {
        TimedRegion t(c, 4);
        launch_sens(c, c->vb.as<double>(), c->vtmp.as<double>(), gd);
    }
    if (where == MASTG_HOST)
        HIPC(c, hipMemcpyAsync(g, gd, c->S.n_elems * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return MASTG_OK;
}

}  // extern "C"

namespace {

// d D3(E = 1, nu) / d nu: derivative of the isotropic 3D material matrix at unit
// modulus (IsotropicMaterialProperty::StiffnessMatrix3D::derivative,
// src/property_cards/isotropic_material_property_card.cpp:533-567, with dE = 0)
void material_dD3_dnu_unitE(double nu, double D[36]) {
    std::memset(D, 0, 36 * sizeof(double));
    const double d = 1. - nu - 2. * nu * nu, dd = -1. - 4. * nu;
    const double diag = (-d - (1. - nu) * dd) / (d * d);   // d/dnu (1 - nu) / d
    const double off = (d - nu * dd) / (d * d);            // d/dnu nu / d
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) D[i * 6 + j] = (i == j) ? diag : off;
        D[(i + 3) * 6 + i + 3] = -0.5 / ((1. + nu) * (1. + nu));   // d/dnu 1 / (2 (1 + nu))
    }
}

// dR/dp at the active state Xa for a global parameter p, into `out` (active,
// constrained entries 0).  Every residual term is linear in E (elasticity, the
// plate's sections), in k0, h and the emissivity (heat) and in the pressure
// (plate loads), so dR/dp is the term's residual at a unit parameter; nu
// enters through D (and the plate's G), whose nu-derivative replaces it in the
// element kernel.  The pressure (plate; hex8 sides) enters through the load only.
int global_dRdp(mastg_ctx* c, int param, const double* Xa, double* out) {
    const Structure& S = c->S;
    TRY(dalloc(c, c->Rel, size_t(S.n_elems) * S.npe * S.nav * sizeof(double)));
    const int ph = c->d.physics;
    double scale = 1.0;
    bool load_term = false;
    if (ph == MASTG_HEX8_ELASTIC) {
        Hex8Args a = hex8_args(c, Xa, false);
        if (param == MASTG_PARAM_E) {
            if (c->d.E == 0) return fail(c, MASTG_ERR_INVALID, "dR/dE needs E != 0");
            scale = 1.0 / c->d.E;
        } else if (param == MASTG_PARAM_NU) {
            material_dD3_dnu_unitE(c->d.nu, a.Dhat);
        } else if (param == MASTG_PARAM_PRESSURE && c->has_pressure) {
            load_term = true;   // dR/dp = -(load of a unit pressure on the marked sides)
        } else {
            return fail(c, MASTG_ERR_INVALID,
                        "parameter not defined for hex8 elasticity (E, nu; pressure with pressure sides)");
        }
        if (!load_term) launch_hex8_K(a, c->stream);
    } else if (ph == MASTG_QUAD4_MINDLIN) {
        Quad4Args a = quad4_args(c, Xa, false);
        const double E = c->d.E, nu = c->d.nu;
        if (param == MASTG_PARAM_E) {
            if (E == 0) return fail(c, MASTG_ERR_INVALID, "dR/dE needs E != 0");
            scale = 1.0 / E;
        } else if (param == MASTG_PARAM_NU) {
            // d/dnu of the plane-stress D and of G = E / (2 (1 + nu))
            // (isotropic_material_property_card.cpp:334-345, 431-493 with dE = 0)
            const double q = 1. - nu * nu;
            std::memset(a.Dps, 0, sizeof(a.Dps));
            a.Dps[0] = a.Dps[4] = 2. * E * nu / (q * q);
            a.Dps[1] = a.Dps[3] = E * (1. + nu * nu) / (q * q);
            a.Dps[8] = -E / (2. * (1. + nu) * (1. + nu));
            a.G = -E / (2. * (1. + nu) * (1. + nu));
        } else if (param == MASTG_PARAM_PRESSURE) {
            load_term = true;   // dR/dp = -(load of a unit pressure)
        } else {
            return fail(c, MASTG_ERR_INVALID, "parameter not defined for the Mindlin plate (E, nu, pressure)");
        }
        if (!load_term) launch_quad4_K(a, c->stream);
    } else {
        HeatArgs a = heat_args(c, Xa, true, false);
        a.k0 = 0.0;
        a.h = 0.0;
        a.rad_se = 0.0;
        a.src_scale = 0.0;
        if (param == MASTG_PARAM_K0) a.k0 = 1.0;
        else if (param == MASTG_PARAM_CONV_H) a.h = 1.0;
        else if (param == MASTG_PARAM_EMISSIVITY) {
            if (!c->has_rad) return fail(c, MASTG_ERR_INVALID, "no radiating sides");
            a.rad_se = c->d.rad_sigma;
        } else {
            return fail(c, MASTG_ERR_INVALID, "parameter not defined for heat conduction (k0, h, emissivity)");
        }
        launch_heat_K(a, c->stream);
    }
    // -(unit pressure load) for the pressure, else the elements' rows
    launch_node_residual(S.n_nodes, S.nav, S.npe, c->ninc_ptr.as<int32_t>(), c->ninc.as<int32_t>(),
                         load_term ? nullptr : c->Rel.as<double>(), load_term ? c->Fp.as<double>() : nullptr,
                         c->con.as<uint8_t>(), out, c->stream);
    if (scale != 1.0) launch_axpby(S.n_active, 0.0, out, scale, out, c->stream);
    HIPC(c, hipGetLastError());
    return MASTG_OK;
}

}  // namespace

extern "C" {

int mastg_sensitivity_rhs(mastg_ctx* c, int param, const double* X, double* dRdp, int where) {
    if (!c || !X || !dRdp) return MASTG_ERR_INVALID;
    TRY(to_active(c, X, where, c->vb.as<double>()));
    TRY(global_dRdp(c, param, c->vb.as<double>(), c->vtmp.as<double>()));
    return from_active(c, c->vtmp.as<double>(), dRdp, where);
}

int mastg_sensitivity_global(mastg_ctx* c, int param, const double* X, const double* lam, int where,
                             double* value) {
    if (!c || !X || !lam || !value) return MASTG_ERR_INVALID;
    TRY(to_active(c, X, where, c->vb.as<double>()));
    TRY(global_dRdp(c, param, c->vb.as<double>(), c->vtmp.as<double>()));
    TRY(to_active(c, lam, where, c->vz.as<double>()));
    return dot_active(c, c->vz.as<double>(), c->vtmp.as<double>(), value);
}

int mastg_output_noalloc(mastg_ctx* c, int kind, const double* X, double* value) {
    if (!c || !X || !value) return MASTG_ERR_INVALID;
    if (!c->scratch_ptr) return fail(c, MASTG_ERR_DEVICE, "no scratch address for this device");
    const int64_t n = c->S.n_active;
    const int nparts = int(std::min<int64_t>(kScratchPartials, std::max<int64_t>(1, (n + 255) / 256)));
    if (nparts > kScratchPartials)
        return fail(c, MASTG_ERR_INVALID, "reduction needs more partial sums than the scratch holds");
    if (kind == MASTG_OUTPUT_MEAN_T) {
        if (c->d.physics != MASTG_HEX8_HEAT) return fail(c, MASTG_ERR_INVALID, "mean temperature needs heat physics");
        // J = sum_i w_i T_i, w = (1/V) int N (every node; constrained ones carry T = 0)
        launch_gather_dot(n, c->wmean.as<double>(), nullptr, c->dof_of_active.as<int64_t>(), X, c->scratch_ptr,
                          nparts, value, c->stream);
    } else if (kind == MASTG_OUTPUT_COMPLIANCE) {
        // C = F . X over the free active DOFs
        launch_gather_dot(n, c->F.as<double>(), c->con.as<uint8_t>(), c->dof_of_active.as<int64_t>(), X,
                          c->scratch_ptr, nparts, value, c->stream);
    } else {
        return fail(c, MASTG_ERR_INVALID, "output not available");
    }
    HIPC(c, hipGetLastError());
    return MASTG_OK;
}

int mastg_set_ks_stress(mastg_ctx* c, double k, double sigma_ref) {
    if (!c) return MASTG_ERR_INVALID;
    if (!(k > 0) || !(sigma_ref > 0) || !std::isfinite(k) || !std::isfinite(sigma_ref))
        return fail(c, MASTG_ERR_INVALID, "KS aggregation constant and reference stress must be positive");
    c->ks_k = k;
    c->ks_sref = sigma_ref;
    return MASTG_OK;
}

int mastg_set_stress_aggregation(mastg_ctx* c, double p, double p_weight, double rho, double sigma0) {
    if (!c) return MASTG_ERR_INVALID;
    if (!(p > 0) || !(sigma0 > 0) || !std::isfinite(p) || !std::isfinite(p_weight) || !std::isfinite(rho) ||
        !std::isfinite(sigma0))
        return fail(c, MASTG_ERR_INVALID, "stress aggregation: p and the reference stress must be positive, all finite");
    c->agg_p = p;
    c->agg_pw = p_weight;
    c->agg_rho = rho;
    c->agg_s0 = sigma0;
    return MASTG_OK;
}

}  // extern "C"

namespace {

bool is_stress_output(int kind) {
    return kind == MASTG_OUTPUT_KS_STRESS || kind == MASTG_OUTPUT_PNORM_STRESS || kind == MASTG_OUTPUT_RAMP_STRESS;
}

StressAgg stress_agg(const mastg_ctx* c, int kind) {
    StressAgg g{};
    g.lim = 100.0;   // StressStrainOutputBase::_exp_arg_lim
    if (kind == MASTG_OUTPUT_KS_STRESS) {
        g.kind = kAggKS;
        g.p = c->ks_k;
        g.s0 = c->ks_sref;
    } else {
        g.kind = kind == MASTG_OUTPUT_PNORM_STRESS ? kAggPNorm : kAggRamp;
        g.p = c->agg_p;
        g.pw = c->agg_pw;
        g.rho = c->agg_rho;
        g.s0 = c->agg_s0;
    }
    return g;
}

// a stress aggregate at the active state in c->vb: I and V into c->ks_IV
// (device), value on the host
int ks_value(mastg_ctx* c, int kind, double* value) {
    if (c->d.physics != MASTG_HEX8_ELASTIC)
        return fail(c, MASTG_ERR_INVALID, "the stress outputs are defined for hex8 elasticity");
    if (!c->params_set) return fail(c, MASTG_ERR_INVALID, "design parameters not set");
    if (kind != MASTG_OUTPUT_KS_STRESS && !(c->agg_s0 > 0))
        return fail(c, MASTG_ERR_INVALID, "set the reference stress first (mastg_set_stress_aggregation)");
    const StressAgg g = stress_agg(c, kind);
    TRY(dalloc(c, c->ks_part, 2 * 1024 * sizeof(double)));
    TRY(dalloc(c, c->ks_IV, 2 * sizeof(double)));
    launch_hex8_ks_sum(hex8_args(c, c->vb.as<double>(), false), g, c->ks_part.as<double>(), c->ks_IV.as<double>(),
                       c->stream);
    HIPC(c, hipMemcpyAsync(c->h_pin, c->ks_IV.p, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    const double I = c->h_pin[0], V = c->h_pin[1];
    if (!(I > 0) || !std::isfinite(I) || !(V > 0) || !std::isfinite(V))
        return fail(c, MASTG_ERR_NAN,
                    "stress aggregate overflowed or vanished: choose the reference stress near the peak stress");
    if (g.kind == kAggKS) *value = std::log(I / V) / g.p;
    else if (g.kind == kAggPNorm) *value = g.s0 * std::pow(I / V, 1.0 / g.p);
    else *value = I - V;
    return MASTG_OK;
}

}  // namespace

extern "C" {

int mastg_csr_full_size(mastg_ctx* c, int64_t* n_rows, int64_t* nnz) {
    if (!c || !n_rows || !nnz) return MASTG_ERR_INVALID;
    if ((c->d.physics != MASTG_HEX8_ELASTIC && c->d.physics != MASTG_QUAD4_MINDLIN) || c->S.n_vars != 6)
        return fail(c, MASTG_ERR_INVALID, "the full-variable matrix is the 6-variable structural system's");
    *n_rows = c->S.n_dofs;
    *nnz = c->S.n_blocks * 36;
    return MASTG_OK;
}

int mastg_export_csr_full(mastg_ctx* c, int64_t* row_ptr, int32_t* col, double* val) {
    if (!c || !row_ptr || !col || !val) return MASTG_ERR_INVALID;
    int64_t n_rows = 0, nnz = 0;
    TRY(mastg_csr_full_size(c, &n_rows, &nnz));
    if (!c->assembled) return fail(c, MASTG_ERR_INVALID, "no assembly: call mastg_assemble first");
    const Structure& S = c->S;
    const int nv = 6, nav = S.nav;
    const bool hex = c->d.physics == MASTG_HEX8_ELASTIC;
    // rows in the caller's numbering: node r's 6 rows at base(r) + v, each
    // with 6 columns per neighbour block
    std::vector<int64_t> rp(size_t(n_rows) + 1, 0);
    for (int64_t r = 0; r < S.n_nodes; r++) {
        const int64_t deg = S.blk_ptr[r + 1] - S.blk_ptr[r];
        for (int v = 0; v < nv; v++) {
            const int64_t d = S.node_base_r[r] + v;
            if (d < 0 || d >= n_rows) return fail(c, MASTG_ERR_INVALID, "node DOFs outside the DOF range");
            rp[size_t(d) + 1] = deg * nv;
        }
    }
    for (int64_t i = 0; i < n_rows; i++) rp[size_t(i) + 1] += rp[size_t(i)];
    if (rp[size_t(n_rows)] != nnz) return fail(c, MASTG_ERR_INVALID, "the nodes' DOFs overlap");
    FullExportArgs a{};
    a.n_nodes = S.n_nodes;
    a.nav = nav;
    a.nv = nv;
    uint32_t inactive = 0;
    for (int v = 0; v < nv; v++) {
        a.act[v] = -1;
        for (int k = 0; k < nav; k++)
            if (S.active_var[k] == v) a.act[v] = int8_t(k);
        if (a.act[v] < 0) inactive |= 1u << v;
    }
    DevBuf rpf, colf, valf, diag, base, conv, mxd;
    auto release_all = [&]() {
        for (DevBuf* b : {&rpf, &colf, &valf, &diag, &base, &conv, &mxd}) {
            c->device_bytes -= int64_t(b->bytes);
            b->release();
        }
    };
    int rc = upload(c, rpf, rp);
    if (!rc) rc = upload(c, base, S.node_base_r);
    if (!rc) rc = upload(c, conv, S.constrained_var);
    if (!rc) rc = dalloc(c, colf, size_t(nnz) * sizeof(int32_t));
    if (!rc) rc = dalloc(c, valf, size_t(nnz) * sizeof(double));
    if (!rc) rc = dalloc(c, diag, size_t(S.n_nodes) * nv * sizeof(double));
    if (!rc && hex) {
        // the rotations' diagonal needs each element matrix's largest diagonal entry
        if (!c->sE.p) rc = dalloc(c, c->sE, size_t(S.n_elems) * sizeof(double));
        if (!rc) rc = dalloc(c, mxd, size_t(S.n_elems) * sizeof(double));
        if (!rc) {
            ensure_scales(c);
            launch_hex8_max_diag(S.n_elems, c->xyz.as<double>(), c->econn.as<int32_t>(), c->sE.as<double>(),
                                 c->Dhat[1], c->Dhat[21], mxd.as<double>(), c->stream);
        }
    }
    if (rc) {
        release_all();
        return rc;
    }
    launch_inactive_diag(S.n_nodes, nv, S.npe, inactive, c->ninc_ptr.as<int32_t>(), c->ninc.as<int32_t>(),
                         hex ? mxd.as<double>() : nullptr, hex ? 1.0e-20 : 1.0e-8, conv.as<uint8_t>(),
                         diag.as<double>(), c->stream);
    a.blk_ptr = c->blk_ptr.as<int64_t>();
    a.blk_col = c->blk_col.as<int32_t>();
    a.row_ptr_a = c->row_ptr.as<int64_t>();
    a.val_a = c->val.as<double>();
    a.diag = diag.as<double>();
    a.base_r = base.as<int64_t>();
    a.row_ptr_f = rpf.as<int64_t>();
    a.col_f = colf.as<int32_t>();
    a.val_f = valf.as<double>();
    launch_export_full(a, c->stream);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) {
        HipLock l(g_hip_mu);
        e = hipMemcpy(col, colf.p, size_t(nnz) * sizeof(int32_t), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(val, valf.p, size_t(nnz) * sizeof(double), hipMemcpyDeviceToHost);
    }
    release_all();
    if (e != hipSuccess) return fail(c, MASTG_ERR_DEVICE, std::string("export: ") + hipGetErrorString(e));
    std::copy(rp.begin(), rp.end(), row_ptr);
    return MASTG_OK;
}

int mastg_output_sensitivity(mastg_ctx* c, int kind, const double* X, int where, double* dJdp) {
    if (!c || !X || !dJdp) return MASTG_ERR_INVALID;
    double* gd = where == MASTG_HOST ? c->vg.as<double>() : dJdp;
    if (kind == MASTG_OUTPUT_COMPLIANCE || kind == MASTG_OUTPUT_MEAN_T) {
        // the loads and the mean-temperature weights do not depend on the
        // element parameters
        launch_fill(c->S.n_elems, 0.0, gd, c->stream);
    } else if (is_stress_output(kind)) {
        TRY(to_active(c, X, where, c->vb.as<double>()));
        double v;
        TRY(ks_value(c, kind, &v));
        TRY(dalloc(c, c->Rel, size_t(c->S.n_elems) * 24 * sizeof(double)));
        launch_hex8_ks_grad(hex8_args(c, c->vb.as<double>(), false), stress_agg(c, kind), c->ks_IV.as<double>(), gd,
                            c->stream);
    } else {
        return fail(c, MASTG_ERR_INVALID, "output not available");
    }
    HIPC(c, hipGetLastError());
    if (where == MASTG_HOST)
        HIPC(c, hipMemcpyAsync(dJdp, gd, c->S.n_elems * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return MASTG_OK;
}

int mastg_output(mastg_ctx* c, int kind, const double* X, int where, double* value, double* dv) {
    if (!c || !X || !value) return MASTG_ERR_INVALID;
    if (is_stress_output(kind)) {
        TRY(to_active(c, X, where, c->vb.as<double>()));
        TRY(ks_value(c, kind, value));
        if (dv) {
            TRY(dalloc(c, c->Rel, size_t(c->S.n_elems) * 24 * sizeof(double)));
            launch_hex8_ks_grad(hex8_args(c, c->vb.as<double>(), false), stress_agg(c, kind), c->ks_IV.as<double>(),
                                nullptr, c->stream);
            launch_node_residual(c->S.n_nodes, c->S.nav, c->S.npe, c->ninc_ptr.as<int32_t>(),
                                 c->ninc.as<int32_t>(), c->Rel.as<double>(), nullptr, c->con.as<uint8_t>(),
                                 c->vtmp.as<double>(), c->stream);
            TRY(from_active(c, c->vtmp.as<double>(), dv, where));
        }
        HIPC(c, hipGetLastError());
        return MASTG_OK;
    }
    if (kind == MASTG_OUTPUT_MEAN_T) {
        if (c->d.physics != MASTG_HEX8_HEAT) return fail(c, MASTG_ERR_INVALID, "mean temperature needs heat physics");
        TRY(to_active(c, X, where, c->vb.as<double>()));
        // J = sum_i w_i T_i with w = (1/V) int N (over all nodes; the constrained
        // ones carry T = 0), dJ/dT = w with constrained entries removed
        TRY(dot_active(c, c->wmean.as<double>(), c->vb.as<double>(), value));
        if (dv) {
            HIPC(c, hipMemcpyAsync(c->vtmp.p, c->wmean.p, c->S.n_active * sizeof(double),
                                   hipMemcpyDeviceToDevice, c->stream));
            launch_mask_constrained(c->S.n_active, c->con.as<uint8_t>(), c->vtmp.as<double>(), c->stream);
            TRY(from_active(c, c->vtmp.as<double>(), dv, where));
        }
        return MASTG_OK;
    }
    if (kind != MASTG_OUTPUT_COMPLIANCE) return fail(c, MASTG_ERR_INVALID, "output not available");
    TRY(to_active(c, X, where, c->vb.as<double>()));
    // external load vector with constrained entries removed
    launch_node_residual(c->S.n_nodes, c->S.nav, c->S.npe, c->ninc_ptr.as<int32_t>(), c->ninc.as<int32_t>(),
                         nullptr, c->F.as<double>(), c->con.as<uint8_t>(), c->vtmp.as<double>(), c->stream);
    launch_axpby(c->S.n_active, 0.0, c->vtmp.as<double>(), -1.0, c->vtmp.as<double>(), c->stream);
    TRY(dot_active(c, c->vtmp.as<double>(), c->vb.as<double>(), value));
    if (dv) TRY(from_active(c, c->vtmp.as<double>(), dv, where));
    return MASTG_OK;
}

int mastg_design_eval(mastg_ctx* c, const double* params, double* dobj, int where,
                      const mastg_solve_opts* o, mastg_design_result* out) {
    if (!c || !params) return MASTG_ERR_INVALID;
    const Structure& S = c->S;
    TRY(mastg_set_params(c, params, where));
    if (c->d.physics == MASTG_HEX8_HEAT) return heat_design_eval(c, dobj, where, o, out);
    hipEvent_t e[4];
    for (auto& x : e) HIPC(c, ev_create(&x));
    const int solver = o ? o->solver : MASTG_SOLVER_CG_JACOBI;
    TRY(check_cg_solver(c, solver));
    // hex8 multigrid on a lattice is matrix-free end to end (lattice operator,
    // row information and Galerkin products from the class matrix): only the
    // load vector is assembled, and no assembled matrix is left behind
    bool matrix_free = false;
    if (solver == MASTG_SOLVER_MG) {
        TRY(mg_build(c));   // refusals come before any assembly
        matrix_free = c->d.physics == MASTG_HEX8_ELASTIC && !c->mg_power;
    }
    HIPC(c, hipEventRecord(e[0], c->stream));
    TRY(assemble_active(c, nullptr, true, !matrix_free));   // R(0) = -F -> b = F (constrained)
    launch_axpby(S.n_active, 0.0, c->vtmp.as<double>(), -1.0, c->vtmp.as<double>(), c->stream);
    HIPC(c, hipMemcpyAsync(c->vb.p, c->vtmp.p, S.n_active * sizeof(double), hipMemcpyDeviceToDevice,
                           c->stream));
    HIPC(c, hipEventRecord(e[1], c->stream));
    mastg_solve_stats st{};
    int rc = solver == MASTG_SOLVER_MG ? mg_solve(c, c->vb.as<double>(), o, &st)
                                       : cg_solve(c, c->vb.as<double>(), o, &st, cg_op(solver));
    HIPC(c, hipEventRecord(e[2], c->stream));
    double C = 0;
    TRY(dot_active(c, c->vb.as<double>(), c->vx.as<double>(), &C));
    // self-adjoint compliance: lambda = u, dC/dp = -u . dR/dp
    double* gd = (where == MASTG_HOST || !dobj) ? c->vg.as<double>() : dobj;
    {
        TimedRegion t(c, 4);
        launch_sens(c, c->vx.as<double>(), c->vx.as<double>(), gd);
    }
    launch_axpby(S.n_elems, 0.0, gd, -1.0, gd, c->stream);
    HIPC(c, hipEventRecord(e[3], c->stream));
    if (dobj && where == MASTG_HOST)
        HIPC(c, hipMemcpyAsync(dobj, gd, S.n_elems * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    float t01 = 0, t12 = 0, t23 = 0;
    (void)hipEventElapsedTime(&t01, e[0], e[1]);
    (void)hipEventElapsedTime(&t12, e[1], e[2]);
    (void)hipEventElapsedTime(&t23, e[2], e[3]);
    for (auto& x : e) (void)ev_destroy(x);
    if (out) {
        out->objective = C;
        out->iters = st.iters;
        out->newton_iters = 1;
        out->rel_res = st.rel_res;
        out->t_assemble = t01 * 1e-3;
        out->t_solve = t12 * 1e-3;
        out->t_sens = t23 * 1e-3;
    }
    return rc;
}

int mastg_device_alloc(mastg_ctx* c, size_t bytes, void** p) {
    if (!c || !p) return MASTG_ERR_INVALID;
    HipLock l(g_hip_mu);
    HIPC(c, hipMalloc(p, bytes));
    return MASTG_OK;
}
int mastg_device_free(mastg_ctx* c, void* p) {
    if (!c) return MASTG_ERR_INVALID;
    HipLock l(g_hip_mu);
    HIPC(c, hipFree(p));
    return MASTG_OK;
}
int mastg_copy(mastg_ctx* c, void* dst, const void* src, size_t bytes, int kind) {
    if (!c) return MASTG_ERR_INVALID;
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost
                                                                          : hipMemcpyDeviceToDevice;
    HIPC(c, hipMemcpyAsync(dst, src, bytes, k, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return MASTG_OK;
}
int mastg_synchronize(mastg_ctx* c) {
    if (!c) return MASTG_ERR_INVALID;
    HIPC(c, hipStreamSynchronize(c->stream));
    return MASTG_OK;
}
int mastg_enable_kernel_timing(mastg_ctx* c, int on) {
    if (!c) return MASTG_ERR_INVALID;
    c->timing = on != 0;
    return MASTG_OK;
}
int mastg_set_option(mastg_ctx* c, int option, int value) {
    if (!c) return MASTG_ERR_INVALID;
    switch (option) {
        case MASTG_OPT_GRAPHS: c->use_graphs = value != 0; return MASTG_OK;
        case MASTG_OPT_ROW_ASSEMBLY:
            if (value < 0 || value > 2) return fail(c, MASTG_ERR_INVALID, "row assembly option: 0, 1 or 2");
            c->row_fused = value != 0;
            c->row_plan_ok = value == 1;
            return MASTG_OK;
        case MASTG_OPT_MG_STEPS_FINE:
        case MASTG_OPT_MG_STEPS_COARSE:
        case MASTG_OPT_MG_CHEB_RATIO:
            if (option == MASTG_OPT_MG_CHEB_RATIO) {
                if (value != 0 && (value < 2 || value > 10000))
                    return fail(c, MASTG_ERR_INVALID, "Chebyshev interval ratio out of range");
                c->mg_ratio = value;
            } else {
                if (value < 1 || value > kMGMaxCheb) return fail(c, MASTG_ERR_INVALID, "smoothing steps out of range");
                (option == MASTG_OPT_MG_STEPS_FINE ? c->mg_k0 : c->mg_kc) = value;
                if (option == MASTG_OPT_MG_STEPS_FINE) c->mg_k0_set = true;
            }
            // the captured V-cycles and the Chebyshev coefficients change
            drop_graphs(c);
            c->mg_valid = false;
            return MASTG_OK;
        default: return fail(c, MASTG_ERR_INVALID, "unknown option");
    }
}

int mastg_get_tuning(const mastg_ctx* c, int* spmv_lanes, int* grid) {
    if (!c) return MASTG_ERR_INVALID;
    if (spmv_lanes) *spmv_lanes = c->spmv_sw;
    if (grid) *grid = c->nparts;
    return MASTG_OK;
}

int mastg_set_tuning(mastg_ctx* c, int spmv_lanes, int grid) {
    if (!c || (spmv_lanes != 4 && spmv_lanes != 8 && spmv_lanes != 16 && spmv_lanes != 32 &&
               spmv_lanes != 64) || grid < 64 || grid > 8192)
        return MASTG_ERR_INVALID;
    c->spmv_sw = spmv_lanes;
    c->nparts = grid;
    c->tuned = true;
    drop_graphs(c);   // CG, BiCGStab and multigrid graphs all read the launch shapes
    return MASTG_OK;
}

int mastg_bench_kernel(mastg_ctx* c, int kind, int n_launch, double* avg_ms) {
    if (!c || !avg_ms || n_launch <= 0) return MASTG_ERR_INVALID;
    if (!c->params_set) return fail(c, MASTG_ERR_INVALID, "set the design parameters first");
    if ((kind == 1 || kind == 2) && !c->assembled) return fail(c, MASTG_ERR_INVALID, "assemble first");
    CGArgs a = cg_args(c);
    // kind 7: a 256 MiB device-to-device copy (read 256 MiB, write 256 MiB)
    constexpr int64_t kCopyDoubles = int64_t(1) << 25;
    DevBuf copy_src, copy_dst;
    ScatterArgs sa{c->S.n_nodes, c->S.nav, c->blk_ptr.as<int64_t>(), c->cptr.as<int32_t>(),
                   c->cidx.as<int32_t>(), c->Kel.as<double>(), c->val.as<double>()};
    // the CG scalars must not read "done" for the replayed SpMV
    HIPC(c, hipMemsetAsync(c->scal.p, 0, sizeof(CGScalars), c->stream));
    auto launch = [&]() {
        switch (kind) {
            case 0:
                if (c->d.physics == MASTG_HEX8_HEAT) launch_elements(c, c->vT.as<double>(), true);
                else launch_elements(c, nullptr, true);
                break;
            case 1: launch_scatter(sa, c->stream); break;
            case 2: launch_spmv_csr(a, a.p, a.q, 1, c->stream); break;
            case 3:
                ensure_scales(c);
                launch_mf_apply(a, a.p, a.p, a.q, 1, c->stream);
                break;
            case 4: launch_sens(c, a.x, a.x, c->vg.as<double>()); break;
            case 5: launch_hex8_K_f32(hex8_args(c, nullptr, false), c->Kel32.as<float>(), c->stream); break;
            case 6: launch_lattice_apply(a, c->mg[0]->b_lat.as<double>(), c->mg[0]->y_lat.as<double>(), c->stream); break;
            case 7: launch_copy_f64(kCopyDoubles, copy_src.as<double>(), copy_dst.as<double>(), c->stream); break;
            case 9: launch_spmv_nb64(a, a.p, a.q, 1, c->stream); break;
            case 10: (void)assemble_active(c, c->vx.as<double>(), true, true); break;
            case 8: {
                const MGLevelArgs& L0 = c->mg[0]->a;
                launch_mg_fine(a, L0, MG_STEP, 1, L0.da32, L0.db32, nullptr, c->stream);
                break;
            }
            default: break;
        }
    };
    if (kind < 0 || kind > 10) return fail(c, MASTG_ERR_INVALID, "unknown kernel kind");
    if (kind == 0 && c->d.physics != MASTG_HEX8_HEAT)   // the element kernel alone, into its own blocks
        TRY(dalloc(c, c->Kel, size_t(c->S.n_elems) * c->S.npe * c->S.npe * c->S.nav * c->S.nav * sizeof(double)));
    if (kind == 10) {
        // the whole FP64 assembly at the state in the solution vector:
        // element K and R, scatter into the CSR, inverse diagonal, node
        // residual (allocates on the first call only)
        TRY(ensure_vectors(c));
        TRY(assemble_active(c, c->vx.as<double>(), true, true));
    }
    if (kind == 9 && (c->S.nav != 5 || !c->assembled))
        return fail(c, MASTG_ERR_INVALID, "the node-block FP64 product is the plates', after an assembly");
    if (kind == 8) {
        // one Chebyshev step of the hex8 problem level's FP32 smoothing on a
        // zero state (x, res, directions) with the current assembly's bounds
        TRY(mg_build(c));
        a.S = nullptr;
        if (!lattice_args(c, a)) return fail(c, MASTG_ERR_INVALID, "no lattice operator for this mesh");
        ensure_scales(c);
        TRY(mg_setup(c));
        const MGLevelDev& L0 = *c->mg[0];
        for (const DevBuf* b : {&L0.x32, &L0.res32, &L0.da32, &L0.db32})
            HIPC(c, hipMemsetAsync(b->p, 0, b->bytes, c->stream));
    }
    if (kind == 7) {
        TRY(dalloc(c, copy_src, kCopyDoubles * sizeof(double)));
        TRY(dalloc(c, copy_dst, kCopyDoubles * sizeof(double)));
        HIPC(c, hipMemsetAsync(copy_src.p, 0, kCopyDoubles * sizeof(double), c->stream));
    }
    if (kind == 6) {
        TRY(mg_build(c));
        a.S = nullptr;
        if (!lattice_args(c, a)) return fail(c, MASTG_ERR_INVALID, "no lattice operator for this mesh");
        ensure_scales(c);
        launch_fill(c->S.n_active, 1.0, c->mg[0]->b_lat.as<double>(), c->stream);
    }
    if (kind == 5) {
        if (c->d.physics != MASTG_HEX8_ELASTIC) return fail(c, MASTG_ERR_INVALID, "FP32 element kernel is hex8 only");
        TRY(dalloc(c, c->Kel32, size_t(c->S.n_elems) * 64 * 9 * sizeof(float)));
    }
    if (kind == 1 && !c->Kel.p) return fail(c, MASTG_ERR_INVALID, "no FP64 element blocks resident");
    if (kind == 3 && (c->S.n_classes == 0 || c->d.physics != MASTG_HEX8_ELASTIC))
        return fail(c, MASTG_ERR_INVALID, "no matrix-free operator");
    launch();   // warm
    HIPC(c, hipEventRecord(c->ev0, c->stream));
    for (int i = 0; i < n_launch; i++) launch();
    HIPC(c, hipEventRecord(c->ev1, c->stream));
    HIPC(c, hipEventSynchronize(c->ev1));
    float ms = 0;
    HIPC(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *avg_ms = double(ms) / n_launch;
    // the copy kernel's buffers are released with this frame
    c->device_bytes -= int64_t(copy_src.bytes + copy_dst.bytes);
    HIPC(c, hipGetLastError());
    return MASTG_OK;
}

int mastg_kernel_time(const mastg_ctx* c, int kind, double* ms, int64_t* n) {
    if (!c || kind < 0 || kind >= 8) return MASTG_ERR_INVALID;
    if (ms) *ms = c->kt[kind].ms;
    if (n) *n = c->kt[kind].launches;
    return MASTG_OK;
}

}  // extern "C"
