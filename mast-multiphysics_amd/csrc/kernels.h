// Kernel argument blocks and launchers shared by kernels.hip and context.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mastg {

struct Hex8Args {
    int64_t n_elems;
    const double* xyz;      // [n_nodes*3] by node rank
    const int32_t* econn;   // [n_elems*8] node ranks
    const double* param;    // [n_elems] rho_e
    double E0, penalty;
    double Dhat[36];        // material matrix at E = 1
    const uint8_t* con;     // [n_active] Dirichlet mask, null = no constraints
    const double* X;        // [n_active] state or null
    double* Rel;            // [n_elems*8*3] element residual rows (when X)
    double* Kel;            // [n_elems*64*9] element blocks or null
};

// quad4 Mindlin plate: active vars u v w tx ty (nav 5, every element in a
// z = const plane) or u v w tx ty tz (nav 6, elements in their own frame)
struct Quad4Args {
    int64_t n_elems;
    int nav;                // 5 or 6
    const double* xyz;      // [n_nodes*3] by node rank
    const int32_t* econn;   // [n_elems*4]
    const double* param;    // [n_elems] thickness h_e
    double Dps[9];          // plane-stress material matrix
    double G;               // shear modulus
    double kappa, off;      // shear factor, section offset
    const uint8_t* con;     // [n_active] or null
    const double* X;        // [n_active] state or null
    double* Rel;            // [n_elems*4*nav] element residual rows (when X)
    double* Kel;            // [n_elems*16*nav*nav] element blocks or null
};

// hex8 steady conduction, k(T) = k0 (1 + alpha T)
struct HeatArgs {
    int64_t n_elems;
    const double* xyz;      // [n_nodes*3] by node rank
    const int32_t* econn;   // [n_elems*8]
    const double* param;    // [n_elems] volumetric source q_e
    double k0, alpha;
    double h, Tinf;         // convection coefficient / ambient temperature
    const uint8_t* sides;   // [n_elems] bit s = side s convects
    const uint8_t* con;     // [n_active] or null
    const double* X;        // [n_active] temperature (state), required
    double* Rel;            // [n_elems*8] element residual or null
    double* Kel;            // [n_elems*64] element Jacobian or null
    // surface radiation: flux rad_se ((T + rad_T0)^4 - (rad_Tamb + rad_T0)^4)
    // on the sides marked in rsides (null: none); rad_se = sigma * emissivity
    const uint8_t* rsides;  // [n_elems] bit s = side s radiates
    double rad_se, rad_Tamb, rad_T0;
    double src_scale;       // multiplies the source q_e (1; 0 leaves the source out)
    // congruence classes: per class kHeatGeo doubles (gradients [q][n][d],
    // then JxW [q]); null = geometry per element
    const double* geo;
    const uint8_t* eclass;  // [n_elems] class of each element (with geo)
};
constexpr int kHeatGeo = 8 * 8 * 3 + 8;

// fused hex8 assembly by CSR row owner (K, R, scatter, inverse diagonal)
struct Hex8RowArgs {
    int64_t n_nodes;
    const double* xyz;        // [n_nodes*3] by node rank
    const int32_t* econn;     // [n_elems*8]
    const int32_t* ninc_ptr;  // [n_nodes+1] node -> (element, local node) incidences
    const int32_t* ninc;      // e*8 + la
    const int64_t* blk_ptr;   // [n_nodes+1] node-block CSR
    const int32_t* blk_col;
    const int64_t* row_ptr;   // scalar CSR (node r's rows start its value chunk)
    const double* sE;         // [n_elems] element scales E0 rho_e^p
    double lam, mu;           // Lame constants of the material matrix at E = 1
    // element classes (null: every element's blocks from its own geometry):
    // K_e = sE[e] Khat[eclass[e]], Khat in node-block layout [class][a][b][3][3]
    const double* Khat;
    const uint8_t* eclass;
    int n_classes;
    const uint8_t* con;       // [n_active] or null
    const double* X;          // [n_active] state or null (no residual from the elements)
    const double* F;          // [n_active] loads or null
    double* val;              // CSR values
    double* R;                // [n_active] residual or null
    double* dinv;             // [n_active] inverse diagonal or null
    // row plan of the class kernel (null: positions searched per pass):
    // [n_nodes*64] position in node r's block row of local node b of the
    // element in incidence slot s, at r*64 + s*8 + b; -1 past the incidences
    const int8_t* row_pos;
    const uint8_t* cmask;     // [n_nodes] constraint bits of the node's 3 variables (with row_pos)
};
// rows with more neighbour blocks than this take the element + scatter path
int hex8_row_assembly_max_blocks();
void launch_hex8_row_assemble(const Hex8RowArgs& a, int max_blocks_per_row, hipStream_t s);
// the row plan (needs at most 8 incidences per node)
void launch_hex8_row_plan(const Hex8RowArgs& a, int8_t* pos, uint8_t* cmask, hipStream_t s);

// the MAST-exact full-variable matrix from the active assembly (export.hip)
struct FullExportArgs {
    int64_t n_nodes;
    int nav, nv;                     // active and caller variables per node
    int8_t act[6];                   // caller variable -> active index, -1 inactive
    const int64_t* blk_ptr;
    const int32_t* blk_col;
    const int64_t* row_ptr_a;        // active scalar CSR
    const double* val_a;
    const double* diag;              // [n_nodes*nv] inactive-variable diagonal
    const int64_t* base_r;           // [n_nodes] caller DOF of (rank, variable 0)
    const int64_t* row_ptr_f;        // [n_dofs+1] full-variable rows, caller numbering
    int32_t* col_f;
    double* val_f;
};
void launch_hex8_max_diag(int64_t n_elems, const double* xyz, const int32_t* econn, const double* sE, double lam,
                          double mu, double* out, hipStream_t s);
void launch_inactive_diag(int64_t n_nodes, int nv, int npe, uint32_t inactive, const int32_t* ninc_ptr,
                          const int32_t* ninc, const double* per_elem, double coef, const uint8_t* con_var,
                          double* diag, hipStream_t s);
void launch_export_full(const FullExportArgs& a, hipStream_t s);

struct ScatterArgs {
    int64_t n_nodes;
    int nav;
    const int64_t* blk_ptr;
    const int32_t* cptr;
    const int32_t* cidx;
    const double* Kel;
    double* val;
};

struct CGScalars {
    double rz, pq, rr, bb, alpha, beta, tol2;
    int iter, max_it, done, pad;
};

// node-centric matrix-free operator over congruence classes (hex8, nav = 3)
struct MFArgs {
    int64_t n_nodes;
    int n_classes;
    const double* Khat;       // [n_classes*64*9] class matrices at E = 1, block layout
    const double* sE;         // [n_elems] element scale E0 rho^p
    const uint8_t* eclass;    // [n_elems]
    const int32_t* econn;     // [n_elems*8]
    const int32_t* ninc_ptr;  // [n_nodes+1]
    const int32_t* ninc;      // e*8 + la, la ascending per node
    const uint8_t* nodecon;   // [n_nodes] bit v = variable v constrained
    int64_t n_elems;
    double* Fel;              // [24][n_elems] element forces (SoA), phase 1 -> phase 2
};

// matrix-free hex8 operator on a lattice mesh, evaluated per node with
// vectors in lattice order: y_n = sum over the (up to 8) cells at n of
// s_e Khat[a_n][b] x_b, constrained rows as diag = cells at the node (inputs
// must be zero on constrained DOFs)
struct LatticeOpArgs {
    int nx, ny, nz;                    // cells per dimension ((nx+1)(ny+1)(nz+1) < 2^29)
    const double* sE_cell;             // [cells] element scales in lattice-cell order
                                       // (cell i + nx (j + ny k))
    const double* Khat;                // [64*9] the (single) class matrix at E = 1
    const float* Khat32;               // its FP32 copy (the FP32 smoothing steps)
    const uint8_t* nodecon;            // [lattice points] bit v = variable v constrained
};
// partial sums the lattice operator writes (its grid)
int64_t lattice_partials(const LatticeOpArgs& l);

struct CGArgs {
    int64_t n;
    const int64_t* row_ptr;
    const int32_t* col;
    const double* val;
    const double* dinv;
    double *x, *r, *z, *p, *q;
    double *part0, *part1;
    int nparts;
    int nparts_op;            // partial sums written by the operator kernel
    int nvec;                 // grid of the CG vector kernels (and their partials)
    CGScalars* S;
    int use_mf;               // operator: 0 CSR SpMV, 1 matrix-free, 2 FP32 node-block SpMV,
                              // 3 hex8 lattice operator, 4 FP64 node-block SpMV
    int spmv_sw;              // CSR SpMV lanes per row (4, 8, 16, 32, 64)
    MFArgs mf;
    // FP32 copy of the CSR values read by node blocks (operator 2): node rank
    // r owns rows r*nav..r*nav+nav-1 and the blocks blk_ptr[r]..blk_ptr[r+1]
    const int64_t* blk_ptr;
    const int32_t* blk_col;
    const float* val32;
    int nav;
    LatticeOpArgs lat;        // operator 3: per-node lattice matrix-free product
};

// right-preconditioned BiCGStab + Jacobi on a scalar CSR (PETSc KSPBCGS role)
struct BiCGScalars {
    double rho, rho_old, alpha, omega, beta, rr, bb, tol2;
    int iter, max_it, done, pad;
};
struct BiCGArgs {
    int64_t n;
    const int64_t* row_ptr;
    const int32_t* col;
    const double* val;        // A or A^T values in the pattern of A (pattern symmetric)
    const double* dinv;       // inverse diagonal (same for A and A^T)
    double *x, *r, *r0, *p, *v, *s, *t, *ph, *sh;
    double *part0, *part1;
    int nparts, nvec, spmv_sw;
    BiCGScalars* S;
};
void launch_bicg_start(const BiCGArgs& a, const double* b, double tol2, int max_it, hipStream_t s);
void launch_bicg_iteration(const BiCGArgs& a, hipStream_t s);
// the iteration in three parts (p and ph = dinv p | v = A ph, alpha, s and sh =
// dinv s | t = A sh, x, r, check) so a preconditioner can replace ph and sh
void launch_bicg_part(const BiCGArgs& a, int part, hipStream_t s);
void launch_gather_values(int64_t n, const int32_t* idx, const double* src, double* dst, hipStream_t s);

void launch_to_f32(int64_t n, const double* src, float* dst, hipStream_t s);
// dst[0:n) = src[0:n) (n even, 16-byte aligned buffers)
void launch_copy_f64(int64_t n, const double* src, double* dst, hipStream_t s);
void launch_spmv_nb32(const CGArgs& c, const double* x, double* y, int want_dot, hipStream_t s);
void launch_spmv_nb64(const CGArgs& c, const double* x, double* y, int want_dot, hipStream_t s);
// residual replacement: r = rt (a true FP64 residual), z = dinv r, fresh r.z /
// r.r, p = z when restart; done = 1 if ||r|| meets the tolerance, else 0 (2
// past max_it)
void launch_cg_replace(const CGArgs& c, const double* rt, int restart, hipStream_t s);
void launch_hex8_K(const Hex8Args& a, hipStream_t s);
void launch_hex8_sens(const Hex8Args& a, const double* lam, double* g, hipStream_t s);
// aggregates of the von Mises stress: KS (p = the aggregation constant k),
// the weighted p-norm of StressStrainOutputBase (p, weight exponent pw, rho,
// exponent clamp lim) and the smooth ramp (p); s0 the reference stress
constexpr int kAggKS = 0, kAggPNorm = 1, kAggRamp = 2;
struct StressAgg {
    int kind;
    double p, pw, rho, s0, lim;
};
// at the state a.X: IV[0] = I, IV[1] = V, the aggregate's two sums over the
// quadrature points (part: 2 x 1024 doubles of partial sums); then the
// element rows of the state derivative into a.Rel and, when dpar is not
// null, the direct density derivative
void launch_hex8_ks_sum(const Hex8Args& a, const StressAgg& g, double* part, double* IV, hipStream_t s);
void launch_hex8_ks_grad(const Hex8Args& a, const StressAgg& g, const double* IV, double* dpar, hipStream_t s);
void launch_hex8_K_f32(const Hex8Args& a, float* Kel32, hipStream_t s);
void launch_scatter_f32(const ScatterArgs& a, const float* Kel32, hipStream_t s);
void launch_scatter(const ScatterArgs& a, hipStream_t s);
void launch_quad4_K(const Quad4Args& a, hipStream_t s);
void launch_heat_K(const HeatArgs& a, hipStream_t s);
void launch_heat_geometry(int n_classes, const double* xyz, const int32_t* econn, const int32_t* class_rep,
                          double* geo, hipStream_t s);
void launch_heat_source_dot(const HeatArgs& a, const double* lam, double scale, double* g, hipStream_t s);
void launch_quad4_sens(const Quad4Args& a, const double* lam, double* g, hipStream_t s);
void launch_node_residual(int64_t n_nodes, int nav, int npe, const int32_t* ninc_ptr,
                          const int32_t* ninc, const double* Rel, const double* F,
                          const uint8_t* con, double* R, hipStream_t s);
void launch_gather(int64_t n, const int64_t* idx, const double* src, double* dst, hipStream_t s);
void launch_scatter_vec(int64_t n, const int64_t* idx, const double* src, double* dst, hipStream_t s);
void launch_fill(int64_t n, double v, double* x, hipStream_t s);
void launch_axpby(int64_t n, double a, const double* x, double b, double* y, hipStream_t s);
void launch_jacobi_inv(int64_t n, const int64_t* diag_pos, const double* val, double* dinv,
                       hipStream_t s);
void launch_dot(int64_t n, const double* a, const double* b, double* partials, int nparts,
                double* out, hipStream_t s);
void launch_spmv_csr(const CGArgs& c, const double* x, double* y, int want_dot, hipStream_t s);
// *out = sum_i w_i x[idx[i]] over the i with skip[i] == 0 (skip may be null),
// partial sums of the nparts work-groups in `partials`, reduced in a fixed order
void launch_gather_dot(int64_t n, const double* w, const uint8_t* skip, const int64_t* idx, const double* x,
                       double* partials, int nparts, double* out, hipStream_t s);
void launch_cg_start(const CGArgs& c, const double* b, double tol2, int max_it, hipStream_t s);
void launch_cg_iteration(const CGArgs& c, hipStream_t s);
void launch_elem_scale(int64_t n, const double* rho, double E0, double penalty, double* sE,
                       hipStream_t s);
// x_elem must be zero on constrained DOFs; x_node feeds the constrained rows
void launch_mf_apply(const CGArgs& c, const double* x_elem, const double* x_node, double* y,
                     int want_dot, hipStream_t s);
void launch_mask_constrained(int64_t n, const uint8_t* con, double* x, hipStream_t s);
int mf_node_blocks(int64_t n_nodes);

// ---- geometric multigrid V-cycle (hex8 elasticity on lattice meshes) ------
constexpr int kMGMaxCheb = 8;          // Chebyshev smoothing steps at most
constexpr int kMGChebStride = 16;      // coefficients per level: 1/theta, (c1_i, c2_i)

struct MGLevelArgs {
    int64_t n_nodes;
    int nav, dim;                      // variables per node (3 hex8, 5 plate), lattice dimension
    int nx, ny, nz;                    // lattice cells per dimension (nz = 0 in 2D)
    const int32_t* rank_of_lat;        // vector index of lattice point (null: lattice order)
    int planar;                        // 1: vectors are component-major in lattice order
                                       // (DOF v of point p at v n_nodes + p; the hex8
                                       // problem level), 0: node-major (p nav + v)
    const uint8_t* con;                // [nav n_nodes] constrained DOFs
    const int64_t* blk_ptr;            // operator of levels >= 1: node-block CSR, FP64
    const int32_t* blk_col;
    const double* val;
    const float* val32;                // FP32 copy of val read by the smoother (null = FP64)
    // the smoother's values in FP16 (raw bits), each row divided by its
    // largest magnitude rs16[row] (null = FP32): the plate problem level's
    // node blocks, the heat problem level's 27-point stencil
    const uint16_t* val16;
    const float* rs16;
    const double* dinv;                // [nav n_nodes] inverse diagonal
    double *r, *x, *res, *da, *db;     // right side, iterate, residual, smoothing directions
    const double* cheb;                // this level's Chebyshev coefficients
    // the hex8 problem level (planar) smooths in FP32: iterate, residual,
    // directions and inverse diagonal (r stays the FP64 CG residual)
    float *x32, *res32, *da32, *db32;
    const float* dinv32;
    // the plate problem level smooths with block Jacobi: the inverses of its
    // 5 x 5 diagonal node blocks ([n_nodes][25], null: point Jacobi)
    const float* dinvb;
};

struct MGGalerkinArgs {                // coarse element matrices = sum_c Q_c^T K_c Q_c
    int dim, nav;                      // 3, 3 (hex8) or 2, 5 (quad4 plate)
    int64_t n_coarse;                  // coarse elements (lattice cell order)
    int cnx, cny;                      // coarse cells in x, y
    int fnx, fny;                      // fine cells in x, y
    const int32_t* f_elem_of_cell;     // null = identity
    const int32_t* f_econn;            // fine element node ranks
    const uint8_t* f_con;              // fine constraints by rank*3+v
    const int32_t* c_econn;
    const uint8_t* c_con;
    const double* Kf;                  // [n_fine][npe^2][nav^2] fine element blocks, or
    const double* Khat;                // (non-null) fine blocks sE_cell[cell] Khat: one
    const double* sE_cell;             // element class on a lattice, scales by cell
    double* Kc;                        // [n_coarse][npe^2][nav^2]
    // one-class lattice (hex8): Gc[c] = Q_c^T Khat Q_c for the 8 children, so a
    // coarse element with no constrained DOF among its 27 fine nodes (dirty[E]
    // = 0) is sum_c s_c Gc[c]; the masked product handles the listed others
    const double* Gc;                  // [8][576]
    const uint8_t* dirty;              // [n_coarse]
    const int32_t* elems;              // (non-null) the masked product runs on these
    int64_t n_elems_list;              //   n_elems_list coarse elements only
};

// smoothing modes (row-local epilogue after y = A in):
//   MG_STEP:       x += din; res -= y; dout = c1_i din + c2_i dinv res   (in = din)
//   MG_FINAL:      x += din; res -= y                                     (in = din)
//   MG_RESID_INIT: res = r - y; dout = dinv res / theta                   (in = x)
enum { MG_STEP = 1, MG_FINAL = 2, MG_RESID_INIT = 3 };

void launch_mg_galerkin(const MGGalerkinArgs& g, hipStream_t s);
// the coarse elements with dirty[E] = 0 from g.Gc and g.sE_cell (hex8)
void launch_mg_galerkin_class(const MGGalerkinArgs& g, hipStream_t s);
// dinv (levels >= 1, from the node-block values) and the Gershgorin bound
// max_i sum_j |a_ij| / a_ii into *lam; level 0 reads the scalar CSR
void launch_mg_cheb(const double* lam, double* cheb, int n_levels, int k, double ratio, hipStream_t s);
void launch_mg_dense_gather(const MGLevelArgs& L, double* A, hipStream_t s);
void launch_mg_invert(int n, double* A, hipStream_t s);
// V-cycle pieces; every kernel returns at once when *done != 0 (done may be null)
void launch_mg_pre_init(const MGLevelArgs& L, const int* done, hipStream_t s);
// the FP16 row-scaled copy of a lattice-order 27-point stencil ([27][n])
void launch_sten27_to_f16(int64_t n, const float* sten, uint16_t* sten16, float* srow, hipStream_t s);
// the FP16 row-scaled copy of a node-block matrix (val16, rs16 as above)
void launch_to_f16_rows(int64_t n_nodes, int nav, const int64_t* blk_ptr, const double* val, uint16_t* val16,
                        float* rs16, hipStream_t s);
void launch_mg_bsr(const MGLevelArgs& L, int mode, int step, const double* din, double* dout, const int* done,
                   hipStream_t s);
// a lattice-ordered hex8 coarse level as a 27-point FP32 stencil
// (sten[(o 9 + v 3 + w) n_nodes + r], o = the neighbour offset), built from its
// FP32 node blocks, and a smoothing step with it (same modes as the node blocks)
// heat levels in lattice order: 27-point FP32 stencil of a lattice-ordered
// coarse level from its rows, its smoothing step, and the rank <-> lattice
// permutations of node-major vectors
void launch_mg_to_stencil27(const MGLevelArgs& L, float* sten, hipStream_t s);
// the same stencil from precomputed row positions (kpos[p 27 + o], 255 = none)
void launch_mg_gather_stencil(const MGLevelArgs& L, const int32_t* rank_of_lat, const uint8_t* kpos, float* sten,
                              hipStream_t s);
void launch_mg_stencil_rank(const MGLevelArgs& L, const float* sten, int mode, int step, const double* din,
                            double* dout, const int* done, hipStream_t s);
void launch_perm_rank_lat(int64_t n_nodes, int nav, const int32_t* rank_of_lat, const double* src, double* dst,
                          int to_lat, const int* done, hipStream_t s);
// lambda_max(D^-1 A) of the heat problem level by power iteration on its
// lattice-ordered stencil (L: the level in lattice order, its inverse diagonal)
void launch_mg_power_heat(const MGLevelArgs& L, const float* sten, const int32_t* rank_of_lat, int n_it,
                          double safety, double* w0, double* w1, double* part, double* scal, double* lam,
                          hipStream_t s);
void launch_mg_combine_rank(int64_t n_nodes, int nav, const int32_t* rank_of_lat, const double* x0,
                            const double* dlast, double* out, const int* done, hipStream_t s);
void launch_mg_to_stencil(const MGLevelArgs& L, float* sten, hipStream_t s);
// inverses of the 5 x 5 diagonal node blocks of a plate level (FP64 values)
void launch_mg_block_dinv5(const MGLevelArgs& L, float* dinvb, hipStream_t s);
void launch_mg_gather_stencil3(const MGLevelArgs& L, const int32_t* rank_of_lat, const uint8_t* kpos, float* sten,
                               hipStream_t s);
constexpr int kMGHalfStencil = 123;   // FP32 entries per point of the symmetric half 27-point 3x3 stencil
void launch_mg_stencil(const MGLevelArgs& L, const float* sten, int mode, int step, const double* din, double* dout,
                       const int* done, hipStream_t s);
// level 0 (hex8): FP32 smoothing step with the lattice operator (c.lat) and
// the row-local epilogue on L's FP32 state
void launch_mg_fine(const CGArgs& c, const MGLevelArgs& L, int mode, int step, const float* din, float* dout,
                    const int* done, hipStream_t s);
// y = A x with the lattice operator (lattice-ordered vectors) and partial sums
// of x.y into c.part0
void launch_lattice_apply(const CGArgs& c, const double* x, double* y, hipStream_t s);
// inverse diagonal of the lattice operator (component-major lattice order,
// FP64 and FP32) and lambda_max(D^-1 A) <= max Gershgorin ratio into *lam, with
// the assembled entries formed from the class matrix and the cell scales
void launch_lattice_rowinfo(const LatticeOpArgs& l, double* dinv, float* dinv32, double* part, double* lam,
                            hipStream_t s);
// rank order <-> lattice order: gather dst[p] = src[map[p]] or scatter dst[map[p]] = src[p]
void launch_perm_nodes(int64_t n_nodes, const int32_t* map, const double* src, double* dst, int scatter,
                       hipStream_t s);
void launch_mg_rowinfo_nb(int nav, int64_t n_nodes, const int64_t* blk_ptr, const int32_t* blk_col,
                          const double* val, double* dinv, double* part, double* lam, hipStream_t s);
// estimate of lambda_max(D^-1 A) by power iteration on the FP64 node blocks of
// L (n_it steps from a fixed positive start vector), times `safety`, into
// *lam; w0, w1: work vectors of nav n_nodes; scal: 2 device doubles
void launch_mg_power(const MGLevelArgs& L, const double* val, int n_it, double safety, double* w0, double* w1,
                     double* part, double* scal, double* lam, hipStream_t s);
void launch_mg_restrict(const MGLevelArgs& F, const MGLevelArgs& C, const int* done, hipStream_t s);
// (F planar: the fine vectors are F's FP32 state)
void launch_mg_prolong(const MGLevelArgs& F, const MGLevelArgs& C, const double* dlast, const int* done,
                       hipStream_t s);
void launch_mg_coarse_solve(int n, const double* Ainv, const double* r, double* x, const int* done,
                            hipStream_t s);
// preconditioned CG around the V-cycle
void launch_pcg_update(const CGArgs& c, hipStream_t s);
// the same, also starting the pre-smoothing of the FP32 problem level L (hex8)
void launch_pcg_update_mg(const CGArgs& c, const MGLevelArgs& L, hipStream_t s);
void launch_pcg_rz(const CGArgs& c, const double* x0, const double* dlast, hipStream_t s);
// the same from the FP32 problem-level state of the hex8 V-cycle
void launch_pcg_rz32(const CGArgs& c, const float* x0, const float* dlast, hipStream_t s);
void launch_pcg_start(const CGArgs& c, const double* b, double tol2, int max_it, hipStream_t s);
void launch_pcg_start2(const CGArgs& c, hipStream_t s);
// r.r of the current r (a restart's true residual): done = 1 / 0 / 2 as it meets tol
void launch_pcg_recheck(const CGArgs& c, hipStream_t s);
void launch_pcg_beta(const CGArgs& c, hipStream_t s);
// out = x0 + dlast (a V-cycle's level-0 output), skipped when *done
void launch_mg_combine(int64_t n, const double* x0, const double* dlast, double* out, const int* done,
                       hipStream_t s);
// pieces of the CG iteration shared with the preconditioned variant
void launch_mf_elem(const CGArgs& c, const double* x, hipStream_t s);   // element phase only
void launch_cg_op_alpha(const CGArgs& c, hipStream_t s);                // q = A p, p.q, alpha
void launch_cg_p(const CGArgs& c, hipStream_t s);                       // p = z + beta p

}  // namespace mastg
