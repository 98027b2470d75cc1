/*
 * mastg — MI355X-native element assembly, solve and sensitivity for MAST.
 *
 * C ABI of the drop-in boundary.  Plain pointers and sizes only; every entry
 * point returns an int status (MASTG_OK = 0, negative on error) and never
 * throws.  mastg_last_error() returns the message of the last failure on a
 * context (or of the last failed create when ctx is NULL).
 *
 * The reference path this replaces (MAST, files relative to the reference root):
 *   - MAST::NonlinearImplicitAssembly::residual_and_jacobian
 *       src/base/nonlinear_implicit_assembly.cpp:64-266       -> mastg_assemble
 *   - MAST::NonlinearImplicitAssembly::sensitivity_assemble +
 *     AssemblyBase::calculate_output_adjoint_sensitivity_multiple_parameters_no_direct
 *       src/base/nonlinear_implicit_assembly.cpp:515-686,
 *       src/base/assembly_base.cpp:594-639                     -> mastg_sensitivity
 *   - MAST::NonlinearSystem::solve / adjoint_solve (libMesh SNES + PETSc KSP)
 *       src/base/nonlinear_system.cpp:209-235, 798-847         -> mastg_solve
 *   - MAST::ComplianceOutput::evaluate / output_total
 *       src/elasticity/compliance_output.cpp:71-102, 239-248   -> mastg_output
 *   - the element-operation contract
 *       src/base/nonlinear_implicit_assembly_elem_operations.h:35-116,
 *       src/elasticity/structural_nonlinear_assembly.cpp:112-197,
 *       src/heat_conduction/heat_conduction_nonlinear_assembly.cpp:81-118
 *     is realised per physics by the element kernels selected with
 *     mastg_problem_desc.physics.
 *
 * Vectors passed across the boundary use the caller's DofMap numbering
 * (DOF(node, var) = node_dof_base[node] + var, as libMesh's var-major numbering
 * with one variable group gives).  The global matrix is exported over the
 * "active" variables (hex8 elasticity: ux uy uz; Mindlin plate: u v w tx ty,
 * all six when an element leaves the z = const plane; heat: T), with rows and
 * columns in the caller's order: it is the caller's
 * full-coupling sparsity pattern restricted to those variables.
 */
#ifndef MASTG_MASTG_H
#define MASTG_MASTG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status -- */
#define MASTG_OK 0
#define MASTG_ERR_INVALID (-1)        /* bad argument / unsupported combination   */
#define MASTG_ERR_OOM (-2)            /* device allocation failed                  */
#define MASTG_ERR_NOT_CONVERGED (-3)  /* iterative solve hit its iteration limit  */
#define MASTG_ERR_NAN (-4)            /* NaN / Inf detected in inputs or iterates  */
#define MASTG_ERR_DEVICE (-5)         /* HIP runtime error                         */
#define MASTG_ERR_ADMISSION (-6)      /* request exceeds the server's HBM budget   */
#define MASTG_ERR_BUSY (-7)           /* server queue full; resubmit later         */

typedef struct mastg_ctx mastg_ctx;

/* physics of the element kernels (the elem-ops class the caller would attach) */
#define MASTG_HEX8_ELASTIC 1   /* MAST::StructuralElement3D, linear strain, SIMP E_e = rho_e^p E */
#define MASTG_QUAD4_MINDLIN 2  /* MAST::StructuralElement2D + MindlinBendingOperator, param h_e */
#define MASTG_HEX8_HEAT 3      /* MAST::HeatConductionElementBase, k(T) = k0(1+alpha T), param q_e */

/* Problem description.  Arrays are read during mastg_ctx_create only; the
 * context keeps its own device-resident copies (mesh, DofMap, CSR structure,
 * element->CSR position maps).  Replaces what MAST reads from the libMesh
 * MeshBase / DofMap / boundary info and the property cards
 * (src/elasticity/structural_system_initialization.cpp:37-55,
 *  src/base/physics_discipline_base.cpp:45-121, 218-363). */
typedef struct {
    int32_t physics;             /* MASTG_* physics                                   */
    int64_t n_nodes;
    int64_t n_elems;
    const double* xyz;           /* [n_nodes*3]                                        */
    const int32_t* conn;         /* [n_elems*npe] libMesh local node order (npe 8 / 4) */
    int32_t n_vars;              /* variables per node of the caller's system: 6 (MAST's
                                    structural system; hex8 also 3, plates in one z = const
                                    plane also 5) or 1 (heat).  A plate element whose nodes
                                    do not share one z is computed in its own frame, which
                                    needs all 6 */
    const int64_t* node_dof_base;/* [n_nodes] or NULL: libMesh var-major first-touch  */
    int64_t n_constrained;       /* homogeneous Dirichlet DOFs (caller numbering)      */
    const int64_t* constrained_dofs;
    int64_t n_sides;             /* boundary sides: (elem, libMesh side, boundary id)  */
    const int32_t* side_elem;
    const int32_t* side_index;
    const int32_t* side_id;
    const double* point_loads;   /* [n_dofs] nodal loads F (R -= F) or NULL            */
    /* material / section / boundary data */
    double E, nu, penalty;       /* elasticity; SIMP penalty p (E_e = rho^p E)         */
    double kappa, offset;        /* Mindlin section shear factor, offset (0 supported) */
    double pressure;             /* Mindlin: uniform pressure along local -z; hex8: on
                                    the sides in pres_bid_mask (below)               */
    double k0, alpha;            /* conduction k(T) = k0 (1 + alpha T)                 */
    double conv_h, conv_tinf;    /* convection coefficient / ambient temperature       */
    uint32_t conv_bid_mask;      /* boundary ids carrying convection (bit mask)        */
    int32_t flags;               /* reserved, 0                                        */
    /* surface radiation (heat): flux sigma eps ((T + T0)^4 - (T_amb + T0)^4) on the
     * sides whose boundary id is in rad_bid_mask, with its Jacobian
     * 4 sigma eps (T + T0)^3 -- HeatConductionElementBase::surface_radiation_residual
     * (src/heat_conduction/heat_conduction_elem_base.cpp:1402-1463) with the
     * boundary condition's emissivity, ambient_temperature,
     * reference_zero_temperature and stefan_bolzmann_constant parameters */
    double rad_emissivity, rad_tamb, rad_tzero, rad_sigma;
    uint32_t rad_bid_mask;       /* boundary ids carrying radiation (bit mask)         */
    /* hex8 elasticity: uniform pressure on the sides whose boundary id is in
     * pres_bid_mask, R -= int N (p n) dA with n the side's outward normal --
     * StructuralElement3D::surface_pressure_residual
     * (src/elasticity/solid_element_3d.cpp:815-876); MASTG_PARAM_PRESSURE
     * is its parameter */
    uint32_t pres_bid_mask;
} mastg_problem_desc;

typedef struct {
    int64_t n_dofs;         /* caller DOFs                                */
    int64_t n_active;       /* DOFs of the active variables (matrix size) */
    int64_t nnz;            /* nonzeros of the active CSR                 */
    int64_t n_blocks;       /* node-pair blocks (nnz / nav^2)             */
    int32_t n_active_vars;  /* nav                                        */
    int32_t nodes_per_elem;
    int64_t device_bytes;   /* resident device footprint of the context   */
} mastg_sizes;

/* pointer-location flags for vectors passed to the entry points */
#define MASTG_HOST 0
#define MASTG_DEVICE 1

int mastg_ctx_create(const mastg_problem_desc* desc, mastg_ctx** out);
int mastg_ctx_destroy(mastg_ctx* ctx);
const char* mastg_last_error(const mastg_ctx* ctx);
int mastg_get_sizes(const mastg_ctx* ctx, mastg_sizes* out);

/* Sparsity of the active-variable matrix: row_ptr [n_active+1], col_idx [nnz],
 * rows/cols numbered in the caller's order restricted to active variables.
 * Replaces libMesh SparsityPattern::Build as attached by
 * MAST::NonlinearSystem::init_data (src/base/nonlinear_system.cpp:114-159). */
int mastg_get_csr_pattern(const mastg_ctx* ctx, int64_t* row_ptr, int32_t* col_idx);
/* active index of each caller DOF (-1 for inactive variables), [n_dofs] */
int mastg_get_active_map(const mastg_ctx* ctx, int64_t* active_of_dof);

/* The MAST-exact matrix of the 6-variable structural system after the last
 * FP64 assembly, as MAST's residual_and_jacobian leaves it in the PETSc
 * matrix (src/base/nonlinear_implicit_assembly.cpp:111-186): every caller DOF
 * a row, libMesh's full-coupling pattern over all 6 variables of the nodes,
 * the assembled values on the active variables, and on the others the small
 * diagonal MAST adds -- hex8: 1e-20 x each element matrix's largest diagonal
 * entry on the rotations (src/elasticity/solid_element_3d.cpp:413-417),
 * plate: 1e-8 per element on theta_z (src/elasticity/structural_element_2d.cpp:
 * 1215-1219) -- with libMesh's constraint convention on constrained DOFs.
 * mastg_csr_full_size gives the row count (n_dofs) and nnz; the export
 * writes host arrays row_ptr [n_rows+1], col_idx [nnz], val [nnz]. */
int mastg_csr_full_size(mastg_ctx* ctx, int64_t* n_rows, int64_t* nnz);
int mastg_export_csr_full(mastg_ctx* ctx, int64_t* row_ptr, int32_t* col_idx, double* val);

/* Per-element design parameter (rho_e / h_e / q_e), [n_elems]; `where` =
 * MASTG_HOST or MASTG_DEVICE.  Replaces the per-element FieldFunction
 * (examples/structural/example_6/example_6.cpp:100-137). */
int mastg_set_params(mastg_ctx* ctx, const double* params, int where);

/* element precision for the Jacobian build: FP64 everywhere, or (hex8
 * elasticity) FP32 element matrices on the matrix cores, summed into the FP64
 * global matrix; entries then carry ~1e-7 relative rounding */
#define MASTG_ELEM_FP64 0
#define MASTG_ELEM_FP32 1

/* NonlinearImplicitAssembly::residual_and_jacobian.  X, R: [n_dofs] caller
 * numbering (X NULL = zero state); J_vals: [nnz] in mastg_get_csr_pattern
 * order (NULL = residual only).  Element loop, libMesh constraint
 * application (constrained rows/cols zeroed, unit diagonal per element),
 * deterministic segmented-reduction scatter (no atomics) and point loads.
 * The assembled matrix also stays resident for mastg_solve. */
int mastg_assemble(mastg_ctx* ctx, const double* X, double* R, double* J_vals,
                   int where, int elem_precision);

/* ------------------------------------------------------------ solve ---- */
#define MASTG_SOLVER_CG_JACOBI 0        /* FP64 CSR CG + Jacobi (plain path)           */
#define MASTG_SOLVER_MIXED 1            /* FP32-SpMV PCG + FP64 reliable updates        */
#define MASTG_SOLVER_BICGSTAB_JACOBI 2  /* FP64 BiCGStab + Jacobi (non-symmetric J)     */
#define MASTG_SOLVER_MATFREE 3          /* FP64 CG + Jacobi on the matrix-free operator
                                           over congruent-element classes (the same
                                           matrix as the CSR; hex8, <= 8 element shapes) */
#define MASTG_SOLVER_MG 4               /* FP64 Krylov solver preconditioned by a geometric
                                           multigrid V-cycle: trilinear transfers, Galerkin
                                           coarse operators re-formed per parameter set,
                                           Chebyshev/Jacobi smoothing, dense coarsest solve.
                                           hex8 elasticity: CG, matrix-free lattice operator
                                           on the problem level (FP32 smoothing); hex8
                                           heat: BiCGStab on the assembled Jacobian;
                                           Mindlin plate: CG, bilinear transfers, assembled
                                           problem level.  Meshes that fill a lattice whose
                                           cell counts halve to <= 1536 coarse DOFs */

typedef struct {
    int32_t solver;       /* MASTG_SOLVER_*                                   */
    int32_t max_iters;    /* 0 -> 100000                                      */
    double rtol;          /* ||b - A x||_2 / ||b||_2 target; 0 -> 1e-10        */
    double reliable_delta;/* mixed: FP64 residual replacement trigger, 0 -> 0.1 */
    int32_t transpose;    /* solve A^T x = b (adjoint_solve)                  */
    int32_t check_every;  /* convergence polling interval in iterations, 0->32*/
} mastg_solve_opts;

typedef struct {
    int32_t iters;          /* Krylov iterations                     */
    int32_t replacements;   /* FP64 residual replacements (mixed)    */
    double rel_res;         /* final true relative residual (FP64)   */
    double seconds;         /* device time of the solve              */
} mastg_solve_stats;

/* Solve J x = rhs with the matrix of the last mastg_assemble; rhs, x in the
 * caller numbering (inactive variables of x are set to 0). */
int mastg_solve(mastg_ctx* ctx, const double* rhs, double* x, int where,
                const mastg_solve_opts* opts, mastg_solve_stats* stats);

/* J dX with the Jacobian of the last mastg_assemble, op 0 = assembled CSR,
 * op 1 = matrix-free operator (same matrix), op 2 = the FP32-rounded values
 * read by node blocks (the mixed-precision solver's operator), op 3 = the
 * per-node lattice matrix-free operator of the multigrid solver applied to dX
 * with its constrained entries zeroed (hex8 lattice meshes), op 4 = the FP64
 * values read by node blocks, one column index per block (the plates'
 * multigrid CG operator; structural Jacobians).  Replaces
 * NonlinearImplicitAssembly::linearized_jacobian_solution_product
 * (src/base/nonlinear_implicit_assembly.cpp:271-383) for the linear elements. */
int mastg_jacobian_product(mastg_ctx* ctx, const double* dX, double* JdX, int where, int op);

/* g_e = lambda_e . dR_e/dp_e for every element (constrained), the batched form of
 * sensitivity_assemble + the per-parameter dot of
 * calculate_output_adjoint_sensitivity_multiple_parameters_no_direct
 * (element parameters: rho_e, h_e, and the heat source q_e, whose dR/dq_e =
 * -int N is volume_heat_source_residual_sensitivity,
 * src/heat_conduction/heat_conduction_elem_base.cpp:1749-1791). */
int mastg_sensitivity(mastg_ctx* ctx, const double* X, const double* lambda, double* g,
                      int where);

/* Global parameters: every element's residual depends on them. */
#define MASTG_PARAM_E 1           /* Young's modulus (hex8 elasticity, plate)           */
#define MASTG_PARAM_NU 2          /* Poisson's ratio (hex8 elasticity, plate)           */
#define MASTG_PARAM_PRESSURE 3    /* plate surface pressure; hex8 side pressure         */
#define MASTG_PARAM_K0 4          /* conductivity k0 of k(T) = k0 (1 + alpha T) (heat)   */
#define MASTG_PARAM_CONV_H 5      /* convection coefficient h (heat)                    */
#define MASTG_PARAM_EMISSIVITY 6  /* radiation emissivity (heat)                        */

/* dR/dp at state X for a global parameter, [n_dofs] caller numbering
 * (constrained and inactive entries 0): NonlinearImplicitAssembly::
 * sensitivity_assemble (src/base/nonlinear_implicit_assembly.cpp:515-686) for a
 * parameter every element depends on, through the elements'
 * elem_sensitivity_calculations (e.g. StructuralElement3D::internal_residual_sensitivity,
 * src/elasticity/solid_element_3d.cpp:566-788; HeatConductionElementBase
 * surface_convection/radiation_residual_sensitivity,
 * src/heat_conduction/heat_conduction_elem_base.cpp:1160-1275, 1519-1634). */
int mastg_sensitivity_rhs(mastg_ctx* ctx, int param, const double* X, double* dRdp, int where);
/* lambda . dR/dp for a global parameter (one value): the per-parameter dot of
 * AssemblyBase::calculate_output_adjoint_sensitivity_multiple_parameters_no_direct
 * (src/base/assembly_base.cpp:594-639); the output's total derivative is
 * dJ/dp = dJ/dp|_direct - lambda . dR/dp. */
int mastg_sensitivity_global(mastg_ctx* ctx, int param, const double* X, const double* lambda,
                             int where, double* value);
/* Forward sensitivity: solve J dX/dp = -dR/dp with the Jacobian of the last
 * mastg_assemble (assemble it at X first) -- NonlinearSystem::sensitivity_solve
 * (src/base/nonlinear_system.cpp:742-794).  dXdp: [n_dofs] caller numbering. */
int mastg_sensitivity_solve(mastg_ctx* ctx, int param, const double* X, double* dXdp, int where,
                            const mastg_solve_opts* opts, mastg_solve_stats* stats);

/* outputs */
#define MASTG_OUTPUT_COMPLIANCE 1   /* C = F . X (external loads, incl. point loads) */
#define MASTG_OUTPUT_MEAN_T 2       /* (1/V) int T dV                                */
#define MASTG_OUTPUT_KS_STRESS 3    /* hex8 elasticity: (1/k) ln( sum_qp exp(k (svm - s_ref)
                                       / s_ref) JxW / sum_qp JxW ), svm the von Mises stress of
                                       E_e D B u at the quadrature points --
                                       KSStressStrainOutput (src/elasticity/ks_stress_output.cpp:53-243)
                                       over StructuralElement3D::calculate_stress
                                       (src/elasticity/solid_element_3d.cpp:1147-1236)  */
#define MASTG_OUTPUT_PNORM_STRESS 4 /* hex8 elasticity: s0 (sum_qp (svm/s0)^p w JxW / sum_qp w JxW)^(1/p),
                                       w = exp(min(rho ((svm - s0)/s0)^pw, 100)) --
                                       StressStrainOutputBase::functional_for_all_elems
                                       (src/elasticity/stress_output_base.cpp:852-908; the stress
                                       constraint of examples/structural/example_6/example_6.cpp:515-587) */
#define MASTG_OUTPUT_RAMP_STRESS 5  /* hex8 elasticity: sum_qp ((1 + (svm/s0)^p)^(1/p) - 1) JxW --
                                       SmoothRampStressStrainOutput::functional_for_all_elems
                                       (src/elasticity/smooth_ramp_stress_output.cpp:36-87) */
int mastg_output(mastg_ctx* ctx, int kind, const double* X, int where, double* value,
                 double* dvalue_dX /* [n_dofs] or NULL */);
/* KS stress output's aggregation constant k and reference stress s_ref
 * (KSStressStrainOutput's _p_norm_stress and _sigma0; defaults 10 and 1) */
int mastg_set_ks_stress(mastg_ctx* ctx, double k, double sigma_ref);
/* p-norm and smooth-ramp stress outputs' coefficients: the stress exponent p,
 * the weight exponent pw, rho and the reference stress s0 --
 * StressStrainOutputBase::set_aggregation_coefficients(p1, p2, rho, sigma0)
 * (src/elasticity/stress_output_base.h:246-252; defaults 2, 1, 1, and s0 has
 * no default: the outputs refuse until it is set).  The ramp uses p and s0. */
int mastg_set_stress_aggregation(mastg_ctx* ctx, double p, double p_weight, double rho, double sigma0);
/* Direct derivative of an output with respect to every element parameter at
 * state X, [n_elems] (the partial dJ/dp_e at fixed X; the total derivative
 * adds -lambda . dR/dp_e from mastg_sensitivity with lambda solving
 * J^T lambda = dJ/dX): zero for compliance and mean temperature; for the
 * stress outputs the rho_e^p dependence of the stress --
 * KSStressStrainOutput::functional_sensitivity_for_elem (ks_stress_output.cpp:98-136),
 * StressStrainOutputBase::functional_sensitivity_for_elem
 * (stress_output_base.cpp:973-1043), SmoothRampStressStrainOutput::
 * functional_sensitivity_for_elem (smooth_ramp_stress_output.cpp:90-130). */
int mastg_output_sensitivity(mastg_ctx* ctx, int kind, const double* X, int where, double* dJdp);

/* mastg_output for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops): no counterpart in the reference.  X
 * ([n_dofs], caller numbering) and value (one double) are DEVICE pointers; the
 * work is enqueued on the context's stream and not waited for (value is ready
 * after mastg_synchronize).  It calls no hipMalloc / hipFree and no allocating
 * library routine: the reduction's partial sums (at most 4096) live in a
 * file-scope device array of 32 KiB per GPU, held until the process ends and
 * shared by every call on that GPU, so calls that overlap on one GPU race.
 * Any vector length fits (grid-stride reduction); outputs as mastg_output. */
int mastg_output_noalloc(mastg_ctx* ctx, int kind, const double* X, double* value);

/* ------------------------------------------- one design evaluation ---- */
/* assemble + solve + output + adjoint sensitivities with everything resident
 * on the device: compliance and dC/dp for the elastic physics, mean
 * temperature and d(mean T)/dq for heat (Newton for the k(T) nonlinearity).
 * params / dobj are [n_elems]; `where` applies to both.  With MASTG_SOLVER_MG
 * on a hex8 lattice mesh the evaluation is matrix-free (lattice operator, row
 * information and Galerkin products from the element class matrix): no
 * element or global matrix is formed and none is left assembled afterwards
 * (mastg_assemble forms one). */
typedef struct {
    double objective;
    int32_t iters;          /* linear iterations (sum over Newton steps)    */
    int32_t newton_iters;
    double rel_res;
    double t_assemble, t_solve, t_sens;   /* seconds (device events)       */
} mastg_design_result;

int mastg_design_eval(mastg_ctx* ctx, const double* params, double* dobj, int where,
                      const mastg_solve_opts* opts, mastg_design_result* out);

/* ------------------------------------ batched design evaluations ---- */
/* Several designs of one mesh at once.  MAST evaluates one design per
 * FunctionEvaluation::evaluate (src/optimization/function_evaluation.h:107;
 * examples/structural/example_6/example_6.cpp:450); a batch holds n_lanes
 * (1..16) contexts over the same description, each with its own HIP stream
 * and host thread, and the designs of one mastg_batch_eval are taken in turn
 * by whichever lane is free, so their solves overlap on the device.  The
 * first design of the first call runs on lane 0 alone and every lane then
 * keeps lane 0's launch shapes: a design's result is bitwise what one
 * context alone gives, whatever its lane and whatever runs beside it.
 * params / dobj are [n_designs][n_elems] (dobj may be null), out and seconds
 * (each design's wall time on its lane) are [n_designs] or null; `where` as
 * for mastg_design_eval.  Every design is evaluated; the return value is the
 * error of the lowest-numbered failing design (its message in
 * mastg_batch_last_error).  Calls on one batch must not overlap. */
typedef struct mastg_batch mastg_batch;
int mastg_batch_create(const mastg_problem_desc* desc, int n_lanes, mastg_batch** out);
int mastg_batch_destroy(mastg_batch* b);
int mastg_batch_lanes(const mastg_batch* b);
int mastg_batch_lane(mastg_batch* b, int lane, mastg_ctx** ctx);   /* a lane's context (read-mostly use) */
int mastg_batch_set_option(mastg_batch* b, int option, int value);  /* mastg_set_option on every lane */
int mastg_batch_eval(mastg_batch* b, int n_designs, const double* params, double* dobj, int where,
                     const mastg_solve_opts* opts, mastg_design_result* out, double* seconds);
const char* mastg_batch_last_error(mastg_batch* b);

/* device vector access for callers that keep data resident */
int mastg_device_alloc(mastg_ctx* ctx, size_t bytes, void** ptr);
int mastg_device_free(mastg_ctx* ctx, void* ptr);
int mastg_copy(mastg_ctx* ctx, void* dst, const void* src, size_t bytes, int kind /*0 h2d 1 d2h 2 d2d*/);
int mastg_synchronize(mastg_ctx* ctx);

/* The HIP stream (a hipStream_t, as void*) the context enqueues its work on.
 * By default a non-blocking stream the context created; mastg_set_stream
 * moves the context onto the caller's stream (null: back to its own) after
 * the work already enqueued has finished, so that the caller can order the
 * context's kernels with its own device work (events, graphs).  Entry points
 * with host results still synchronise that stream before they return; the
 * stream must be of default priority and outlive its use by the context.  No
 * counterpart in the reference (PETSc and libMesh run on the host); for
 * callers that own the device stream. */
int mastg_get_stream(const mastg_ctx* ctx, void** stream);
int mastg_set_stream(mastg_ctx* ctx, void* stream);

/* Device time (milliseconds) of named kernels over the context's lifetime,
 * measured with HIP events on the stream the kernels run on. kind:
 * 0 element kernel, 1 scatter, 2 SpMV (FP64 CSR), 3 matrix-free operator,
 * 4 sensitivity; returns total ms and launch count. */
int mastg_kernel_time(const mastg_ctx* ctx, int kind, double* total_ms, int64_t* launches);
int mastg_enable_kernel_timing(mastg_ctx* ctx, int on);
/* Replays one kernel of the resident state n_launch times back to back on the
 * context's stream and returns the average duration from HIP events (same
 * kinds as mastg_kernel_time, plus 5 = the FP32 matrix-core element kernel,
 * hex8 only, into a buffer of its own, 6 = the per-node lattice operator, 7 =
 * a plain 256 MiB device copy, the known-traffic kernel that hardware byte
 * counters are scaled against, 8 = one FP32 multigrid smoothing step on
 * the lattice, hex8 only, and 9 = the FP64 operator over node blocks, the
 * plates' multigrid CG product).  For roofline measurement. */
int mastg_bench_kernel(mastg_ctx* ctx, int kind, int n_launch, double* avg_ms);
/* Launch shapes: CSR SpMV lanes per row and grid size (work-groups of the
 * SpMV / CG vector kernels).  Autotuned on the first solve unless set. */
int mastg_get_tuning(const mastg_ctx* ctx, int* spmv_lanes, int* grid);
/* options: MASTG_OPT_GRAPHS (1 = capture solver iterations in hipGraphs, the
 * default; 0 = launch kernels directly) */
#define MASTG_OPT_GRAPHS 1
/* multigrid smoothing: Chebyshev steps per pre-/post-smoothing on the problem
 * level (MASTG_OPT_MG_STEPS_FINE) and on the coarse levels
 * (MASTG_OPT_MG_STEPS_COARSE); 1..8, defaults: hex8 elasticity 6 and 2
 * (3 and 2 on meshes whose cells are not all congruent), plate 4 and 2,
 * heat 3 and 2 */
#define MASTG_OPT_MG_STEPS_FINE 2
#define MASTG_OPT_MG_STEPS_COARSE 3
/* multigrid smoothing interval [lam / ratio, lam] on D^-1 A (lam: the
 * Gershgorin bound, except 1.1 x a power-iteration estimate on every plate
 * level and on the coarse levels of hex8 meshes whose cells are not all
 * congruent); 2..10000, 0 = the default: hex8 elasticity 80 (30 on meshes
 * whose cells are not all congruent), plate 150, heat 5 */
#define MASTG_OPT_MG_CHEB_RATIO 4
/* hex8 elasticity FP64 assembly: 1 (default) = fused kernel by CSR row owner
 * (element blocks, residual, scatter and inverse diagonal in one pass, no
 * element matrices in memory), over a precomputed row plan when the elements
 * fall into congruence classes and no node has more than 8; 2 = the fused
 * kernel without the plan (neighbour positions searched per node); 0 =
 * element kernel + segmented scatter (also used whenever the element matrices
 * are needed: multigrid on an assembled problem level, the FP32 element
 * path).  All three give the same matrix. */
#define MASTG_OPT_ROW_ASSEMBLY 5
int mastg_set_option(mastg_ctx* ctx, int option, int value);
int mastg_set_tuning(mastg_ctx* ctx, int spmv_lanes, int grid);

/* ------------------------------------------ element-wise design space --- */
/* The steps of a SIMP topology-optimisation loop after the design evaluation,
 * on the device:
 *   - density filter over the element centroids (the points carrying the
 *     element-wise design variables): rho~_e = sum_{|c_e - c_f| <= r}
 *     (r - d_ef) rho_f / sum (r - d_ef) -- MAST::FilterBase::
 *     compute_filtered_values with the weights of FilterBase::_init / _init2
 *     (src/level_set/filter_base.cpp:35-142, 244-420), on the element field;
 *     transpose = 1 applies its transpose (the chain rule of a sensitivity);
 *   - material volume V = sum_e V_e rho_e and dV/drho_e = V_e (element volume,
 *     or area for a plate's thickness variable) -- example_6's _evaluate_volume
 *     (examples/structural/example_6/example_6.cpp:714-791) with element densities;
 *   - one optimality-criteria update: rho_new = clamp(rho (B_e / lambda)^(1/2),
 *     max(rho_min, rho - move), min(1, rho + move)), B_e = -dC/drho_e / dV/drho_e,
 *     lambda such that V(rho_new) = volfrac * V_total (MAST drives external
 *     optimisers through the src/optimization classes; this is the standard SIMP step).
 * Built from a problem description's mesh (xyz, conn, physics); vectors are
 * [n_elems] on the host or device (`where`). */
typedef struct mastg_design mastg_design;
int mastg_design_create(const mastg_problem_desc* desc, double filter_radius, mastg_design** out);
int mastg_design_destroy(mastg_design* d);
const char* mastg_design_last_error(const mastg_design* d);
int mastg_design_info(const mastg_design* d, int64_t* n_elems, int64_t* filter_nnz, double* v_total);
int mastg_design_filter(mastg_design* d, const double* in, double* out, int transpose, int where);
int mastg_design_volume(mastg_design* d, const double* rho, int where, double* volume,
                        double* dvol /* [n_elems] or NULL */);
int mastg_design_oc_update(mastg_design* d, const double* rho, const double* dc, const double* dv,
                           double volfrac, double move, double rho_min, double* rho_new, int where,
                           double* lambda /* or NULL */);

/* ---------------------------------------------------- request server --- */
/* Continuous batching of design evaluations on one GPU (no counterpart in the
 * reference, whose optimizer drives one evaluation at a time through
 * FunctionEvaluation::evaluate, e.g. examples/structural/example_6/example_6.cpp:450-655).
 * Requests name a structured workload (kind 1 hex8 cantilever nx x ny x nz,
 * kind 2 Mindlin plate nx x ny) and carry their design parameters; n_slots
 * requests run concurrently on separate HIP streams; per-mesh structure is
 * cached and reused; admission is by estimated device footprint against
 * hbm_budget (idle contexts of other meshes are evicted LRU first). */
typedef struct mastg_server mastg_server;
typedef struct {
    int32_t n_slots;          /* concurrent requests on the device                 */
    int32_t solver;           /* MASTG_SOLVER_* for hex8 requests; plates take
                                 MASTG_SOLVER_MG when it is chosen, else CG +
                                 Jacobi (they refuse the hex8-only solvers)       */
    int64_t hbm_budget;       /* device bytes the server may hold; 0 = 64 GiB      */
    int32_t queue_capacity;   /* pending requests before submit returns BUSY       */
    double rtol;              /* hex8 solve tolerance, 0 -> 1e-10                  */
    double rtol_plate;        /* plate solve tolerance, 0 -> 1e-7 (thin plates put
                                 ||b - K u||/||b|| near 1e-8 in FP64 at 384^2)     */
} mastg_server_opts;
typedef struct {
    int32_t kind, nx, ny, nz;
    const double* params;     /* [n_elems], copied at submit                       */
    double* grad_out;         /* [n_elems] written when done; caller keeps it alive */
} mastg_request;
typedef struct {
    int32_t status;           /* 0 queued, 1 running, 2 done, < 0 MASTG_ERR_*      */
    int32_t iters;
    double objective;
    double t_queue, t_run;    /* seconds waiting / running                          */
} mastg_request_result;
int mastg_server_create(const mastg_server_opts* opts, mastg_server** out);
int mastg_server_submit(mastg_server* s, const mastg_request* req, int64_t* id);
/* non-blocking; a done / failed result is returned once and then retired */
int mastg_server_poll(mastg_server* s, int64_t id, mastg_request_result* res);
int mastg_server_wait(mastg_server* s, int64_t id, mastg_request_result* res);
int mastg_server_stats(mastg_server* s, int64_t* bytes_in_use, int32_t* n_contexts, int32_t* queued);
int mastg_server_destroy(mastg_server* s);   /* drains the queue first */
/* message of the most recent failed request ("" if none) */
const char* mastg_server_last_error(mastg_server* s);
int mastg_request_footprint(const mastg_request* req, int64_t* bytes);

/* ---------------------------------------------------- Nastran models --- */
/* Reads a Nastran bulk-data file (.bdf) into a problem description:
 * MAST::NastranIO::read (src/mesh/nastran_io.cpp:217-255; nodes :120-143,
 * elements :146-206, SPC sets :209-223), which parses through pyNastran.
 * Nodes are ordered by Nastran id and elements by EID (the libMesh ids the
 * reference gives them); DOF bases are handed out first-touch over the
 * elements, 6 structural variables per node (MAST's
 * StructuralSystemInitialization).  Cards: GRID (basic system), CHEXA (8
 * nodes) or CQUAD4 (one element type per model), SPC1 / SPC (homogeneous:
 * every set is applied), FORCE (basic frame: nodal loads), MAT1 (E, nu),
 * PSHELL (plate thickness -> the element parameter h_e), PSOLID; small-,
 * large- and free-field formats.  The element parameter of a hex8 model is
 * 1 (a solid density).  The description points into the returned object,
 * which owns its arrays until mastg_nastran_destroy. */
typedef struct mastg_mesh_file mastg_mesh_file;
int mastg_nastran_read(const char* path, mastg_mesh_file** out);
const char* mastg_nastran_last_error(void);   /* of the calling thread's last failed read */
const mastg_problem_desc* mastg_nastran_desc(const mastg_mesh_file* m);
const double* mastg_nastran_params(const mastg_mesh_file* m);   /* [n_elems] */
/* Nastran ids of the nodes [n_nodes] and of the elements and their property
 * ids [n_elems], in the description's order (any pointer may be NULL) */
int mastg_nastran_ids(const mastg_mesh_file* m, int64_t* node_ids, int64_t* elem_ids, int32_t* elem_pids);
int mastg_nastran_spc_sets(const mastg_mesh_file* m);
int mastg_nastran_destroy(mastg_mesh_file* m);

/* ------------------------------------------------ benchmark workloads -- */
/* Structured problems of the benchmark configurations, built the way libMesh
 * MeshTools::Generation::build_cube / build_square number meshes (what MAST's
 * examples use, e.g. examples/structural/example_6/example_6.cpp), with the
 * configurations' boundary conditions, loads and SplitMix64-seeded parameters:
 *   kind 1: hex8 cantilever 1.0 x 0.5 x 0.5 m, nx x ny x nz, x=0 clamped,
 *           traction -1e6 Pa (z) on x=1 as consistent nodal loads, rho ~ U[lo,hi)
 *   kind 2: Mindlin plate 1 x 1 m, nx x ny quad4, edges u=v=w=0, p = 1e3 Pa,
 *           h ~ U[0.002, 0.01)
 *   kind 3: heat unit cube nx^3, T=0 on z=0, convection on the other faces,
 *           q ~ U[0, 1e6)
 * mast_exact_vars selects the MAST variable count (6 structural) over the
 * active one.  The workload owns the arrays its description points to. */
typedef struct mastg_workload mastg_workload;
int mastg_workload_create(int kind, int nx, int ny, int nz, uint64_t seed, double lo, double hi,
                          int mast_exact_vars, mastg_workload** out);
const mastg_problem_desc* mastg_workload_desc(const mastg_workload* w);
const double* mastg_workload_params(const mastg_workload* w);
int mastg_workload_destroy(mastg_workload* w);

#ifdef __cplusplus
}
#endif
#endif /* MASTG_MASTG_H */
