// FP64 CPU oracle for the MAST element-assembly + sensitivity path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
// calls this code; it is used by tests/, by __graft_entry__.smoke() as the
// checker, and by bench.py's cpu_baseline leg.
//
// It restates, in plain C++17 (no libMesh / PETSc / Eigen), the algorithms of
// the reference files cited beside each function:
//   - libMesh-convention structured meshes, DofMap numbering, sparsity and
//     Dirichlet constraint application (the libMesh behaviour that MAST relies
//     on; libMesh itself is not part of the reference, see the parity notes);
//   - MAST::StructuralElement3D (hex8), MAST::StructuralElement2D with
//     MAST::MindlinBendingOperator (quad4 plate), MAST::HeatConductionElementBase
//     (hex8 conduction, convection, source);
//   - MAST::NonlinearImplicitAssembly::residual_and_jacobian / sensitivity_assemble,
//     MAST::ComplianceOutput, AssemblyBase adjoint sensitivities;
//   - CG + Jacobi, BiCGStab + Jacobi and Newton, as the libMesh/PETSc solves.
//
// Parity status of each part (what pins it) is recorded in the header comment
// of each function and in tests/test_oracle_*.py.
#pragma once
#include <cstdint>
#include <vector>
#include <string>

namespace oracle {

// ---------------------------------------------------------------- mesh ----
// Structured meshes built the way libMesh MeshTools::Generation::build_cube /
// build_square number nodes, elements, element nodes and boundary sides.
// (libMesh is not under the reference; MAST's examples call build_cube /
// build_square, e.g. examples/structural/example_6/example_6.cpp.)
// Parity: unpinned against libMesh itself; pinned by hand-enumerated tiny
// meshes in tests/test_oracle_mesh.py.
struct Mesh {
    int dim = 3;                 // 2 (quad4) or 3 (hex8)
    int npe = 8;                 // nodes per element
    int nx = 0, ny = 0, nz = 0;
    int64_t n_nodes = 0, n_elems = 0;
    std::vector<double>  xyz;    // [n_nodes*3]
    std::vector<int32_t> conn;   // [n_elems*npe], libMesh local node order
    // boundary sides (elem, side, boundary id), libMesh side numbering
    std::vector<int32_t> bs_elem, bs_side, bs_id;
};

Mesh build_cube(int nx, int ny, int nz, double xmin, double xmax, double ymin,
                double ymax, double zmin, double zmax);
Mesh build_square(int nx, int ny, double xmin, double xmax, double ymin, double ymax);

// local node lists of the element sides (libMesh HEX8 / QUAD4 side_nodes_map)
extern const int HEX8_SIDE_NODES[6][4];
extern const int QUAD4_SIDE_NODES[4][2];

// -------------------------------------------------------------- dofmap ----
// libMesh DofMap::distribute_local_dofs_var_major restated for one variable
// group (all variables share the LAGRANGE FIRST FE type, as MAST's
// StructuralSystemInitialization declares 6 of them,
// src/elasticity/structural_system_initialization.cpp:37-55): node DOF bases
// are handed out in first-touch order of the element traversal and a node's
// variables are contiguous.  Element dof_indices are variable-major
// (var0 of nodes 0..npe-1, var1 ...), the column layout of
// MAST::FEMOperatorMatrix (src/numerics/fem_operator_matrix.h:291-294).
struct DofMap {
    int n_vars = 1;
    int64_t n_dofs = 0;
    std::vector<int64_t> node_base;        // [n_nodes]
    std::vector<uint8_t> constrained;      // [n_dofs] homogeneous Dirichlet
    void elem_dofs(const Mesh& m, int64_t e, int64_t* out) const;  // npe*n_vars
};
DofMap build_dofmap(const Mesh& m, int n_vars);
// Constrain variables `vars` (bit mask) on every node of every side whose
// boundary id is in `bids` (bit mask): libMesh DirichletBoundary with a zero
// function on those boundaries (src/boundary_condition/dirichlet_boundary_condition.cpp:28-145).
void add_dirichlet(const Mesh& m, DofMap& d, uint32_t bid_mask, uint32_t var_mask);

// ----------------------------------------------------------------- CSR ----
struct CSR {
    int64_t n = 0;
    std::vector<int64_t> row_ptr;
    std::vector<int32_t> col;
    std::vector<double>  val;
    int64_t nnz() const { return row_ptr.empty() ? 0 : row_ptr.back(); }
    int64_t find(int64_t r, int64_t c) const;   // position, -1 if absent
};
// libMesh SparsityPattern::Build with full variable coupling: row i holds the
// sorted union of the DOFs of every element that has DOF i.
CSR build_sparsity(const Mesh& m, const DofMap& d);
// The same pattern restricted to one row (for sampled checks of big meshes).
std::vector<int32_t> sparsity_row(const Mesh& m, const DofMap& d,
                                  const std::vector<std::vector<int32_t>>& node_elems,
                                  int64_t dof);
std::vector<std::vector<int32_t>> node_to_elems(const Mesh& m);

// ------------------------------------------------------------ materials ----
// MAST::IsotropicMaterialProperty (src/property_cards/isotropic_material_property_card.cpp)
void material_D3(double E, double nu, double D[36]);                  // :512-528
void material_dD3(double E, double nu, double dE, double dnu, double D[36]); // :533-567
void material_D2_plane_stress(double E, double nu, double D[9]);      // :395-426
void material_dD2_plane_stress(double E, double nu, double dE, double dnu, double D[9]); // :431-493
void material_G(double E, double nu, double G[4]);                    // :334-345
// MAST::Solid2DSectionProperty (src/property_cards/solid_2d_section_element_property_card.cpp)
void section_A(const double Dps[9], double h, double A[9]);           // :393-420
void section_B(const double Dps[9], double h, double off, double B[9]); // :448-476
void section_Db(const double Dps[9], double h, double off, double Db[9]); // :499-531
void section_S(const double G[4], double h, double kappa, double S[4]);   // :105-157

// ------------------------------------------------------------- elements ----
// hex8 linear elasticity, MAST::StructuralElement3D::internal_residual
// (src/elasticity/solid_element_3d.cpp:135-423) with linear strain.
// u: 48 (var-major ux,uy,uz,tx,ty,tz); f: 48; K: 48x48 row-major (may be null).
void hex8_elastic(const double X[8][3], const double D[36], const double* u,
                  double* f, double* K);
// calculate_stress (solid_element_3d.cpp:1147-1236): stress D B u at the 8
// quadrature points (xi fastest) and their JxW; u: 24 translations
void hex8_stress(const double X[8][3], const double D[36], const double* u, double sigma[8][6],
                 double jxw[8]);
// internal_residual_sensitivity (solid_element_3d.cpp:566-788): f' = sum B^T D' B u
void hex8_elastic_sens(const double X[8][3], const double dD[36], const double* u,
                       double* df);

// quad4 Mindlin plate, StructuralElement2D::internal_residual
// (src/elasticity/structural_element_2d.cpp:1075-1227, 1765-2019) with
// MindlinBendingOperator (src/elasticity/mindlin_bending_operator.cpp:57-361).
// u: 24 (var-major u,v,w,tx,ty,tz); f: 24; K: 24x24 (may be null).
struct PlateSection { double A[9], B[9], Db[9], S[4]; };
void quad4_mindlin(const double X[4][3], const PlateSection& s, const double* u,
                   double* f, double* K, bool add_thetaz_diag);
// d(internal residual)/dp with the section derivatives in `ds`
// local frame of a plate element out of the x-y plane (false: none needed)
bool plate_local_frame(const double X[4][3], double T[9], double XL[4][3]);
void plate_to_local(const double T[9], const double* g, double* l);        // 24-vectors, 6 variables
void plate_to_global(const double T[9], const double* l, double* g);
void plate_matrix_to_global(const double T[9], const double* KL, double* KG);   // 24 x 24
void quad4_mindlin_sens(const double X[4][3], const PlateSection& ds, const double* u,
                        double* df);
// uniform surface pressure p along local -z, StructuralElementBase::surface_pressure_residual
// (src/elasticity/structural_element_base.cpp:1167-1235): f -= int N^T (p n), n=(0,0,-1)
void quad4_pressure(const double X[4][3], double p, double* f);

// hex8 heat conduction, HeatConductionElementBase (src/heat_conduction/heat_conduction_elem_base.cpp)
// k(T) = k0*(1+alpha*T) evaluated at the quadrature-point temperature;
// internal residual + Jacobian incl. the dk/dT term of :124-142.
void hex8_heat(const double X[8][3], double k0, double alpha, const double* T,
               double* f, double* K);
// surface_convection_residual (:1099-1154) on side s
// uniform pressure p on side s of a hex8, StructuralElement3D::surface_pressure_residual:
// f -= int N p n dA (outward n), variable-major f
void hex8_pressure(const double X[8][3], int side, double p, double* f);
void hex8_convection(const double X[8][3], int side, double h, double Tinf,
                     const double* T, double* f, double* K);
// surface_radiation_residual (:1402-1463) on side s: flux sigma eps ((T+T0)^4 - (Ta+T0)^4)
void hex8_radiation(const double X[8][3], int side, double sigma, double emiss, double Tamb,
                    double T0, const double* T, double* f, double* K);
// volume_heat_source_residual (:1709-1745): f -= int N q ; its d/dq = -int N
void hex8_source(const double X[8][3], double q, double* f);

// ------------------------------------------------------------ problems ----
// A complete static problem on one mesh: what MAST's NonlinearImplicitAssembly
// + NonlinearSystem see for the configurations exercised here.
enum Physics { HEX8_ELASTIC = 1, QUAD4_MINDLIN = 2, HEX8_HEAT = 3 };

struct Problem {
    Physics phys = HEX8_ELASTIC;
    Mesh mesh;
    DofMap dof;
    // material
    double E = 72e9, nu = 0.33, penalty = 3.0;   // SIMP: E_e = rho^p E
    double kappa = 5.0 / 6.0, off = 0.0;         // plate section
    double k0 = 200.0, alpha = 1e-3;             // heat
    double conv_h = 50.0, conv_Tinf = 300.0;
    uint32_t conv_bid_mask = 0;
    // surface radiation on the sides whose boundary id is in rad_bid_mask
    double rad_sigma = 5.670374419e-8, rad_emiss = 0.0, rad_Tamb = 0.0, rad_T0 = 0.0;
    uint32_t rad_bid_mask = 0;
    double pressure = 0.0;                        // plate; hex8 on the pres_bid_mask sides
    uint32_t pres_bid_mask = 0;                   // hex8: boundary ids carrying the pressure
    std::vector<double> param;                    // rho_e / h_e / q_e
    std::vector<double> point_load;               // [n_dofs] nodal loads F (R -= F)
    bool active_only = false;  // structural: keep only the active variables
};

// Config builders (sizes, loads and BCs of the workloads; seeds -> SplitMix64)
Problem make_cantilever(int nx, int ny, int nz, uint64_t seed, double rho_lo,
                        double rho_hi, bool mast_exact_vars);
Problem make_plate(int nx, int ny, uint64_t seed, bool mast_exact_vars);
Problem make_heat(int n, uint64_t seed);

// SplitMix64 -> uniform doubles in [lo,hi): u = (x>>11)*2^-53
void uniform_fill(uint64_t seed, int64_t n, double lo, double hi, double* out);

// NonlinearImplicitAssembly::residual_and_jacobian
// (src/base/nonlinear_implicit_assembly.cpp:64-266): element loop, libMesh
// constrain_element_matrix_and_vector, add into CSR by per-entry row search
// (MatSetValues), point loads.  K pattern must come from build_sparsity.
void assemble(const Problem& P, const double* X, double* R, CSR* K, int n_threads);
// NonlinearImplicitAssembly::sensitivity_assemble restricted to the one element
// a per-element parameter touches (ElemParameterDependence pruning,
// :562-618), dotted with lambda: g_e = lambda . dR/dp_e  (AssemblyBase
// calculate_output_adjoint_sensitivity_multiple_parameters_no_direct, assembly_base.cpp:594-639)
void sensitivity_dots(const Problem& P, const double* X, const double* lambda,
                      double* g, int n_threads);

// ------------------------------------------------------------- solvers ----
struct SolveStats { int iters = 0; double rel_res = 0; bool converged = false; };
void spmv(const CSR& A, const double* x, double* y, int n_threads);
SolveStats cg_jacobi(const CSR& A, const double* b, double* x, double rtol,
                     int max_it, int n_threads);
SolveStats bicgstab_jacobi(const CSR& A, const double* b, double* x, double rtol,
                           int max_it, int n_threads);
CSR transpose(const CSR& A);

struct ElasticResult {
    std::vector<double> u;
    double compliance = 0;
    std::vector<double> dc;   // dC/dparam_e
    SolveStats st;
};
// Static solve + compliance + adjoint sensitivity (compliance is self-adjoint,
// lambda = u; examples/structural/example_6/example_6.cpp:877-939,
// src/elasticity/compliance_output.cpp:71-152).
// `pattern`: the mesh's sparsity (build_sparsity), or null to build it here.
ElasticResult solve_elastic(const Problem& P, double rtol, int max_it, int n_threads,
                            const CSR* pattern = nullptr);

struct HeatResult {
    std::vector<double> T;
    double mean_T = 0;
    std::vector<double> dJ;   // d(mean T)/dq_e
    int newton_iters = 0;
    std::vector<double> res_hist;
};
HeatResult solve_heat(const Problem& P, double newton_rtol, double lin_rtol,
                      int n_threads);

}  // namespace oracle
