// C++ mirror of MAST's assembly interfaces for the hot path, implemented on
// the mastg C ABI (include/mastg/mastg.h).  Class and method names, argument
// meaning and error behaviour follow the reference:
//   MAST::AssemblyElemOperations / NonlinearImplicitAssemblyElemOperations
//     (src/base/assembly_elem_operation.h:42-203,
//      src/base/nonlinear_implicit_assembly_elem_operations.h:35-116)
//   MAST::StructuralNonlinearAssemblyElemOperations
//     (src/elasticity/structural_nonlinear_assembly.h:34-154)
//   MAST::HeatConductionNonlinearAssemblyElemOperations
//     (src/heat_conduction/heat_conduction_nonlinear_assembly.h)
//   MAST::NonlinearImplicitAssembly::residual_and_jacobian / sensitivity_assemble /
//     linearized_jacobian_solution_product (src/base/nonlinear_implicit_assembly.h:36-166)
//   MAST::AssemblyBase::calculate_output / calculate_output_adjoint_sensitivity_...
//     (src/base/assembly_base.h:51-370)
//   MAST::NonlinearSystem::solve / adjoint_solve / sensitivity_solve
//     (src/base/nonlinear_system.h:133-160)
//   MAST::ComplianceOutput (src/elasticity/compliance_output.h)
//   MAST::KSStressStrainOutput (src/elasticity/ks_stress_output.h)
//   MAST::NastranIO::read (src/mesh/nastran_io.h)
//   MAST::FilterBase (src/level_set/filter_base.h) and example_6's volume
// Element work is batched on the GPU: the per-element virtual calls of the
// reference become one device pass over all elements; vectors and the Jacobian
// live in the caller's DofMap numbering (SerialVector / CsrMatrix below stand
// in for libMesh NumericVector / SparseMatrix).  Failures throw
// std::runtime_error, as libmesh_error does in the reference.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "mastg/mastg.h"

namespace MAST {
namespace GPU {

using RealVector = std::vector<double>;

inline void check(int rc, const mastg_ctx* c, const char* what) {
    if (rc != MASTG_OK)
        throw std::runtime_error(std::string(what) + ": " + mastg_last_error(c) + " (status " +
                                 std::to_string(rc) + ")");
}

// global matrix over the active variables, caller numbering
struct CsrMatrix {
    std::vector<int64_t> row_ptr;
    std::vector<int32_t> col;
    std::vector<double> val;
    int64_t n() const { return int64_t(row_ptr.size()) - 1; }
};

// Description of the system: what MAST's SystemInitialization / DofMap /
// PhysicsDisciplineBase provide to the assembly
class SystemDescription {
public:
    explicit SystemDescription(const mastg_problem_desc& d) : _d(d) {}
    const mastg_problem_desc& desc() const { return _d; }
private:
    mastg_problem_desc _d;
};

// ---- element operations --------------------------------------------------
// The physics chosen for the element kernels; each subclass corresponds to the
// reference elem-ops class of the same name.  init / set_elem_solution /
// elem_calculations are performed for all elements at once by the assembly.
class AssemblyElemOperations {
public:
    virtual ~AssemblyElemOperations() = default;
    virtual int32_t physics() const = 0;
};

class NonlinearImplicitAssemblyElemOperations : public AssemblyElemOperations {};

class StructuralNonlinearAssemblyElemOperations : public NonlinearImplicitAssemblyElemOperations {
public:
    explicit StructuralNonlinearAssemblyElemOperations(bool plate = false) : _plate(plate) {}
    int32_t physics() const override { return _plate ? MASTG_QUAD4_MINDLIN : MASTG_HEX8_ELASTIC; }
private:
    bool _plate;
};

class HeatConductionNonlinearAssemblyElemOperations : public NonlinearImplicitAssemblyElemOperations {
public:
    int32_t physics() const override { return MASTG_HEX8_HEAT; }
};

// ---- assembly -------------------------------------------------------------
class NonlinearImplicitAssembly {
public:
    NonlinearImplicitAssembly() = default;
    NonlinearImplicitAssembly(const NonlinearImplicitAssembly&) = delete;
    ~NonlinearImplicitAssembly() { clear_discipline_and_system(); }

    // set_discipline_and_system + set_elem_operation_object (assembly_base.h)
    void set_discipline_and_system(const SystemDescription& sys) {
        clear_discipline_and_system();
        mastg_ctx* c = nullptr;
        check(mastg_ctx_create(&sys.desc(), &c), nullptr, "mastg_ctx_create");
        _ctx = c;
        check(mastg_get_sizes(_ctx, &_sizes), _ctx, "mastg_get_sizes");
        _n_elems = sys.desc().n_elems;
    }
    void set_elem_operation_object(NonlinearImplicitAssemblyElemOperations& ops) {
        if (!_ctx) throw std::runtime_error("set_discipline_and_system first");
        _ops = &ops;
    }
    void clear_discipline_and_system() {
        if (_ctx) mastg_ctx_destroy(_ctx);
        _ctx = nullptr;
        _ops = nullptr;
    }
    mastg_ctx* context() const { return _ctx; }
    const mastg_sizes& sizes() const { return _sizes; }

    // the per-element design parameters (rho_e / h_e / q_e) the element
    // FieldFunctions would return
    void set_element_parameters(const RealVector& p) {
        ready();
        check(mastg_set_params(_ctx, p.data(), MASTG_HOST), _ctx, "mastg_set_params");
    }

    // the Jacobian's sparsity, what SparsityPattern::Build gives libMesh
    void init_matrix(CsrMatrix& J) const {
        ready();
        J.row_ptr.resize(_sizes.n_active + 1);
        J.col.resize(_sizes.nnz);
        J.val.assign(_sizes.nnz, 0.0);
        check(mastg_get_csr_pattern(_ctx, J.row_ptr.data(), J.col.data()), _ctx, "mastg_get_csr_pattern");
    }

    // residual_and_jacobian(X, R, J, S): R and/or J may be null
    void residual_and_jacobian(const RealVector& X, RealVector* R, CsrMatrix* J) {
        ready();
        if (R) R->assign(_sizes.n_dofs, 0.0);
        if (J && J->val.size() != size_t(_sizes.nnz)) init_matrix(*J);
        check(mastg_assemble(_ctx, X.empty() ? nullptr : X.data(), R ? R->data() : nullptr,
                             J ? J->val.data() : nullptr, MASTG_HOST, MASTG_ELEM_FP64),
              _ctx, "residual_and_jacobian");
    }

    // linearized_jacobian_solution_product: J dX with the last assembled J
    void linearized_jacobian_solution_product(const RealVector& dX, RealVector& JdX) {
        ready();
        JdX.assign(_sizes.n_dofs, 0.0);
        check(mastg_jacobian_product(_ctx, dX.data(), JdX.data(), MASTG_HOST, 0), _ctx,
              "linearized_jacobian_solution_product");
    }

    // calculate_output_adjoint_sensitivity_multiple_parameters_no_direct for the
    // element parameters: sens[e] = adj . dR/dp_e for every element at once
    void calculate_output_adjoint_sensitivity_multiple_parameters_no_direct(
        const RealVector& X, const RealVector& adj, RealVector& sens) {
        ready();
        sens.assign(size_t(n_elems()), 0.0);
        check(mastg_sensitivity(_ctx, X.data(), adj.data(), sens.data(), MASTG_HOST), _ctx,
              "calculate_output_adjoint_sensitivity");
    }

    // sensitivity_assemble(p, rhs) for a parameter every element depends on
    // (MASTG_PARAM_*): rhs = dR/dp at X
    void sensitivity_assemble(int param, const RealVector& X, RealVector& dRdp) {
        ready();
        dRdp.assign(_sizes.n_dofs, 0.0);
        check(mastg_sensitivity_rhs(_ctx, param, X.data(), dRdp.data(), MASTG_HOST), _ctx, "sensitivity_assemble");
    }

    // calculate_output_adjoint_sensitivity for one global parameter: adj . dR/dp
    double adjoint_sensitivity(int param, const RealVector& X, const RealVector& adj) {
        ready();
        double v = 0;
        check(mastg_sensitivity_global(_ctx, param, X.data(), adj.data(), MASTG_HOST, &v), _ctx,
              "calculate_output_adjoint_sensitivity");
        return v;
    }

    int64_t n_elems() const { return _n_elems; }

private:
    void ready() const {
        if (!_ctx || !_ops) throw std::runtime_error("assembly has no system / elem operations attached");
    }
    mastg_ctx* _ctx = nullptr;
    NonlinearImplicitAssemblyElemOperations* _ops = nullptr;
    mastg_sizes _sizes{};
    int64_t _n_elems = 0;
};

// ---- outputs --------------------------------------------------------------
class ComplianceOutput {
public:
    // evaluate + output_total: C = F . X (external loads incl. point loads)
    double evaluate(NonlinearImplicitAssembly& a, const RealVector& X) {
        double v = 0;
        check(mastg_output(a.context(), MASTG_OUTPUT_COMPLIANCE, X.data(), MASTG_HOST, &v, nullptr),
              a.context(), "ComplianceOutput::evaluate");
        return v;
    }
    // output_derivative_for_elem, assembled: dC/dX = F (constrained entries 0)
    void output_derivative(NonlinearImplicitAssembly& a, const RealVector& X, RealVector& dq_dX) {
        dq_dX.assign(a.sizes().n_dofs, 0.0);
        double v = 0;
        check(mastg_output(a.context(), MASTG_OUTPUT_COMPLIANCE, X.data(), MASTG_HOST, &v, dq_dX.data()),
              a.context(), "ComplianceOutput::output_derivative");
    }
};

// KS aggregate of the von Mises stress (hex8): output_total is the value,
// output_derivative dKS/dX, output_sensitivity the direct dKS/drho_e
class KSStressStrainOutput {
public:
    KSStressStrainOutput(double p_norm = 10.0, double sigma0 = 1.0) : _p(p_norm), _s0(sigma0) {}
    void set_participating_elements_to_all() {}
    double evaluate(NonlinearImplicitAssembly& a, const RealVector& X) {
        check(mastg_set_ks_stress(a.context(), _p, _s0), a.context(), "KSStressStrainOutput");
        double v = 0;
        check(mastg_output(a.context(), MASTG_OUTPUT_KS_STRESS, X.data(), MASTG_HOST, &v, nullptr), a.context(),
              "KSStressStrainOutput::evaluate");
        return v;
    }
    void output_derivative(NonlinearImplicitAssembly& a, const RealVector& X, RealVector& dq_dX) {
        check(mastg_set_ks_stress(a.context(), _p, _s0), a.context(), "KSStressStrainOutput");
        dq_dX.assign(a.sizes().n_dofs, 0.0);
        double v = 0;
        check(mastg_output(a.context(), MASTG_OUTPUT_KS_STRESS, X.data(), MASTG_HOST, &v, dq_dX.data()),
              a.context(), "KSStressStrainOutput::output_derivative");
    }
    void output_sensitivity(NonlinearImplicitAssembly& a, const RealVector& X, RealVector& dq_dp) {
        check(mastg_set_ks_stress(a.context(), _p, _s0), a.context(), "KSStressStrainOutput");
        dq_dp.assign(size_t(a.n_elems()), 0.0);
        check(mastg_output_sensitivity(a.context(), MASTG_OUTPUT_KS_STRESS, X.data(), MASTG_HOST, dq_dp.data()),
              a.context(), "KSStressStrainOutput::output_sensitivity");
    }
private:
    double _p, _s0;
};

// The weighted p-norm of the von Mises stress (StressStrainOutputBase, the
// stress constraint of example_6) and the smooth ramp
// (SmoothRampStressStrainOutput): set_aggregation_coefficients(p1, p2, rho,
// sigma0) as in MAST; output_total, output_derivative (d/dX) and
// output_sensitivity (direct d/drho_e) as for the KS output
template <int KIND>
class AggregatedStressOutput {
public:
    void set_participating_elements_to_all() {}
    void set_aggregation_coefficients(double p1, double p2, double rho, double sigma0) {
        _p = p1;
        _pw = p2;
        _rho = rho;
        _s0 = sigma0;
    }
    double evaluate(NonlinearImplicitAssembly& a, const RealVector& X) {
        configure(a);
        double v = 0;
        check(mastg_output(a.context(), KIND, X.data(), MASTG_HOST, &v, nullptr), a.context(), "stress output");
        return v;
    }
    void output_derivative(NonlinearImplicitAssembly& a, const RealVector& X, RealVector& dq_dX) {
        configure(a);
        dq_dX.assign(a.sizes().n_dofs, 0.0);
        double v = 0;
        check(mastg_output(a.context(), KIND, X.data(), MASTG_HOST, &v, dq_dX.data()), a.context(),
              "stress output derivative");
    }
    void output_sensitivity(NonlinearImplicitAssembly& a, const RealVector& X, RealVector& dq_dp) {
        configure(a);
        dq_dp.assign(size_t(a.n_elems()), 0.0);
        check(mastg_output_sensitivity(a.context(), KIND, X.data(), MASTG_HOST, dq_dp.data()), a.context(),
              "stress output sensitivity");
    }
private:
    void configure(NonlinearImplicitAssembly& a) {
        check(mastg_set_stress_aggregation(a.context(), _p, _pw, _rho, _s0), a.context(),
              "set_aggregation_coefficients");
    }
    double _p = 2.0, _pw = 1.0, _rho = 1.0, _s0 = 0.0;   // StressStrainOutputBase's defaults
};
using StressStrainOutputBase = AggregatedStressOutput<MASTG_OUTPUT_PNORM_STRESS>;
using SmoothRampStressStrainOutput = AggregatedStressOutput<MASTG_OUTPUT_RAMP_STRESS>;

// ---- mesh input -------------------------------------------------------------
// NastranIO::read: the model's mesh, DofMap, SPCs, loads and properties as a
// SystemDescription (owned by this object) and the element parameters
class NastranIO {
public:
    NastranIO() = default;
    NastranIO(const NastranIO&) = delete;
    ~NastranIO() { if (_m) mastg_nastran_destroy(_m); }
    void read(const std::string& filename) {
        if (_m) mastg_nastran_destroy(_m);
        _m = nullptr;
        if (mastg_nastran_read(filename.c_str(), &_m) != MASTG_OK)
            throw std::runtime_error(std::string("NastranIO::read: ") + mastg_nastran_last_error());
    }
    SystemDescription system() const { return SystemDescription(*mastg_nastran_desc(_m)); }
    RealVector element_parameters() const {
        const mastg_problem_desc* d = mastg_nastran_desc(_m);
        const double* p = mastg_nastran_params(_m);
        return RealVector(p, p + d->n_elems);
    }
private:
    mastg_mesh_file* _m = nullptr;
};

// ---- design space ---------------------------------------------------------
// FilterBase::compute_filtered_values over element centroids, and the volume
// of example_6 (element densities)
class FilterBase {
public:
    FilterBase(const SystemDescription& sys, double radius) {
        if (mastg_design_create(&sys.desc(), radius, &_d) != MASTG_OK)
            throw std::runtime_error("FilterBase: mastg_design_create failed");
    }
    FilterBase(const FilterBase&) = delete;
    ~FilterBase() { mastg_design_destroy(_d); }
    void compute_filtered_values(const RealVector& input, RealVector& output, bool transpose = false) const {
        output.assign(input.size(), 0.0);
        if (mastg_design_filter(_d, input.data(), output.data(), transpose ? 1 : 0, MASTG_HOST) != MASTG_OK)
            throw std::runtime_error(std::string("FilterBase: ") + mastg_design_last_error(_d));
    }
    double volume(const RealVector& rho, RealVector* grad) const {
        double v = 0;
        if (grad) grad->assign(rho.size(), 0.0);
        if (mastg_design_volume(_d, rho.data(), MASTG_HOST, &v, grad ? grad->data() : nullptr) != MASTG_OK)
            throw std::runtime_error(std::string("FilterBase volume: ") + mastg_design_last_error(_d));
        return v;
    }
    mastg_design* handle() const { return _d; }
private:
    mastg_design* _d = nullptr;
};

// ---- system ---------------------------------------------------------------
class NonlinearSystem {
public:
    explicit NonlinearSystem(NonlinearImplicitAssembly& a) : _a(a) {}

    // solve(elem_ops, assembly): one Newton step for the linear elements from
    // X = 0 (SNES + KSP CG/Jacobi in the reference); solution in `solution`
    void solve(RealVector& solution, const mastg_solve_opts* opts = nullptr) {
        RealVector R;
        _a.residual_and_jacobian(RealVector(), &R, nullptr);
        for (auto& r : R) r = -r;
        solution.assign(R.size(), 0.0);
        mastg_solve_stats st{};
        check(mastg_solve(_a.context(), R.data(), solution.data(), MASTG_HOST, opts, &st), _a.context(),
              "NonlinearSystem::solve");
        last = st;
    }

    // adjoint_solve: J^T lambda = -dq/dX (rhs = -dq/dX as MAST forms it)
    void adjoint_solve(const RealVector& dq_dX, RealVector& adj, const mastg_solve_opts* opts = nullptr) {
        RealVector rhs(dq_dX);
        for (auto& r : rhs) r = -r;
        mastg_solve_opts o = opts ? *opts : mastg_solve_opts{};
        o.transpose = 1;
        adj.assign(rhs.size(), 0.0);
        mastg_solve_stats st{};
        check(mastg_solve(_a.context(), rhs.data(), adj.data(), MASTG_HOST, &o, &st), _a.context(),
              "NonlinearSystem::adjoint_solve");
        last = st;
    }

    // sensitivity_solve(X, p): J dX/dp = -dR/dp for a global parameter with the
    // Jacobian assembled at X (element parameters take the batched adjoint
    // route of the assembly, every element's sensitivity in one pass)
    void sensitivity_solve(const RealVector& X, int param, RealVector& dXdp, const mastg_solve_opts* opts = nullptr) {
        dXdp.assign(X.size(), 0.0);
        mastg_solve_stats st{};
        check(mastg_sensitivity_solve(_a.context(), param, X.data(), dXdp.data(), MASTG_HOST, opts, &st),
              _a.context(), "NonlinearSystem::sensitivity_solve");
        last = st;
    }
    mastg_solve_stats last{};

private:
    NonlinearImplicitAssembly& _a;
};

// ---- several design evaluations at once ---------------------------------
// MAST::FunctionEvaluation::evaluate (src/optimization/function_evaluation.h:107;
// the SIMP driver's evaluate, examples/structural/example_6/example_6.cpp:450)
// for several design points: each design's objective (compliance, or mean
// temperature for heat) and, on request, its gradient, evaluated side by side
// on n_lanes contexts (mastg_batch_*); each result equals one lone evaluation
// bitwise
class FunctionEvaluationBatch {
public:
    FunctionEvaluationBatch(const SystemDescription& sys, int n_lanes) : _n_elems(sys.desc().n_elems) {
        const int rc = mastg_batch_create(&sys.desc(), n_lanes, &_b);
        if (rc != MASTG_OK)
            throw std::runtime_error(std::string("mastg_batch_create: ") + mastg_last_error(nullptr) +
                                     " (status " + std::to_string(rc) + ")");
    }
    FunctionEvaluationBatch(const FunctionEvaluationBatch&) = delete;
    ~FunctionEvaluationBatch() { mastg_batch_destroy(_b); }

    // dvars[i]: the element parameters of design i; obj[i], obj_grad[i]
    void evaluate(const std::vector<RealVector>& dvars, std::vector<double>& obj, bool eval_obj_grad,
                  std::vector<RealVector>& obj_grad, const mastg_solve_opts* opts = nullptr) {
        const size_t n = dvars.size();
        RealVector p(n * size_t(_n_elems)), g(eval_obj_grad ? n * size_t(_n_elems) : 0);
        for (size_t i = 0; i < n; i++) {
            if (int64_t(dvars[i].size()) != _n_elems)
                throw std::runtime_error("FunctionEvaluationBatch::evaluate: design of the wrong length");
            std::copy(dvars[i].begin(), dvars[i].end(), p.begin() + i * size_t(_n_elems));
        }
        std::vector<mastg_design_result> r(n);
        const int rc = mastg_batch_eval(_b, int(n), p.data(), eval_obj_grad ? g.data() : nullptr, MASTG_HOST, opts,
                                        r.data(), nullptr);
        if (rc != MASTG_OK)
            throw std::runtime_error(std::string("FunctionEvaluationBatch::evaluate: ") + mastg_batch_last_error(_b) +
                                     " (status " + std::to_string(rc) + ")");
        obj.resize(n);
        obj_grad.assign(eval_obj_grad ? n : 0, RealVector());
        for (size_t i = 0; i < n; i++) {
            obj[i] = r[i].objective;
            if (eval_obj_grad)
                obj_grad[i].assign(g.begin() + i * size_t(_n_elems), g.begin() + (i + 1) * size_t(_n_elems));
        }
        last = r;
    }
    std::vector<mastg_design_result> last;

private:
    mastg_batch* _b = nullptr;
    int64_t _n_elems;
};

}  // namespace GPU
}  // namespace MAST
