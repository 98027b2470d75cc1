// Compliance and its adjoint sensitivity through the MAST-style C++ layer,
// following the call sequence of the reference's SIMP driver
// (examples/structural/example_6/example_6.cpp:450-655, 877-939):
// attach system + elem ops, solve, evaluate compliance, adjoint solve with
// rhs = -dC/dX, then sens_e = lambda . dR/drho_e for every element.  Prints one
// JSON line and cross-checks the result against mastg_design_eval.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "mast_gpu.hpp"

int main(int argc, char** argv) {
    const int nx = argc > 1 ? std::atoi(argv[1]) : 12;
    mastg_workload* w = nullptr;
    if (mastg_workload_create(1, nx, nx / 2, nx / 2, 3, 0.01, 1.0, 1, &w)) return 2;
    const mastg_problem_desc& d = *mastg_workload_desc(w);
    MAST::GPU::RealVector rho(mastg_workload_params(w), mastg_workload_params(w) + d.n_elems);
    try {
        MAST::GPU::SystemDescription sys(d);
        MAST::GPU::StructuralNonlinearAssemblyElemOperations ops;
        MAST::GPU::NonlinearImplicitAssembly assembly;
        assembly.set_discipline_and_system(sys);
        assembly.set_elem_operation_object(ops);
        assembly.set_element_parameters(rho);
        MAST::GPU::NonlinearSystem system(assembly);
        MAST::GPU::RealVector X;
        system.solve(X);
        MAST::GPU::ComplianceOutput compliance;
        const double C = compliance.evaluate(assembly, X);
        MAST::GPU::RealVector dq, adj, g1;
        compliance.output_derivative(assembly, X, dq);
        system.adjoint_solve(dq, adj);
        assembly.calculate_output_adjoint_sensitivity_multiple_parameters_no_direct(X, adj, g1);
        // NonlinearSystem::adjoint_solve forms rhs = -dq/dX, so lambda = -X for
        // compliance and g1 = lambda . dR/drho is dC/drho itself (example_6 passes
        // lambda = X instead and uses -g1; the direct term g2 is 0 for point loads).
        // Cross-check against the fused device pipeline:
        MAST::GPU::RealVector g(d.n_elems);
        mastg_design_result r{};
        MAST::GPU::check(mastg_design_eval(assembly.context(), rho.data(), g.data(), MASTG_HOST, nullptr, &r),
                         assembly.context(), "design_eval");
        double emax = 0, gmax = 0;
        for (int64_t e = 0; e < d.n_elems; e++) {
            emax = std::fmax(emax, std::fabs(g1[e] - g[e]));
            gmax = std::fmax(gmax, std::fabs(g[e]));
        }
        const double ec = std::fabs(C - r.objective) / std::fabs(r.objective);
        // the same design and two others evaluated side by side on two lanes:
        // design 0 agrees with the lone evaluation (bitwise when the lanes
        // run the launch shapes the assembly's context tuned; to rounding else)
        std::vector<MAST::GPU::RealVector> designs{rho, rho, rho};
        for (int64_t e = 0; e < d.n_elems; e++) {
            designs[1][e] = 0.5 * (rho[e] + rho[(e + 1) % d.n_elems]);
            designs[2][e] = std::fmin(1.0, 1.1 * rho[e]);
        }
        MAST::GPU::FunctionEvaluationBatch batch(sys, 2);
        std::vector<double> obj;
        std::vector<MAST::GPU::RealVector> obj_grad;
        batch.evaluate(designs, obj, true, obj_grad);
        bool same = obj[0] == r.objective;
        double bmax = 0;
        for (int64_t e = 0; e < d.n_elems; e++) {
            same = same && obj_grad[0][e] == g[e];
            bmax = std::fmax(bmax, std::fabs(obj_grad[0][e] - g[e]));
        }
        const double bc = std::fabs(obj[0] - r.objective) / std::fabs(r.objective);
        std::printf("{\"n_elems\": %lld, \"compliance\": %.15g, \"design_eval_compliance\": %.15g, "
                    "\"compliance_rel_diff\": %.3g, \"sens_max_rel_diff\": %.3g, \"iters\": %d, "
                    "\"batch_objectives\": [%.15g, %.15g, %.15g], \"batch_design0_rel_diff\": %.3g, "
                    "\"batch_design0_grad_rel_diff\": %.3g, \"batch_design0_bitwise\": %s}\n",
                    (long long)d.n_elems, C, r.objective, ec, emax / gmax, system.last.iters, obj[0], obj[1], obj[2],
                    bc, bmax / gmax, same ? "true" : "false");
        mastg_workload_destroy(w);
        return (ec < 1e-8 && emax / gmax < 1e-6 && bc < 1e-12 && bmax / gmax < 1e-10) ? 0 : 1;
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "error: %s\n", ex.what());
        mastg_workload_destroy(w);
        return 3;
    }
}
