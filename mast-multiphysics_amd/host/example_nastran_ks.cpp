// A Nastran model through the MAST-style C++ layer: NastranIO::read, a
// filtered density field (FilterBase), the static solve, the KS von Mises
// stress output and its adjoint sensitivity (output_derivative ->
// adjoint_solve -> direct + adjoint terms, AssemblyBase's
// calculate_output_adjoint_sensitivity pattern, src/base/assembly_base.cpp:
// 594-639), the forward sensitivity of the state to E (NonlinearSystem::
// sensitivity_solve, src/base/nonlinear_system.cpp:742-794) and the
// forward-adjoint check dC/dE = F . dX/dE = -X . dR/dE for the compliance.
// Prints one JSON line; exit status 0 when the checks hold.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "mast_gpu.hpp"

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s model.bdf [filter radius]\n", argv[0]);
        return 2;
    }
    const double radius = argc > 2 ? std::atof(argv[2]) : 0.15;
    try {
        MAST::GPU::NastranIO io;
        io.read(argv[1]);
        MAST::GPU::SystemDescription sys = io.system();
        const int64_t ne = sys.desc().n_elems;
        // a density field: smooth ramp over the element order, then filtered
        MAST::GPU::RealVector x(ne), rho;
        for (int64_t e = 0; e < ne; e++) x[e] = 0.3 + 0.6 * double((e * 37) % ne) / double(ne);
        MAST::GPU::FilterBase filter(sys, radius);
        filter.compute_filtered_values(x, rho);
        const double vol = filter.volume(rho, nullptr);

        MAST::GPU::StructuralNonlinearAssemblyElemOperations ops;
        MAST::GPU::NonlinearImplicitAssembly assembly;
        assembly.set_discipline_and_system(sys);
        assembly.set_elem_operation_object(ops);
        assembly.set_element_parameters(rho);
        MAST::GPU::NonlinearSystem system(assembly);
        MAST::GPU::RealVector X;
        system.solve(X);

        // KS stress: value, dKS/dX, adjoint, total dKS/drho_e
        MAST::GPU::KSStressStrainOutput ks(10.0, 2.0e6);
        const double k = ks.evaluate(assembly, X);
        MAST::GPU::RealVector dq, adj, g_adj, g_dir;
        ks.output_derivative(assembly, X, dq);
        system.adjoint_solve(dq, adj);   // J^T adj = -dKS/dX
        assembly.calculate_output_adjoint_sensitivity_multiple_parameters_no_direct(X, adj, g_adj);
        ks.output_sensitivity(assembly, X, g_dir);
        double gmax = 0;
        for (int64_t e = 0; e < ne; e++) gmax = std::fmax(gmax, std::fabs(g_dir[e] + g_adj[e]));

        // the weighted p-norm of the von Mises stress (StressStrainOutputBase,
        // example_6's stress constraint): dF/dX along X against a central
        // difference of the value
        MAST::GPU::StressStrainOutputBase pn;
        pn.set_aggregation_coefficients(6.0, 1.0, 0.5, 2.0e6);
        const double F0 = pn.evaluate(assembly, X);
        MAST::GPU::RealVector dF, Xp(X), Xm(X);
        pn.output_derivative(assembly, X, dF);
        const double h = 1e-6;
        double an = 0;
        for (size_t i = 0; i < X.size(); i++) {
            Xp[i] *= 1 + h;
            Xm[i] *= 1 - h;
            an += dF[i] * X[i];
        }
        const double fd = (pn.evaluate(assembly, Xp) - pn.evaluate(assembly, Xm)) / (2 * h);
        const double e3 = std::fabs(fd - an) / std::fabs(an);

        // compliance and its E derivative, forward and adjoint
        MAST::GPU::ComplianceOutput compliance;
        const double C = compliance.evaluate(assembly, X);
        MAST::GPU::RealVector F, dXdE;
        compliance.output_derivative(assembly, X, F);
        system.sensitivity_solve(X, MASTG_PARAM_E, dXdE);
        double fwd = 0;
        for (size_t i = 0; i < F.size(); i++) fwd += F[i] * dXdE[i];
        const double adj_E = -assembly.adjoint_sensitivity(MASTG_PARAM_E, X, X);
        const double E = sys.desc().E;
        const double e1 = std::fabs(fwd - adj_E) / std::fabs(adj_E);
        const double e2 = std::fabs(adj_E + C / E) / std::fabs(C / E);   // C is homogeneous of degree -1 in E
        std::printf("{\"n_elems\": %lld, \"volume\": %.12g, \"compliance\": %.12g, \"ks\": %.12g, "
                    "\"ks_grad_max\": %.6g, \"pnorm\": %.12g, \"pnorm_dX_vs_fd\": %.3g, "
                    "\"dCdE_forward_vs_adjoint\": %.3g, \"dCdE_vs_minus_C_over_E\": %.3g, \"iters\": %d}\n",
                    (long long)ne, vol, C, k, gmax, F0, e3, e1, e2, system.last.iters);
        return (e1 < 1e-8 && e2 < 1e-8 && e3 < 1e-6 && std::isfinite(k) && gmax > 0 && F0 > 0) ? 0 : 1;
    } catch (const std::exception& ex) {
        std::fprintf(stderr, "error: %s\n", ex.what());
        return 3;
    }
}
