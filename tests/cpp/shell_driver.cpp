// Drives the drop-in C++ interface (include/edge_montecarlo.h) the way the reference's
// main.cu snippets do, with a fixed seed, and prints machine-readable results.
#include <cstdio>

#include "dynamics.h"
#include "edge_montecarlo.h"
#include "green_function.h"

int main()
{
    edgeMC::monteCarloParameters p;
    p.setGPU_threads_blocks(64, 8);      // 512 walkers
    p.setMCSamples(400);
    p.setBurnIn(200);
    p.seed = 1234;

    std::vector<double> ev = edgeMC::spectrum_compute(p, 4, true, 7);
    std::printf("SPECTRUM");
    for (double e : ev) std::printf(" %.12e", e);
    std::printf("\nSTEP %.9e\nDIMS %d %d\n", p.randomStep, p.subspaceDimension, p.maxDegree);

    auto dsf = edgeMC::dsf_compute(p, 3, true, 7);
    std::printf("DSF");
    for (auto& x : dsf) std::printf(" %.12e %.12e", x.first, x.second);
    std::printf("\n");

    auto U = [] __device__(complex<float>* position) {
        double u = 0;
        for (int i = 0; i < sys_params::particlesNumber; i++) u += pow(re(position[i]) / sys_params::RCl, 2);
        return u;
    };
    vector<int> sectors = {0, 2, 4};
    edgeMC::excitationResponse_compute(p, sectors, U, 0);

    auto F = [](double t) {
        const double z = (t - 2.5) / 1.0;
        return 0.1 * std::exp(-z * z);
    };
    timeEvolution::edgeDensityResponse_compute(sys_params::particlesNumber, sys_params::inverseFilling, {0, 2, 4}, F,
                                               5.0, 0.5, 0);
    // occupation numbers of dL = 2, normalised with the metric of a spectrum run
    edgeMC::monteCarloParameters p2 = p;
    p2.seed = 77;
    std::vector<double> ev2 = edgeMC::spectrum_compute(p2, 2, true, 5);
    cmatrix<double> M2 = edgeMC::metric_compute(p2, 2, false, 5);
    greenFunctionMC::monteCarloParameters gp;
    gp.setGPU_threads_blocks(64, 32);
    gp.setMCSamples(400);
    gp.setBurnIn(200);
    gp.seed = 99;
    greenFunctionMC::occupation_numbers_compute(gp, 2, M2, 5);
    greenFunctionMC::occupation_numbers_compute(gp, 2, 5);
    greenFunctionMC::spectral_function_compute(gp, 3, 6);
    // the reference's own estimator: 64 chains on |p_f Psi_L|^2 for each of the 5 rows of dL = 4
    edgeMC::monteCarloParameters pr;
    pr.setGPU_threads_blocks(64, 1);
    pr.setMCSamples(400);
    pr.setBurnIn(200);
    pr.seed = 4321;
    pr.estimator = LE_EST_ROWWISE;
    std::vector<double> evr = edgeMC::spectrum_compute(pr, 4, false, 0);
    std::printf("ROWWISE");
    for (double e : evr) std::printf(" %.12e", e);
    std::printf("\nROWSTEP %.9e\n", pr.randomStep);
    // the mixture estimator: one population on w |Psi_L|^2, S and H reweighted by 1 / w
    edgeMC::monteCarloParameters pm = pr;
    pm.setGPU_threads_blocks(64, 8);
    pm.seed = 2468;
    pm.randomStep = 0.8f;                // the row-wise run wrote its tuned step back into pr
    pm.estimator = LE_EST_MIXTURE;
    std::vector<double> evm = edgeMC::spectrum_compute(pm, 4, false, 0);
    std::printf("MIXTURE");
    for (double e : evm) std::printf(" %.12e", e);
    std::printf("\nMIXSTEP %.9e\n", pm.randomStep);
    std::printf("DONE\n");
    return 0;
}
