// Drop-in greenFunctionMC::occupation_numbers_compute over the C ABI (two-sided chain
// engine, le_engine_create_occupation).
//
// Same namespace, parameter class and overloads as the reference
// (modules/green_function.h:33-69, 175-183; gf_parameters.cu; gf_occupation_numbers.cu:309-562):
// dL > ansatzMaxAngularMomentum ends the process with a message; orbitals
// l < angularMomentumCutoff = m (N - 1) + ansatzMaxAngularMomentum; each (a, b) entry is
// normalised to sum_l O_ab(l) = N (times metric(a, b) when a metric is given); the tensor
// is made self-adjoint per l; the TSV rows are "a\tb\tl\tre\tim" with a blank line after
// each (a, b), under ../output/N={N}_m={m}/dL={dL}/Statistics/
// OccupationNumbers_{normalized|not_normalized}/fileNumber.tsv.
// What changes underneath: one chain on |Psi(z_0..z_{N-1})| |Psi(z_1..z_N)| serves all
// d^2 pairs (the reference launches one chain per pair).
//
// spectral_function_compute has no reference implementation (README.md:14-15 "coming soon";
// the commented-out spectral_function_sampler, modules/green_function.h:18,90): it writes the
// particle-removal spectral function of the Laughlin state at the edge orbitals,
// A(l, w) = n_l sum_j |<j, N-1| c_l |Psi_L^N>|^2 / n_l delta(w - (E_0^N - E_j^{N-1})),
// from three runs: the occupations n_l (two-sided chain, dL = 0), E_0^N (sector 0 of N
// particles) and the (N-1)-particle sectors dL' = 0..dLmax, orbital l = m (N - 1) - dL'
// (le_removal_weights).  Rows "l\tj\tomega\tweight" under
// ../output/N={N}_m={m}/SpectralFunction/fileNumber.tsv.
#ifndef LE_GREEN_FUNCTION_H
#define LE_GREEN_FUNCTION_H

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "complex_numbers.h"
#include "laughlinedge.h"
#include <sys_params.h>   // on the include path: regenerate it, or put your own first
#include "tensors.h"

namespace greenFunctionMC {

class monteCarloParameters {
public:
    int maxDegree = -1;
    int extendedMaxDegree = -1;
    int subspaceDimension = -1;
    int injectedAngularMomentum = 0;
    float randomStep;
    unsigned int burnInMoves;
    unsigned int MCSamples;
    unsigned int threadsPerBlock;
    unsigned int nBlocks;
    unsigned int gridSize;
    size_t totalSamplesNumber;
    unsigned int targetAcceptanceMoves;
    float targetAcceptance;
    float acceptance_tol;
    // no reference counterpart
    unsigned long long seed = 0;      // 0: seeded from the clock, like the reference
    int precision = LE_PREC_FP32;

    monteCarloParameters()
    {
        randomStep = 0.8f;
        burnInMoves = 1000;
        MCSamples = 50000;
        targetAcceptanceMoves = 100;
        targetAcceptance = 0.50f;
        acceptance_tol = 0.01f;
        threadsPerBlock = 128;
        nBlocks = 32;
        gridSize = nBlocks * threadsPerBlock;
        totalSamplesNumber = (size_t)MCSamples * gridSize;
    }
    void setEdgeSpaceInfo(int maxPolyDegree, int subDim)
    {
        maxDegree = maxPolyDegree;
        extendedMaxDegree = maxDegree + 1;
        subspaceDimension = subDim;
    }
    void setBurnIn(unsigned int n) { burnInMoves = n; }
    void setMCSamples(unsigned int n)
    {
        MCSamples = n;
        totalSamplesNumber = (size_t)MCSamples * gridSize;
    }
    void setGPU_threads_blocks(unsigned int threads, unsigned int blocks)
    {
        threadsPerBlock = threads;
        nBlocks = blocks;
        gridSize = nBlocks * threadsPerBlock;
        totalSamplesNumber = (size_t)MCSamples * gridSize;
    }
    void setAcceptanceBisectionParams(unsigned int moves, float target, float tol)
    {
        targetAcceptanceMoves = moves;
        targetAcceptance = target;
        acceptance_tol = tol;
    }
    void setRandomStep(float s) { randomStep = s; }
};

namespace detail {

inline le_mc_params mc_params(const monteCarloParameters& params)
{
    le_mc_params m;
    le_mc_params_default(&m);
    m.walkers = (int)params.gridSize;
    m.burnin_sweeps = (int)params.burnInMoves;
    m.samples = (int)params.MCSamples;
    m.step = params.randomStep;
    m.tune_sweeps = (int)params.targetAcceptanceMoves;
    m.target_acceptance = params.targetAcceptance;
    m.acceptance_tol = params.acceptance_tol;
    if (params.seed) {
        m.seed = params.seed;
    } else {
        std::srand((unsigned)std::time(nullptr));
        m.seed = (unsigned long long)std::rand();
    }
    m.precision = params.precision;
    return m;
}

inline void check(int rc)
{
    if (rc != LE_OK) {
        std::cout << "\nlaughlinedge: " << le_last_error() << std::endl;
        std::exit(1);
    }
}

inline void on_compute(monteCarloParameters& params, int dL, const cmatrix<double>* metric, int fileNumber)
{
    if (dL > sys_params::ansatzMaxAngularMomentum) {
        std::cout << "The defined parameter ansatzMaxAngularMomentum is smaller than dL. Modify and re-run "
                     "system_parameters.cpp"
                  << std::endl;
        std::exit(1);
    }
    params.injectedAngularMomentum = dL;
    le_system_params s{};
    s.particles_number = sys_params::particlesNumber;
    s.inverse_filling = sys_params::inverseFilling;
    const le_mc_params m = mc_params(params);
    const int lcut = sys_params::angularMomentumCutoff;
    le_engine* e = nullptr;
    check(le_engine_create_occupation(&s, &m, dL, lcut, &e));
    le_engine_info info;
    check(le_engine_get_info(e, &info));
    const int d = info.D / 2;
    params.setEdgeSpaceInfo(info.K, d);
    if (metric && (metric->rows() != d || metric->cols() != d)) {
        std::cout << "\nThe metric must be " << d << " x " << d << std::endl;
        std::exit(1);
    }
    check(le_run(e));
    le_run_stats st;
    check(le_get_stats(e, &st));
    params.setRandomStep((float)st.step);
    std::printf("\tSampled %lld configurations (%d walkers), acceptance %.2f, step %.4f\n", st.samples,
                (int)params.gridSize, st.acceptance, st.step);
    std::vector<double> raw(2 * (size_t)d * d * lcut);
    check(le_get_occupation(e, -1, raw.data()));
    le_engine_destroy(e);
    rank3tensor<complex<double>> O(d, d, lcut);
    for (int a = 0; a < d; ++a)
        for (int b = 0; b < d; ++b) {
            complex<double> sum(0.0);
            for (int l = 0; l < lcut; ++l) {
                const size_t q = ((size_t)a * d + b) * lcut + l;
                sum += complex<double>(raw[2 * q], raw[2 * q + 1]);
            }
            const complex<double> scale =
                metric ? (*metric)(a, b) * (double)sys_params::particlesNumber : complex<double>((double)sys_params::particlesNumber);
            for (int l = 0; l < lcut; ++l) {
                const size_t q = ((size_t)a * d + b) * lcut + l;
                O(a, b, l) = complex<double>(raw[2 * q], raw[2 * q + 1]) / sum * scale;
            }
        }
    rank3tensor<complex<double>> Os(d, d, lcut);
    for (int l = 0; l < lcut; ++l)
        for (int a = 0; a < d; ++a)
            for (int b = 0; b < d; ++b) Os(a, b, l) = (O(a, b, l) + conj(O(b, a, l))) / 2.0;
    const std::string dir = "../output/N=" + std::to_string(sys_params::particlesNumber) + "_m=" +
                            std::to_string(sys_params::inverseFilling) + "/dL=" + std::to_string(dL) +
                            "/Statistics/OccupationNumbers_" + (metric ? "normalized" : "not_normalized");
    std::filesystem::create_directories(dir);
    std::ofstream out(dir + "/" + std::to_string(fileNumber) + ".tsv");
    char buf[160];
    for (int a = 0; a < d; ++a)
        for (int b = 0; b < d; ++b) {
            for (int l = 0; l < lcut; ++l) {
                std::snprintf(buf, sizeof buf, "%d\t%d\t%d\t%.6f\t%.6f\n", a, b, l, re(Os(a, b, l)), im(Os(a, b, l)));
                out << buf;
            }
            out << "\n";
        }
}

}  // namespace detail

inline void occupation_numbers_compute(monteCarloParameters& params, int dL, cmatrix<double>& metric, int fileNumber = 0)
{
    detail::on_compute(params, dL, &metric, fileNumber);
}

inline void occupation_numbers_compute(monteCarloParameters& params, int dL, int fileNumber = 0)
{
    detail::on_compute(params, dL, nullptr, fileNumber);
}

// Particle-removal spectral function of the N-particle Laughlin state at the orbitals
// l = m (N - 1) - dL', dL' = 0..dLmax (dLmax <= ansatzMaxAngularMomentum and <= N - 1).
inline void spectral_function_compute(monteCarloParameters& params, int dLmax, int fileNumber = 0)
{
    const int N = sys_params::particlesNumber, m = sys_params::inverseFilling;
    if (dLmax < 0 || dLmax > sys_params::ansatzMaxAngularMomentum || dLmax > N - 1) {
        std::cout << "dLmax must be in [0, min(ansatzMaxAngularMomentum, N - 1)]. Modify and re-run" << std::endl;
        std::exit(1);
    }
    using detail::check;
    const le_mc_params mc = detail::mc_params(params);
    // 1. occupations n_l of the Laughlin state (two-sided chain, dL = 0), normalised to sum N
    const int lcut = m * (N - 1) + 1;
    le_system_params so{};
    so.particles_number = N;
    so.inverse_filling = m;
    le_engine* e = nullptr;
    check(le_engine_create_occupation(&so, &mc, 0, lcut, &e));
    check(le_run(e));
    std::vector<double> raw(2 * (size_t)lcut);
    check(le_get_occupation(e, -1, raw.data()));
    le_run_stats st;
    check(le_get_stats(e, &st));
    params.setRandomStep((float)st.step);
    le_engine_destroy(e);
    double total = 0.0;
    for (int l = 0; l < lcut; ++l) total += raw[2 * l];
    std::vector<double> nl(lcut);
    for (int l = 0; l < lcut; ++l) nl[l] = raw[2 * l] / total * N;
    // 2. E_0 of the N-particle Laughlin state (sector 0: S = 1)
    le_system_params sn{};
    sn.particles_number = N;
    sn.inverse_filling = m;
    sn.conf_w = sys_params::confinementW;
    sn.conf_eps = sys_params::confinementEps;
    sn.conf_strength = sys_params::confinementStrength;
    sn.conf_exponent = sys_params::confinementExponent;
    const int zero = 0;
    check(le_engine_create(&sn, &mc, &zero, 1, nullptr, &e));
    check(le_run(e));
    double E0 = 0.0;
    check(le_engine_spectrum(e, 0, -1, &E0));
    le_engine_destroy(e);
    // 3. N - 1 particles in the same trap: strength (r / R_N)^e = strength' (r / R_{N-1})^e
    le_system_params s1 = sn;
    s1.particles_number = N - 1;
    s1.conf_strength = sn.conf_strength * std::pow((double)(N - 1) / N, 0.5 * sn.conf_exponent);
    std::vector<int> sectors(dLmax + 1);
    for (int k = 0; k <= dLmax; ++k) sectors[k] = k;
    check(le_engine_create(&s1, &mc, sectors.data(), (int)sectors.size(), nullptr, &e));
    check(le_run(e));
    const std::string dir = "../output/N=" + std::to_string(N) + "_m=" + std::to_string(m) + "/SpectralFunction";
    std::filesystem::create_directories(dir);
    std::ofstream out(dir + "/" + std::to_string(fileNumber) + ".tsv");
    char buf[160];
    std::vector<int> dims(dLmax + 1), offs(dLmax + 2);
    check(le_sector_dims(e, dims.data(), offs.data()));
    for (int k = 0; k <= dLmax; ++k) {
        const int d = dims[k];
        std::vector<double> E(d), w(d);
        int l = 0;
        check(le_removal_weights(e, k, -1, &l, E.data(), w.data()));
        for (int j = 0; j < d; ++j) {
            std::snprintf(buf, sizeof buf, "%d\t%d\t%.6f\t%.6e\n", l, j, E0 - E[j], nl[l] * w[j]);
            out << buf;
        }
    }
    le_engine_destroy(e);
}

}  // namespace greenFunctionMC

#endif
