// Drop-in C++ interface of the edge Monte-Carlo computations, over the C ABI
// (laughlinedge.h).  Header-only; include it where the reference's
// modules/edge_montecarlo.h was included and link liblaughlinedge.so.
//
// Same namespace, class and function signatures as the reference
// (modules/edge_montecarlo.h:45-88, 199-237):
//   edgeMC::monteCarloParameters            em_parameters.cu:9-79 (defaults and setters)
//   edgeMC::metric_compute                  em_metric.cu:200-392
//   edgeMC::spectrum_compute                em_spectrum.cu:210-422
//   edgeMC::dsf_compute                     em_dsf.cu:235-500
//   edgeMC::excitationResponse_compute<F>   em_edgeResponse.tpp:284-601 (needs a HIP
//                                           translation unit for the __device__ functor)
// Behaviour kept: the normalised matrices (unit-diagonal metric, (X + X^H)/2), the
// capacity checks against sys_params.h (message + exit(1)), the time-seeded RNG when no
// seed is given, the step bisection updating params.randomStep, setEdgeSpaceInfo, and
// the TSV files and paths.  What changes underneath: one |Psi_L|^2 sampling run serves
// every row and sector (see laughlinedge.h), so each call costs one run instead of one
// chain per basis state.
#ifndef LE_EDGE_MONTECARLO_H
#define LE_EDGE_MONTECARLO_H

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "complex_numbers.h"
#include "generalized_eigenproblem.h"
#include "laughlinedge.h"
#include <sys_params.h>   // on the include path: regenerate it, or put your own first
#include "tensors.h"

namespace edgeMC {

class monteCarloParameters {
public:
    int maxDegree = -1;
    int extendedMaxDegree = -1;
    int subspaceDimension = -1;
    int injectedAngularMomentum = 0;
    int* injectedAngularMomenta = nullptr;
    int numberOfSectors = 0;
    int mesh_points_number;
    float dr;
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
    // no reference counterpart: reproducible seeds, the FP64 form, time blocks for errors
    unsigned long long seed = 0;      // 0: seeded from the clock, like the reference
    int precision = LE_PREC_FP32;
    int statisticsBlocks = 1;
    // LE_EST_PLAIN: one |Psi_L|^2 population for every row; LE_EST_ROWWISE: the reference's
    // estimator, gridSize chains on |p_f Psi_L|^2 for every row f (em_spectrum.cu:16-137);
    // LE_EST_MIXTURE: one w |Psi_L|^2 population, S and H reweighted by 1 / w (bounded
    // integrands in every sector; spectrum and metric only)
    int estimator = LE_EST_PLAIN;

    monteCarloParameters()
    {
        mesh_points_number = 200;
        dr = 1.7f * sys_params::DCl / mesh_points_number;
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
    void setMesh(int n)
    {
        mesh_points_number = n;
        dr = 1.7f * sys_params::DCl / mesh_points_number;
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

inline void fatal(const std::string& msg)
{
    std::cout << msg << std::endl;
    std::exit(1);
}

inline void check(int rc)
{
    if (rc != LE_OK) fatal(std::string("\nlaughlinedge: ") + le_last_error());
}

inline le_system_params system()
{
    le_system_params s{};
    s.particles_number = sys_params::particlesNumber;
    s.inverse_filling = sys_params::inverseFilling;
    s.conf_w = sys_params::confinementW;
    s.conf_eps = sys_params::confinementEps;
    s.conf_strength = sys_params::confinementStrength;
    s.conf_exponent = sys_params::confinementExponent;
    return s;
}

inline le_mc_params mc(const monteCarloParameters& p)
{
    le_mc_params m;
    le_mc_params_default(&m);
    m.walkers = (int)p.gridSize;
    m.burnin_sweeps = (int)p.burnInMoves;
    m.samples = (int)p.MCSamples;
    m.thin = 1;
    m.step = p.randomStep;
    m.tune_sweeps = (int)p.targetAcceptanceMoves;
    m.target_acceptance = p.targetAcceptance;
    m.acceptance_tol = p.acceptance_tol;
    if (p.seed) {
        m.seed = p.seed;
    } else {
        std::srand((unsigned)std::time(nullptr));
        m.seed = (unsigned long long)std::rand();
    }
    m.precision = p.precision;
    m.estimator = p.estimator;
    m.n_blocks = p.statisticsBlocks > 0 ? p.statisticsBlocks : 1;
    return m;
}

struct BasisInfo {
    int D = 0, K = 0, longest = 0;
    std::vector<int> dl;                       // per state
    std::vector<std::vector<std::pair<int, int>>> pairs;
};

inline BasisInfo basis(const std::vector<int>& sectors)
{
    BasisInfo b;
    check(le_basis_describe(sys_params::particlesNumber, sectors.data(), (int)sectors.size(), 0, &b.D, &b.K,
                            nullptr, nullptr, nullptr));
    std::vector<int> pr(16 * (size_t)b.D), np(b.D);
    b.dl.resize(b.D);
    check(le_basis_describe(sys_params::particlesNumber, sectors.data(), (int)sectors.size(), b.D, &b.D, &b.K,
                            b.dl.data(), pr.data(), np.data()));
    b.pairs.resize(b.D);
    for (int t = 0; t < b.D; ++t) {
        for (int q = 0; q < np[t]; ++q) b.pairs[t].push_back({pr[16 * t + 2 * q], pr[16 * t + 2 * q + 1]});
        if (np[t] > b.longest) b.longest = np[t];
    }
    return b;
}

// capacity checks of em_spectrum.cu:235-252 against the sys_params.h bounds
inline void capacity(monteCarloParameters& p, const BasisInfo& b)
{
    p.setEdgeSpaceInfo(b.K, b.D);
    if (p.extendedMaxDegree > sys_params::ansatzExtendedMaxDegree)
        fatal("\nWith this angular momentum choice, the maximal single-coordinate degree is " +
              std::to_string(p.maxDegree) + "\nThe upper bound in sys_params.h is " +
              std::to_string(sys_params::ansatzExtendedMaxDegree - 1) + "; regenerate sys_params.h");
    if (b.longest > sys_params::ansatzMaxPartitionSize)
        fatal("\nWith this angular momentum choice, the longest (compressed) partition has length " +
              std::to_string(b.longest) + "\nThe upper bound in sys_params.h is " +
              std::to_string(sys_params::ansatzMaxPartitionSize) + "; regenerate sys_params.h");
    if (b.D > sys_params::ansatzSubspaceDimension)
        fatal("\nWith this angular momentum choice, the edge Hilbert space size is " + std::to_string(b.D) +
              "\nThe upper bound in sys_params.h is " + std::to_string(sys_params::ansatzSubspaceDimension) +
              "; regenerate sys_params.h");
}

// One sampling run over `sectors`; owns the engine.
struct Run {
    le_engine* e = nullptr;
    std::vector<int> dims, offs;
    Run(monteCarloParameters& p, const std::vector<int>& sectors, const le_observables* obs, le_weight_fn fn = nullptr,
        void* ctx = nullptr)
    {
        le_system_params s = system();
        le_mc_params m = mc(p);
        check(le_engine_create(&s, &m, sectors.data(), (int)sectors.size(), obs, &e));
        if (fn) check(le_set_weight_callback(e, fn, ctx));
        dims.resize(sectors.size());
        offs.resize(sectors.size() + 1);
        check(le_sector_dims(e, dims.data(), offs.data()));
        check(le_run(e));
        le_run_stats st;
        check(le_get_stats(e, &st));
        p.setRandomStep((float)st.step);
        std::printf("\tSampled %lld configurations (%d walkers), acceptance %.2f, step %.4f\n", st.samples,
                    (int)p.gridSize, st.acceptance, st.step);
    }
    ~Run() { le_engine_destroy(e); }
    Run(const Run&) = delete;
    Run& operator=(const Run&) = delete;

    std::vector<complex<double>> block(int weight, int s1, int s2) const
    {
        std::vector<double> raw(2 * (size_t)dims[s1] * dims[s2]);
        check(le_get_block(e, -1, weight, s1, s2, raw.data()));
        std::vector<complex<double>> out(raw.size() / 2);
        for (size_t q = 0; q < out.size(); ++q) out[q] = complex<double>(raw[2 * q], raw[2 * q + 1]);
        return out;
    }
    // per-state norms sqrt(S_aa) over the whole basis
    std::vector<double> norms() const
    {
        std::vector<double> nrm(offs.back());
        for (size_t s = 0; s < dims.size(); ++s) {
            auto S = block(LE_W_ONE, (int)s, (int)s);
            for (int a = 0; a < dims[s]; ++a) nrm[offs[s] + a] = std::sqrt(re(S[(size_t)a * dims[s] + a]));
        }
        return nrm;
    }
    // place a normalised block into a D x D matrix
    void place(cmatrix<double>& M, int weight, int s1, int s2, const std::vector<double>& nrm) const
    {
        auto B = block(weight, s1, s2);
        for (int a = 0; a < dims[s1]; ++a)
            for (int b = 0; b < dims[s2]; ++b) {
                const int i = offs[s1] + a, j = offs[s2] + b;
                M(i, j) = B[(size_t)a * dims[s2] + b] / (nrm[i] * nrm[j]);
            }
    }
};

inline cmatrix<double> self_adjoint(const cmatrix<double>& m) { return (m + m.adjoint()) / 2.0; }

inline std::string out_dir(const std::string& tail)
{
    return "../output/N=" + std::to_string(sys_params::particlesNumber) + "_m=" +
           std::to_string(sys_params::inverseFilling) + "/" + tail;
}

inline void write_tsv(const std::string& dir, int fileNumber, const cmatrix<double>& M)
{
    std::filesystem::create_directories(dir);
    std::ofstream out(dir + "/" + std::to_string(fileNumber) + ".tsv");
    char buf[128];
    for (int i = 0; i < M.rows(); ++i)
        for (int j = 0; j < M.cols(); ++j) {
            std::snprintf(buf, sizeof buf, "%d\t%d\t%.6f\t%.6f\n", i, j, re(M(i, j)), im(M(i, j)));
            out << buf;
        }
}

inline void sector_matrices(monteCarloParameters& p, int dL, cmatrix<double>& H, cmatrix<double>& M)
{
    std::vector<int> sec = {dL};
    BasisInfo b = basis(sec);
    capacity(p, b);
    Run run(p, sec, nullptr);
    auto nrm = run.norms();
    H = cmatrix<double>(b.D, b.D);
    M = cmatrix<double>(b.D, b.D);
    run.place(H, LE_W_V, 0, 0, nrm);
    run.place(M, LE_W_ONE, 0, 0, nrm);
    H = self_adjoint(H);
    M = self_adjoint(M);
}

}  // namespace detail

inline cmatrix<double> metric_compute(monteCarloParameters& params, int dL, bool saveOutput = false,
                                      int fileNumber = 0)
{
    cmatrix<double> H, M;
    detail::sector_matrices(params, dL, H, M);
    if (saveOutput) detail::write_tsv(detail::out_dir("dL=" + std::to_string(dL) + "/Statistics/Metric"), fileNumber, M);
    return M;
}

inline std::vector<double> spectrum_compute(monteCarloParameters& params, int dL, bool saveOutput = false,
                                            int fileNumber = 0)
{
    cmatrix<double> H, M;
    detail::sector_matrices(params, dL, H, M);
    if (saveOutput) {
        detail::write_tsv(detail::out_dir("dL=" + std::to_string(dL) + "/Statistics/Hamiltonian"), fileNumber, H);
        detail::write_tsv(detail::out_dir("dL=" + std::to_string(dL) + "/Statistics/Metric"), fileNumber, M);
    }
    ges::generalizedEigenproblem eig(H, M);
    return eig.extract_spectrum();
}

inline std::vector<std::pair<double, double>> dsf_compute(monteCarloParameters& params, int dL,
                                                          bool saveOutput = false, int fileNumber = 0)
{
    if (dL == 0) detail::fatal("Select dL>0");
    params.injectedAngularMomentum = dL;
    std::vector<int> sec = {0, dL};
    detail::BasisInfo b = detail::basis(sec);
    detail::capacity(params, b);
    le_observables obs{};
    int harm[1] = {dL};
    obs.n_harmonics = 1;
    obs.harmonics = harm;
    detail::Run run(params, sec, &obs);
    auto nrm = run.norms();
    cmatrix<double> H(b.D, b.D), M(b.D, b.D), VC(b.D, b.D), VS(b.D, b.D);
    for (int s = 0; s < 2; ++s) {
        run.place(H, LE_W_V, s, s, nrm);
        run.place(M, LE_W_ONE, s, s, nrm);
    }
    run.place(VC, LE_W_COS(dL), 0, 1, nrm);
    run.place(VC, LE_W_COS(dL), 1, 0, nrm);
    run.place(VS, LE_W_SIN(dL), 0, 1, nrm);
    run.place(VS, LE_W_SIN(dL), 1, 0, nrm);
    H = detail::self_adjoint(H);
    M = detail::self_adjoint(M);
    VC = detail::self_adjoint(VC);
    VS = detail::self_adjoint(VS);
    if (saveOutput) {
        const std::string base = "dL=" + std::to_string(dL) + "/Statistics/";
        detail::write_tsv(detail::out_dir(base + "Hamiltonian"), fileNumber, H);
        detail::write_tsv(detail::out_dir(base + "Metric"), fileNumber, M);
        detail::write_tsv(detail::out_dir(base + "cosine_excitation"), fileNumber, VC);
        detail::write_tsv(detail::out_dir(base + "sine_excitation"), fileNumber, VS);
    }
    ges::generalizedEigenproblem eig(H, M);
    return eig.extract_dsf_coefficients(VC, VS);
}

#if defined(__HIPCC__)
namespace detail {

// U evaluated per walker on FP32 positions (the functor takes complex<float>*)
template <typename Func>
__global__ void user_weight_kernel(const complex<float>* pos, int walkers, int N, void* w, long ld, int w_double,
                                   Func U)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= walkers) return;
    complex<float>* z = const_cast<complex<float>*>(pos) + (size_t)s * N;
    const double u = U(z);
    if (w_double) reinterpret_cast<double*>(w)[s] = u;
    else reinterpret_cast<float*>(w)[s] = (float)u;
}

static __global__ void to_float_positions(const double* in, float* out, long n)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) out[g] = (float)in[g];
}

template <typename Func>
struct UserWeight {
    Func U;
    float* scratch = nullptr;   // FP32 copy when the engine hands FP64 positions
    long scratch_n = 0;
    static int call(void* ctx, const void* pos, int pos_double, int walkers, int N, void* w, long ld, int w_double,
                    void* stream)
    {
        auto* self = static_cast<UserWeight*>(ctx);
        hipStream_t st = (hipStream_t)stream;
        const float* p = static_cast<const float*>(pos);
        if (pos_double) {
            const long n = 2L * walkers * N;
            if (n > self->scratch_n) {
                if (self->scratch) (void)hipFree(self->scratch);
                if (hipMalloc(&self->scratch, n * sizeof(float)) != hipSuccess) return 1;
                self->scratch_n = n;
            }
            hipLaunchKernelGGL(to_float_positions, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                               static_cast<const double*>(pos), self->scratch, n);
            p = self->scratch;
        }
        hipLaunchKernelGGL((user_weight_kernel<Func>), dim3((walkers + 255) / 256), dim3(256), 0, st,
                           reinterpret_cast<const complex<float>*>(p), walkers, N, w, ld, w_double, self->U);
        return hipGetLastError() == hipSuccess ? 0 : 1;
    }
    ~UserWeight()
    {
        if (scratch) (void)hipFree(scratch);
    }
};

}  // namespace detail

// Response matrices of an external potential U(z_1..z_N) and the density Fourier
// components rho_J = sum_i e^{i L_J theta_i} (L_J = the requested sectors), written as
// TSV for timeEvolution::edgeDensityResponse_compute (em_edgeResponse.tpp:559-594).
template <typename Func>
void excitationResponse_compute(monteCarloParameters& params, vector<int>& angular_momentum_sectors,
                                Func excitation_potential, int fileNumber)
{
    static_assert(std::is_invocable_r_v<double, Func, complex<float>*>,
                  "The excitation potential must be a __device__ functor taking complex<float>* and returning double");
    std::vector<int> sec(angular_momentum_sectors.size());
    for (int q = 0; q < angular_momentum_sectors.size(); ++q) sec[q] = angular_momentum_sectors[q];
    detail::BasisInfo b = detail::basis(sec);
    detail::capacity(params, b);
    if ((int)sec.size() > sys_params::ansatzMultipletSize)
        detail::fatal("\nYou chose " + std::to_string(sec.size()) + " different angular momenta.\nThe upper bound in "
                      "sys_params.h is " + std::to_string(sys_params::ansatzMultipletSize) + "; regenerate sys_params.h");
    params.numberOfSectors = (int)sec.size();
    le_observables obs{};
    obs.n_harmonics = (int)sec.size();
    obs.harmonics = sec.data();
    obs.n_user_weights = 1;
    detail::UserWeight<Func> uw{excitation_potential};
    detail::Run run(params, sec, &obs, &detail::UserWeight<Func>::call, &uw);
    auto nrm = run.norms();
    const int D = b.D, ns = (int)sec.size();
    cmatrix<double> H(D, D), M(D, D), U(D, D);
    std::vector<cmatrix<double>> RC(ns, cmatrix<double>(D, D)), RS(ns, cmatrix<double>(D, D));
    for (int s1 = 0; s1 < ns; ++s1)
        for (int s2 = 0; s2 < ns; ++s2) {
            if (run.dims[s1] == 0 || run.dims[s2] == 0) continue;
            if (sec[s1] == sec[s2]) {
                run.place(H, LE_W_V, s1, s2, nrm);
                run.place(M, LE_W_ONE, s1, s2, nrm);
            }
            run.place(U, LE_W_USER(0), s1, s2, nrm);
            for (int J = 0; J < ns; ++J)
                if (std::abs(sec[s2] - sec[s1]) == sec[J]) {
                    run.place(RC[J], LE_W_COS(sec[J]), s1, s2, nrm);
                    run.place(RS[J], LE_W_SIN(sec[J]), s1, s2, nrm);
                }
        }
    H = detail::self_adjoint(H);
    M = detail::self_adjoint(M);
    U = detail::self_adjoint(U);
    const std::string base = "ExternalExcitation/Statistics/";
    detail::write_tsv(detail::out_dir(base + "Hamiltonian"), fileNumber, H);
    detail::write_tsv(detail::out_dir(base + "Metric"), fileNumber, M);
    detail::write_tsv(detail::out_dir(base + "ExcitationPotential"), fileNumber, U);
    for (int J = 0; J < ns; ++J) {
        const std::string dir = detail::out_dir(base + "EdgeDenstiy_FourierTransform/dL=" + std::to_string(sec[J]));
        std::filesystem::create_directories(dir);
        std::ofstream out(dir + "/" + std::to_string(fileNumber) + ".tsv");
        char buf[192];
        for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j) {
                std::snprintf(buf, sizeof buf, "%d\t%d\t%.6f\t%.6f\t%.6f\t%.6f\n", i, j, re(RC[J](i, j)),
                              im(RC[J](i, j)), re(RS[J](i, j)), im(RS[J](i, j)));
                out << buf;
            }
    }
}
#endif

}  // namespace edgeMC

#endif
