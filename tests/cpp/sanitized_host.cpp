// Host code under AddressSanitizer and UndefinedBehaviorSanitizer (built by
// tests/cpp/build_sanitized.sh with g++/gcc only, no GPU code): the FP64 oracle (chains, both
// estimators, eigensolver, DSF weights, integrator), the basis enumerator, the step search
// and the TSV reader of the time-evolution shell, on small inputs and on their error paths.
// Prints "SANITIZED OK" when every check passes; a sanitizer report aborts the run.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../laughlinedge_amd/csrc/le_basis.hpp"
#include "dynamics.h"

extern "C" {
typedef struct {
    int N, m;
    double w, eps, strength;
    int expo;
    double R;
} or_system;
typedef struct {
    int D, K;
    const int *fo, *fk, *fr;
    int nsec;
    const int* so;
} or_basis;
typedef double (*or_time_fn)(void* ctx, double t);
void or_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]);
int or_block_size(const or_basis* b);
int or_accumulate_fixed(const or_system* sys, const or_basis* b, const double* pos, long n, double* S, double* H,
                        double* Vsum);
int or_accumulate_weighted(const or_system* sys, const or_basis* b, const double* pos, long n, const double* w, int s1,
                           int s2, double* out);
int or_plain_run(const or_system* sys, const or_basis* b, int walkers, int burnin, int samples, int thin, double step,
                 uint64_t seed, int threads, double* S, double* H, double* acceptance, double* final_pos,
                 double* r2_mean);
int or_plain_positions(const or_system* sys, int walkers, int burnin, int samples, int thin, double step,
                       uint64_t seed, double* out, double* acceptance);
int or_rowwise_sector(const or_system* sys, const or_basis* b, int sec, int walkers, int burnin, int samples,
                      double step, uint64_t seed, int threads, int dl_exc, double* Hout, double* Mout, double* VCout,
                      double* VSout, double* acceptance, double* final_pos);
int or_geneig(int d, const double* Sin, const double* Hin, double* evals, double* evecs);
int or_dsf_weights(int dp1, const double* H, const double* M, const double* VC, const double* VS, double* energies,
                   double* weights);
int or_evolve(int D, const double* Min, const double* Hin, const double* Uin, or_time_fn F, void* ctx,
              double end_time, double save_step, double eps_abs, double eps_rel, int nrho, const double* rhoin,
              double* out_rho, double* out_norm, long* n_steps);
int le_step_search(int (*measure)(void*, double, double*), void* ctx, double step0, double target, double tol,
                   double* step_out, int* n_measured);
}

namespace le {
std::string g_msg;
void set_error(const std::string& msg) { g_msg = msg; }
}  // namespace le

static int fails = 0;
#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            std::printf("FAILED %s:%d %s\n", __FILE__, __LINE__, #c);  \
            ++fails;                                                    \
        }                                                               \
    } while (0)

static double flat_drive(void*, double t) { return t < 0.5 ? 0.1 : 0.0; }

struct Curve {
    double scale;      // acceptance = 1 / (1 + (step / scale)^2)
    double floor;      // acceptance never below this (a non-bracketing curve when > target)
};
static int curve_measure(void* ctx, double step, double* a)
{
    const Curve* c = static_cast<const Curve*>(ctx);
    const double x = step / c->scale;
    *a = std::fmax(c->floor, 1.0 / (1.0 + x * x));
    return 0;
}

int main()
{
    // Philox known answer (Random123 kat_vectors: philox4x32_10, counter 0, key 0)
    {
        uint32_t ctr[4] = {0, 0, 0, 0}, key[2] = {0, 0}, out[4];
        or_philox4x32_10(ctr, key, out);
        CHECK(out[0] == 0x6627e8d5u && out[1] == 0xe169c58du && out[2] == 0xbc57ac4cu && out[3] == 0x9b00dbd8u);
    }
    // basis (N = 6, dL 0..4) and the oracle's factor arrays built from it
    le::Basis basis;
    basis.build(6, {0, 1, 2, 3, 4});
    CHECK(basis.D == 1 + 1 + 2 + 3 + 5);
    std::vector<int> fo{0}, fk, fr, so(basis.sector_offset.begin(), basis.sector_offset.end());
    for (const auto& st : basis.states) {
        for (const auto& p : st.pairs) {
            fk.push_back(p.number);
            fr.push_back(p.repetitions);
        }
        fo.push_back((int)fk.size());
    }
    or_basis b{basis.D, basis.K, fo.data(), fk.data(), fr.data(), basis.nsec(), so.data()};
    or_system sys{6, 3, 0.05, 1e-4, 0.0, 0, std::sqrt(36.0)};
    le::Basis big;
    big.build(200, std::vector<int>{1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16});
    CHECK(big.D == 914 && big.factor_table().size() == (size_t)914 * le::kMaxFactors);

    const int walkers = 4, samples = 3, nb = or_block_size(&b);
    std::vector<double> pos((size_t)walkers * samples * 6 * 2);
    double acc = 0;
    CHECK(or_plain_positions(&sys, walkers, 5, samples, 1, 0.8, 7, pos.data(), &acc) == 0);
    CHECK(acc > 0 && acc < 1);
    const long n = walkers * samples;
    std::vector<double> S(2 * (size_t)nb), H(2 * (size_t)nb), S2(2 * (size_t)nb), H2(2 * (size_t)nb);
    double vsum = 0;
    CHECK(or_accumulate_fixed(&sys, &b, pos.data(), n, S.data(), H.data(), &vsum) == 0);
    std::vector<double> wts(n, 1.0), blk(2 * 5 * 3);
    CHECK(or_accumulate_weighted(&sys, &b, pos.data(), n, wts.data(), 4, 3, blk.data()) == 0);
    double r2 = 0;
    std::vector<double> fin((size_t)walkers * 6 * 2);
    CHECK(or_plain_run(&sys, &b, walkers, 5, samples, 1, 0.8, 7, 1, S2.data(), H2.data(), &acc, fin.data(), &r2) == 0);
    // identical chains: the plain run's sums equal the fixed accumulation of its positions
    double dmax = 0;
    for (int q = 0; q < 2 * nb; ++q) dmax = std::fmax(dmax, std::fabs(S[q] - S2[q]));
    CHECK(dmax < 1e-9);
    // row-wise estimator of sector dL = 2 (2 rows) with its DSF sums
    std::vector<double> Hr(8), Mr(8), VC(8), VS(8);
    std::vector<double> rfin((size_t)2 * 2 * 6 * 2);
    CHECK(or_rowwise_sector(&sys, &b, 2, 2, 3, 4, 0.8, 11, 1, 2, Hr.data(), Mr.data(), VC.data(), VS.data(), &acc,
                            rfin.data()) == 0);
    CHECK(std::fabs(Mr[0] - 1.0) < 1e-12 && std::fabs(Mr[6] - 1.0) < 1e-12);
    // eigensolves: sector dL = 4 block, then an indefinite metric (error path)
    const long off4 = 2 * (1 + 1 + 4 + 9);
    std::vector<double> ev(5), vec(2 * 25);
    const int rc4 = or_geneig(5, S.data() + off4, H.data() + off4, ev.data(), vec.data());
    CHECK(rc4 == 0 || rc4 == -1);
    std::vector<double> Sbad = {1, 0, 2, 0, 2, 0, 1, 0}, Hbad = {1, 0, 0, 0, 0, 0, 1, 0}, ev2(2);
    CHECK(or_geneig(2, Sbad.data(), Hbad.data(), ev2.data(), nullptr) == -1);
    std::vector<double> Md = {1, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0.1, 0, 0, 0, 0.1, 0, 1, 0};
    std::vector<double> Hd = {1, 0, 0, 0, 0, 0, 0, 0, 2, 0, 0.2, 0, 0, 0, 0.2, 0, 3, 0};
    std::vector<double> e3(2), w3(2);
    CHECK(or_dsf_weights(3, Hd.data(), Md.data(), Md.data(), Hd.data(), e3.data(), w3.data()) == 0);
    CHECK(or_dsf_weights(1, Hd.data(), Md.data(), Md.data(), Hd.data(), e3.data(), w3.data()) == -1);
    // integrator: 2 x 2, free + kicked, then a singular metric (error path)
    std::vector<double> M2 = {1, 0, 0, 0, 0, 0, 1, 0}, H2m = {0, 0, 0, 0, 0, 0, 1, 0}, U2 = {0, 0, 1, 0, 1, 0, 0, 0};
    std::vector<double> rho(2 * 5), nrm(5);
    long steps = 0;
    CHECK(or_evolve(2, M2.data(), H2m.data(), U2.data(), flat_drive, nullptr, 1.0, 0.2, 1e-10, 1e-8, 1, M2.data(),
                    rho.data(), nrm.data(), &steps) == 0);
    CHECK(std::fabs(nrm[4] - 1.0) < 1e-8 && steps > 0);
    std::vector<double> Msing(8, 0.0);
    CHECK(or_evolve(2, Msing.data(), H2m.data(), U2.data(), flat_drive, nullptr, 1.0, 0.2, 1e-10, 1e-8, 1, M2.data(),
                    rho.data(), nrm.data(), &steps) == -1);
    // step search: converging, non-bracketing and argument errors
    Curve good{1.0, 0.0}, flat{1.0, 0.7};
    double step = 0;
    int nm = 0;
    CHECK(le_step_search(curve_measure, &good, 0.3, 0.5, 0.01, &step, &nm) == 0);
    CHECK(std::fabs(1.0 / (1.0 + step * step) - 0.5) < 0.01 && nm > 1);
    CHECK(le_step_search(curve_measure, &flat, 0.3, 0.5, 0.01, &step, &nm) == 5 && nm == 42);
    CHECK(le_step_search(curve_measure, &good, -1.0, 0.5, 0.01, &step, &nm) == 1);
    // TSV reader: valid, malformed, negative index, missing
    const std::string dir = "sanitized_tsv";
    std::filesystem::create_directories(dir);
    std::ofstream(dir + "/ok.tsv") << "0\t0\t1.0\t0.0\n1\t1\t2.0\t0.5\n0\t1\t0.1\t-0.1\n";
    std::ofstream(dir + "/bad.tsv") << "0\t0\t1.0\tx\n";
    std::ofstream(dir + "/neg.tsv") << "0\t-3\t1.0\t0.0\n";
    int dim = 0;
    auto mat = timeEvolution::detail::read_tsv(dir + "/ok.tsv", 4, dim);
    CHECK(dim == 2 && std::abs(mat[3] - std::complex<double>(2.0, 0.5)) < 1e-15);
    for (const char* f : {"bad.tsv", "neg.tsv", "missing.tsv"}) {
        bool threw = false;
        try {
            timeEvolution::detail::read_tsv(dir + "/" + f, 4, dim);
        } catch (const std::runtime_error&) {
            threw = true;
        }
        CHECK(threw);
    }
    std::filesystem::remove_all(dir);
    if (fails) return 1;
    std::printf("SANITIZED OK\n");
    return 0;
}
