// Drop-in timeEvolution::edgeDensityResponse_compute over the FP64 GPU integrator (le_evolve).
//
// Same signature and files as the reference (modules/dynamics.h:75-79,
// modules/dynamics/dynamics.cpp:149-309): reads the Metric, Hamiltonian and
// ExcitationPotential TSVs and the EdgeDenstiy_FourierTransform/dL=L files written by
// edgeMC::excitationResponse_compute (fileNumber.tsv under
// ../output/N={N}_m={m}/ExternalExcitation/Statistics), starts from the Laughlin state
// C(0) = e_0 (the first sector must be dL = 0, else message + exit(1)), integrates
// M dC/dt = -i (H + F(t) U) C with adaptive Cash-Karp 5(4) (abs 1e-10, rel 1e-8), and
// writes rho_L(t) = C^H rho_L C as "t\tre\tim" to edgeDynamics/dL=L/fileNumber.tsv.
// Unreadable or malformed files throw std::runtime_error, as the reference does.
#ifndef LE_DYNAMICS_H
#define LE_DYNAMICS_H

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iostream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "laughlinedge.h"

namespace timeEvolution {

namespace detail {

inline std::string base_dir(unsigned N, unsigned m)
{
    return "../output/N=" + std::to_string(N) + "_m=" + std::to_string(m) + "/ExternalExcitation/Statistics/";
}

// "i j re im [re2 im2]" rows -> dense complex matrix (two for six columns: first + i second)
inline std::vector<std::complex<double>> read_tsv(const std::string& path, int cols, int& n)
{
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Cannot open file: " + path);
    std::vector<std::tuple<int, int, std::complex<double>>> entries;
    int i, j, mi = -1, mj = -1;
    double a, b, c = 0, d = 0;
    while (true) {
        if (!(in >> i >> j >> a >> b)) break;
        if (cols == 6 && !(in >> c >> d)) break;
        // a negative index would write outside the matrix (the reference's Eigen indexing
        // has no defined result for it): reject it as a parsing error
        if (i < 0 || j < 0) throw std::runtime_error("Parsing error in " + path + ": negative index");
        std::complex<double> v(a, b);
        if (cols == 6) v = std::complex<double>(a, b) + std::complex<double>(0.0, 1.0) * std::complex<double>(c, d);
        entries.emplace_back(i, j, v);
        if (i > mi) mi = i;
        if (j > mj) mj = j;
    }
    if (!in.eof()) throw std::runtime_error("Parsing error in " + path);
    n = std::max(mi, mj) + 1;
    std::vector<std::complex<double>> mat((size_t)n * n, 0.0);
    for (auto& e : entries) mat[(size_t)std::get<0>(e) * n + std::get<1>(e)] = std::get<2>(e);
    return mat;
}

inline double call_F(void* ctx, double t) { return (*static_cast<std::function<double(double)>*>(ctx))(t); }

}  // namespace detail

inline void edgeDensityResponse_compute(unsigned particlesNumber, unsigned inverseFilling, std::vector<int> angular_momenta,
                                        std::function<double(double)> temporal_excitation_profile, double end_time,
                                        double save_step, int fileNumber = 0)
{
    const std::string base = detail::base_dir(particlesNumber, inverseFilling);
    const std::string file = std::to_string(fileNumber) + ".tsv";
    int nM = 0, nH = 0, nU = 0;
    auto M = detail::read_tsv(base + "Metric/" + file, 4, nM);
    auto H = detail::read_tsv(base + "Hamiltonian/" + file, 4, nH);
    auto U = detail::read_tsv(base + "ExcitationPotential/" + file, 4, nU);
    const int D = nH;
    std::cout << "Time-evolution in the [ ";
    for (int l : angular_momenta) std::cout << l << " ";
    std::cout << "] angular momentum subspaces. Number of edge states: " << D << "\n";
    if (angular_momenta.empty() || angular_momenta[0] != 0) {
        std::cout << "The Laughlin state (assumed to be the initial state) is not included in the subspace!" << std::endl;
        std::exit(1);
    }
    if (nM != D || nU != D) throw std::runtime_error("Metric, Hamiltonian and ExcitationPotential sizes differ");
    std::vector<std::vector<std::complex<double>>> rho;
    for (int l : angular_momenta) {
        int nr = 0;
        rho.push_back(detail::read_tsv(base + "EdgeDenstiy_FourierTransform/dL=" + std::to_string(l) + "/" + file, 6, nr));
        if (nr != D) throw std::runtime_error("edge density matrix size differs from the Hamiltonian's");
    }
    // angular-momentum blocks of M and H: partition counts of the sectors (<= N parts)
    int Dcheck = 0, K = 0;
    std::vector<int> dims;
    for (int l : angular_momenta) {
        int d = 0;
        if (le_basis_describe((int)particlesNumber, &l, 1, 0, &d, &K, nullptr, nullptr, nullptr) != LE_OK)
            throw std::runtime_error(le_last_error());
        dims.push_back(d);
        Dcheck += d;
    }
    if (Dcheck != D) throw std::runtime_error("matrix size does not match the sectors' edge-space dimension");
    const long nt = (long)std::ceil(end_time / save_step);
    const int nr = (int)rho.size();
    std::vector<double> Rcat;
    for (auto& r : rho)
        for (auto& v : r) {
            Rcat.push_back(v.real());
            Rcat.push_back(v.imag());
        }
    auto flat = [](const std::vector<std::complex<double>>& a) {
        std::vector<double> o;
        o.reserve(2 * a.size());
        for (auto& v : a) {
            o.push_back(v.real());
            o.push_back(v.imag());
        }
        return o;
    };
    std::vector<double> Mf = flat(M), Hf = flat(H), Uf = flat(U);
    std::vector<double> out(2 * (size_t)nt * nr), norm(nt);
    long steps = 0;
    std::cout << "\tIntegrating on the GPU (FP64)...\n";
    const int rc = le_evolve(D, (int)dims.size(), dims.data(), Mf.data(), Hf.data(), 1, Uf.data(), nr, Rcat.data(),
                             &detail::call_F, &temporal_excitation_profile, end_time, save_step, 1e-10, 1e-8,
                             out.data(), norm.data(), &steps);
    if (rc != LE_OK) {
        std::cout << "\nlaughlinedge: " << le_last_error() << std::endl;
        std::exit(1);
    }
    double drift = 0;
    for (long k = 0; k < nt; ++k) drift = std::max(drift, std::fabs(norm[k] - norm[0]));
    std::printf("\tDone (%ld controller steps, max |C^H M C - C0^H M C0| = %.3e). Saving...\n", steps, drift);
    for (int l = 0; l < nr; ++l) {
        const std::string dir = base + "edgeDynamics/dL=" + std::to_string(angular_momenta[l]);
        std::filesystem::create_directories(dir);
        std::ofstream o(dir + "/" + file);
        char buf[128];
        for (long k = 0; k < nt; ++k) {
            std::snprintf(buf, sizeof buf, "%.6f\t%.6f\t%.6f\n", k * save_step, out[2 * ((size_t)k * nr + l)],
                          out[2 * ((size_t)k * nr + l) + 1]);
            o << buf;
        }
    }
}

}  // namespace timeEvolution

#endif
