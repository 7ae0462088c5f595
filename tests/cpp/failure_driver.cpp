// Forces the reference's failure paths through the drop-in C++ interface, one per run
// (argv[1] names it).  Every path ends before any GPU call: the reference exits (or throws)
// during argument, capacity or file checks, and so does the shell.  Exit code 3 marks a
// std::runtime_error that reached main, with its message on stdout; 0 means no failure.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>

#include "dynamics.h"
#include "edge_montecarlo.h"
#include "green_function.h"

static void put(const std::string& path, const std::string& text)
{
    std::filesystem::create_directories(std::filesystem::path(path).parent_path());
    std::ofstream(path) << text;
}

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    const unsigned N = sys_params::particlesNumber, m = sys_params::inverseFilling;
    const std::string base = "../output/N=" + std::to_string(N) + "_m=" + std::to_string(m) +
                             "/ExternalExcitation/Statistics/";
    auto F = [](double) { return 0.0; };
    try {
        if (mode == "capacity") {
            // dL = 13: largest degree 13 > ansatzExtendedMaxDegree - 1 = 12 (em_spectrum.cu:235-240)
            edgeMC::monteCarloParameters p;
            edgeMC::spectrum_compute(p, 13);
        } else if (mode == "dsf0") {
            // em_dsf.cu:241-244
            edgeMC::monteCarloParameters p;
            edgeMC::dsf_compute(p, 0);
        } else if (mode == "occupation_dl") {
            // dL beyond ansatzMaxAngularMomentum (gf_occupation_numbers.cu:316-319)
            greenFunctionMC::monteCarloParameters p;
            greenFunctionMC::occupation_numbers_compute(p, sys_params::ansatzMaxAngularMomentum + 1);
        } else if (mode == "tsv_missing") {
            // dynamics.cpp:164-166
            std::filesystem::remove_all("../output");
            timeEvolution::edgeDensityResponse_compute(N, m, {0, 1}, F, 1.0, 0.1, 0);
        } else if (mode == "tsv_malformed" || mode == "tsv_negative" || mode == "sectors0") {
            // dynamics.cpp:182-184 (parsing), dynamics.cpp:211-214 (first sector not dL = 0)
            std::filesystem::remove_all("../output");
            const std::string one = "0\t0\t1.000000\t0.000000\n1\t1\t1.000000\t0.000000\n";
            put(base + "Metric/0.tsv", mode == "tsv_malformed" ? "0\t0\t1.0\tabc\n" : one);
            put(base + "Hamiltonian/0.tsv", mode == "tsv_negative" ? "0\t-1\t1.0\t0.0\n" : one);
            put(base + "ExcitationPotential/0.tsv", one);
            const std::vector<int> sec = mode == "sectors0" ? std::vector<int>{1, 2} : std::vector<int>{0, 1};
            timeEvolution::edgeDensityResponse_compute(N, m, sec, F, 1.0, 0.1, 0);
        } else {
            return 2;
        }
    } catch (const std::runtime_error& ex) {
        std::printf("runtime_error: %s\n", ex.what());
        return 3;
    }
    return 0;
}
