// Generator of include/sys_params.h: compile-time system constants for the drop-in C++
// interface (the reference generates modules/sys_params.h the same way,
// system_parameters.cpp:9-68).  The names of the reference header are kept; the
// confinement gains the 0.5 w r^2 + eps r^4 form.  The engine itself is sized at run
// time; the ansatz* bounds are kept only for the capacity checks the reference performs.
//
//   g++ -O2 -o /tmp/sysgen laughlinedge_amd/system_parameters.cpp && /tmp/sysgen include/sys_params.h \
//       N=200 m=3 w=0.05 eps=5e-5
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <string>

int main(int argc, char** argv)
{
    const char* out_path = argc > 1 ? argv[1] : "include/sys_params.h";
    int N = 25, m = 2, exponent = 4, max_degree = 12, max_partition = 5, subspace = 80, multiplet = 5,
        max_dl = 15;
    double strength = 1.0, w = 0.0, eps = 0.0;
    for (int a = 2; a < argc; ++a) {
        std::string kv = argv[a];
        auto eq = kv.find('=');
        if (eq == std::string::npos) continue;
        std::string k = kv.substr(0, eq), v = kv.substr(eq + 1);
        if (k == "N") N = std::atoi(v.c_str());
        else if (k == "m") m = std::atoi(v.c_str());
        else if (k == "exponent") exponent = std::atoi(v.c_str());
        else if (k == "strength") strength = std::atof(v.c_str());
        else if (k == "w") w = std::atof(v.c_str());
        else if (k == "eps") eps = std::atof(v.c_str());
        else if (k == "max_degree") max_degree = std::atoi(v.c_str());
        else if (k == "max_partition") max_partition = std::atoi(v.c_str());
        else if (k == "subspace") subspace = std::atoi(v.c_str());
        else if (k == "multiplet") multiplet = std::atoi(v.c_str());
        else if (k == "max_dl") max_dl = std::atoi(v.c_str());
        else {
            std::fprintf(stderr, "unknown key %s\n", k.c_str());
            return 1;
        }
    }
    const double rcl = std::sqrt(2.0 * m * N);
    std::ofstream o(out_path);
    o << std::fixed << std::setprecision(6);
    o << "// Generated by laughlinedge_amd/system_parameters.cpp\n";
    o << "#ifndef SYSPARAMS_H\n#define SYSPARAMS_H\n\nnamespace sys_params {\n";
    o << "    constexpr int particlesNumber = " << N << ";\n";
    o << "    constexpr int inverseFilling = " << m << ";\n\n";
    o << "    constexpr int confinementExponent = " << exponent << ";\n";
    o << "    constexpr float confinementStrength = " << strength << ";\n";
    o << std::setprecision(12);
    o << "    constexpr double confinementW = " << w << ";\n";
    o << "    constexpr double confinementEps = " << eps << ";\n\n";
    o << std::setprecision(6);
    o << "    constexpr float RCl = " << rcl << ";\n";
    o << "    constexpr float DCl = " << 2.0 * rcl << ";\n";
    o << "    constexpr float RCl_reciprocal = " << 1.0 / rcl << ";\n";
    o << "    constexpr float R0_reciprocal = " << 2.0 / rcl << ";\n\n";
    o << "    constexpr int ansatzExtendedMaxDegree = " << max_degree + 1 << ";\n";
    o << "    constexpr int ansatzMaxPartitionSize = " << max_partition << ";\n";
    o << "    constexpr int ansatzSubspaceDimension = " << subspace << ";\n\n";
    o << "    constexpr int ansatzMultipletSize = " << multiplet << ";\n\n";
    o << "    constexpr int ansatzMaxAngularMomentum = " << max_dl << ";\n";
    o << "    constexpr int angularMomentumCutoff = " << m * (N - 1) + max_dl << ";\n\n";
    o << "}\n\n#endif\n";
    return o.good() ? 0 : 1;
}
