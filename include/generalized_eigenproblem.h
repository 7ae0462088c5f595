// ges::generalizedEigenproblem over the engine's FP64 GPU eigensolver (le_geneig).
//
// Same class, constructor and methods as the reference's
// modules/generalized_eigenproblem.h:14-31: extract_spectrum() returns the ascending
// eigenvalues of H v = E M v (generalized_eigenproblem.cpp:45-52);
// extract_dsf_coefficients(VC, VS) drops row/column 0 (the Laughlin state), solves the
// remaining block with M-normalised eigenvectors and returns (E_j, |(VC - i VS)[0,1:] v_j|^2)
// (generalized_eigenproblem.cpp:54-95).  A metric that is not positive definite is
// reported on stdout and ends the process, as an unrecoverable input error.
#ifndef LE_GENERALIZED_EIGENPROBLEM_H
#define LE_GENERALIZED_EIGENPROBLEM_H

#include <cstdio>
#include <cstdlib>
#include <utility>
#include <vector>

#include "complex_numbers.h"
#include "laughlinedge.h"
#include "tensors.h"

namespace ges {

class generalizedEigenproblem {
    cmatrix<double> H0_, M0_;

    static std::vector<double> pack(const cmatrix<double>& A, int off)
    {
        const int d = A.rows() - off;
        std::vector<double> out(2 * (size_t)d * d);
        for (int i = 0; i < d; ++i)
            for (int j = 0; j < d; ++j) {
                out[2 * ((size_t)i * d + j)] = re(A(i + off, j + off));
                out[2 * ((size_t)i * d + j) + 1] = im(A(i + off, j + off));
            }
        return out;
    }

    static void solve(const cmatrix<double>& H, const cmatrix<double>& M, int off, std::vector<double>& ev,
                      std::vector<double>* vec)
    {
        const int d = H.rows() - off;
        ev.assign(d > 0 ? d : 0, 0.0);
        if (d <= 0) return;
        std::vector<double> S = pack(M, off), Hm = pack(H, off);
        if (vec) vec->assign(2 * (size_t)d * d, 0.0);
        const int rc = le_geneig(d, S.data(), Hm.data(), ev.data(), vec ? vec->data() : nullptr);
        if (rc != LE_OK) {
            std::printf("\nGeneralized eigenproblem failed: %s\n", le_last_error());
            std::exit(1);
        }
    }

public:
    generalizedEigenproblem(cmatrix<double>& H, cmatrix<double>& M) : H0_(H), M0_(M) {}

    std::vector<double> extract_spectrum()
    {
        std::vector<double> ev;
        solve(H0_, M0_, 0, ev, nullptr);
        return ev;
    }

    std::vector<std::pair<double, double>> extract_dsf_coefficients(cmatrix<double>& VC, cmatrix<double>& VS)
    {
        std::vector<double> ev, vec;
        solve(H0_, M0_, 1, ev, &vec);
        const int d = (int)ev.size();
        std::vector<std::pair<double, double>> out(d);
        for (int j = 0; j < d; ++j) {
            double sr = 0, si = 0;
            for (int a = 0; a < d; ++a) {
                // (VC - i VS)[0][1+a] * v_j[a]
                const complex<double> v = VC(0, 1 + a) - complex<double>(0.0, 1.0) * VS(0, 1 + a);
                const complex<double> x(vec[2 * ((size_t)a * d + j)], vec[2 * ((size_t)a * d + j) + 1]);
                const complex<double> p = v * x;
                sr += re(p);
                si += im(p);
            }
            out[j] = {ev[j], sr * sr + si * si};
        }
        return out;
    }
};

}  // namespace ges

#endif
