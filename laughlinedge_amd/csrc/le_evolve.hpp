// Launchers of the FP64 time-evolution kernels (le_evolve.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace le {

hipError_t launch_zgemm(const double* A, const double* B, double* C, int M, int N, int K, int lda, int ldb, int ldc,
                        int conjA, hipStream_t s);
// also clears kz, the stage's k that launch_stage_hemv adds into, and *emax when emax is not null
hipError_t launch_stage_prep(const double* b, const double* ks, int nstage, const double* coef, double h, double t,
                             const double* lam, double* y, double* kz, int D, int nu, unsigned long long* emax,
                             hipStream_t s);
hipError_t launch_hermitian_part(double* W, int D, int nu, hipStream_t s);
// kout += -i F e^{i lam t} W y for Hermitian W (its upper 64 x 64 tiles are read)
hipError_t launch_stage_hemv(const double* W, const double* y, double* kout, double F, double t, const double* lam,
                             int D, int nu, hipStream_t s);
hipError_t launch_rk_combine(const double* b, const double* ks, const double* b5, const double* db, double h,
                             double eps_abs, double eps_rel, double* bnew, unsigned long long* emax, int n,
                             hipStream_t s);
// a: scratch of 2 D nu doubles (the phased coefficients); out[u][l], l <= nrho, is added to
hipError_t launch_observe(const double* b, const double* R, const double* lam, double t, int D, int nu, int nrho,
                          double* a, double* out, hipStream_t s);

}  // namespace le
