// FP64 time evolution of the edge-subspace coefficient vector on the GPU.
//
// Reference (modules/dynamics/dynamics.cpp:20-53, 264-290): dC/dt = -i M^-1 (H + F(t) U) C,
// C(0) = e_0, adaptive Cash-Karp 5(4) with abs 1e-10 / rel 1e-8, observed at t_k = k dt:
// rho_l(t) = C^H rho_l C, norm C^H M C.
//
// Here: the generalized eigenvectors V of (H, M) (H V = M V Lambda, V^H M V = 1, computed
// per angular-momentum block) turn the system into a' = -i (Lambda + F W) a with
// W = V^H U V, C = V a.  The fast phases e^{-i Lambda t} are removed exactly
// (a = e^{-i Lambda t} b), leaving b' = -i F(t) e^{i Lambda t} W e^{-i Lambda t} b, which is
// slow and is integrated with the same adaptive Cash-Karp 5(4) controller.  Observables
// are rho_l(t) = a^H (V^H rho_l V) a and the norm a^H a.  Perturbations U_u are batched:
// every kernel runs over all u at once.
#include <hip/hip_runtime.h>

#include "le_evolve.hpp"

namespace le {

struct z2 {
    double r, i;
};

__device__ __forceinline__ z2 zmul(z2 a, z2 b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ z2 zcmul(z2 a, z2 b) { return {a.r * b.r + a.i * b.i, a.r * b.i - a.i * b.r}; }  // conj(a) b

// C[M x N] = op(A) B, op(A) = A or A^H, complex FP64, row-major, 16x16 tiles through LDS
__global__ __launch_bounds__(256) void k_zgemm(const double2* A, const double2* B, double2* C, int Mr, int Nc, int K,
                                               int lda, int ldb, int ldc, int conjA)
{
    __shared__ double2 sa[16][17], sb[16][17];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int row = blockIdx.y * 16 + ty, col = blockIdx.x * 16 + tx;
    double ar = 0, ai = 0;
    for (int k0 = 0; k0 < K; k0 += 16) {
        // op(A)[row][k] = conjA ? conj(A[k][row]) : A[row][k]
        const int ka = k0 + tx;
        double2 va = {0, 0};
        if (conjA) {
            const int r2 = blockIdx.y * 16 + tx, k2 = k0 + ty;   // load A[k2][r2] coalesced over r2
            if (r2 < Mr && k2 < K) {
                va = A[(size_t)k2 * lda + r2];
                va.y = -va.y;
            }
            sa[tx][ty] = va;   // sa[row-in-tile][k-in-tile]
        } else {
            if (row < Mr && ka < K) va = A[(size_t)row * lda + ka];
            sa[ty][tx] = va;
        }
        double2 vb = {0, 0};
        const int kb = k0 + ty;
        if (kb < K && col < Nc) vb = B[(size_t)kb * ldb + col];
        sb[ty][tx] = vb;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const double2 x = sa[ty][k], y = sb[k][tx];
            ar += x.x * y.x - x.y * y.y;
            ai += x.x * y.y + x.y * y.x;
        }
        __syncthreads();
    }
    if (row < Mr && col < Nc) C[(size_t)row * ldc + col] = {ar, ai};
}

// y_u[m] = e^{-i lam_m t} (b_u[m] + h sum_r a_r k_r,u[m])
__global__ void k_stage_prep(const double2* b, const double2* ks, int nstage, const double* coef, double h, double t,
                             const double* lam, double2* y, int D, int nu)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)D * nu) return;
    const int m = (int)(g % D);
    double2 v = b[g];
    for (int r = 0; r < nstage; ++r) {
        const double2 k = ks[(size_t)r * D * nu + g];
        v.x += h * coef[r] * k.x;
        v.y += h * coef[r] * k.y;
    }
    double s, c;
    sincos(lam[m] * t, &s, &c);
    y[g] = {v.x * c + v.y * s, v.y * c - v.x * s};   // times e^{-i lam t}
}

// k_u[j] = -i F e^{i lam_j t} sum_m W_u[j][m] y_u[m]; one wavefront per (u, j)
__global__ __launch_bounds__(256) void k_stage_matvec(const double2* W, const double2* y, double2* kout, double F,
                                                      double t, const double* lam, int D, int nu)
{
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wave >= D * nu) return;
    const int u = wave / D, j = wave % D;
    const double2* wr = W + ((size_t)u * D + j) * D;
    const double2* yu = y + (size_t)u * D;
    double sr = 0, si = 0;
    for (int m = lane; m < D; m += 64) {
        const double2 a = wr[m], x = yu[m];
        sr += a.x * x.x - a.y * x.y;
        si += a.x * x.y + a.y * x.x;
    }
    for (int off = 32; off > 0; off >>= 1) {
        sr += __shfl_xor(sr, off, 64);
        si += __shfl_xor(si, off, 64);
    }
    if (lane == 0) {
        double s, c;
        sincos(lam[j] * t, &s, &c);
        const double pr = sr * c - si * s, pi = sr * s + si * c;   // times e^{i lam t}
        kout[(size_t)u * D + j] = {F * pi, -F * pr};              // times -i F
    }
}

// 5th-order update, embedded error and the controller's max relative error
__global__ void k_rk_combine(const double2* b, const double2* ks, const double* b5, const double* db, double h,
                             double eps_abs, double eps_rel, double2* bnew, unsigned long long* emax, int n)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double r = 0.0;
    if (g < n) {
        double2 v = b[g];
        double er = 0, ei = 0;
        for (int s = 0; s < 6; ++s) {
            const double2 k = ks[(size_t)s * n + g];
            v.x += h * b5[s] * k.x;
            v.y += h * b5[s] * k.y;
            er += h * db[s] * k.x;
            ei += h * db[s] * k.y;
        }
        bnew[g] = v;
        const double2 b0 = b[g], k0 = ks[g];
        const double den = eps_abs + eps_rel * (sqrt(b0.x * b0.x + b0.y * b0.y) + fabs(h) * sqrt(k0.x * k0.x + k0.y * k0.y));
        r = sqrt(er * er + ei * ei) / den;
    }
    for (int off = 32; off > 0; off >>= 1) r = fmax(r, __shfl_xor(r, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(emax, (unsigned long long)__double_as_longlong(r));  // r >= 0
}

// Observation at time t, in two passes.  k_phase: a_u[m] = e^{-i lam_m t} b_u[m] for every
// perturbation u (one sincos per m) and the norms out[u][nrho] = a_u^H a_u.  k_observe: one
// wavefront per row j of R_l reads that row once and forms (R_l a_u)_j for eight u at a
// time; out[u][l] += conj(a_u[j]) (R_l a_u)_j.
__global__ __launch_bounds__(256) void k_phase(const double2* b, const double* lam, double t, int D, int nu,
                                               int nrho, double2* a, double2* out)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    double s = 0, c = 1;
    if (m < D) sincos(lam[m] * t, &s, &c);
    for (int u = 0; u < nu; ++u) {
        double n2 = 0;
        if (m < D) {
            const double2 bm = b[(size_t)u * D + m];
            const double2 am = {bm.x * c + bm.y * s, bm.y * c - bm.x * s};
            a[(size_t)u * D + m] = am;
            n2 = am.x * am.x + am.y * am.y;
        }
        for (int off = 32; off > 0; off >>= 1) n2 += __shfl_xor(n2, off, 64);
        if (lane == 0) atomicAdd(&out[(size_t)u * (nrho + 1) + nrho].x, n2);
    }
}

__global__ __launch_bounds__(256) void k_observe(const double2* a, const double2* R, int D, int nu, int nrho,
                                                 double2* out)
{
    constexpr int U = 8;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wave >= D * nrho) return;
    const int j = wave % D, l = wave / D;
    const double2* rr = R + ((size_t)l * D + j) * D;
    for (int u0 = 0; u0 < nu; u0 += U) {
        double sr[U], si[U];
#pragma unroll
        for (int k = 0; k < U; ++k) sr[k] = si[k] = 0.0;
        for (int m = lane; m < D; m += 64) {
            const double2 r = rr[m];
#pragma unroll
            for (int k = 0; k < U; ++k) {
                if (u0 + k < nu) {
                    const double2 am = a[(size_t)(u0 + k) * D + m];
                    sr[k] += r.x * am.x - r.y * am.y;
                    si[k] += r.x * am.y + r.y * am.x;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < U; ++k) {
            for (int off = 32; off > 0; off >>= 1) {
                sr[k] += __shfl_xor(sr[k], off, 64);
                si[k] += __shfl_xor(si[k], off, 64);
            }
        }
        if (lane < U && u0 + lane < nu) {
            double xr = 0, xi = 0;
#pragma unroll
            for (int k = 0; k < U; ++k)
                if (k == lane) {
                    xr = sr[k];
                    xi = si[k];
                }
            const int u = u0 + lane;
            const double2 aj = a[(size_t)u * D + j];
            const z2 q = zcmul({aj.x, aj.y}, {xr, xi});
            atomicAdd(&out[(size_t)u * (nrho + 1) + l].x, q.r);
            atomicAdd(&out[(size_t)u * (nrho + 1) + l].y, q.i);
        }
    }
}

hipError_t launch_zgemm(const double* A, const double* B, double* C, int M, int N, int K, int lda, int ldb, int ldc,
                        int conjA, hipStream_t s)
{
    dim3 grid((unsigned)((N + 15) / 16), (unsigned)((M + 15) / 16));
    hipLaunchKernelGGL(k_zgemm, grid, dim3(256), 0, s, (const double2*)A, (const double2*)B, (double2*)C, M, N, K, lda,
                       ldb, ldc, conjA);
    return hipGetLastError();
}

hipError_t launch_stage_prep(const double* b, const double* ks, int nstage, const double* coef, double h, double t,
                             const double* lam, double* y, int D, int nu, hipStream_t s)
{
    const long n = (long)D * nu;
    hipLaunchKernelGGL(k_stage_prep, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (const double2*)b,
                       (const double2*)ks, nstage, coef, h, t, lam, (double2*)y, D, nu);
    return hipGetLastError();
}

hipError_t launch_stage_matvec(const double* W, const double* y, double* kout, double F, double t, const double* lam,
                               int D, int nu, hipStream_t s)
{
    const long waves = (long)D * nu;
    hipLaunchKernelGGL(k_stage_matvec, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, (const double2*)W,
                       (const double2*)y, (double2*)kout, F, t, lam, D, nu);
    return hipGetLastError();
}

hipError_t launch_rk_combine(const double* b, const double* ks, const double* b5, const double* db, double h,
                             double eps_abs, double eps_rel, double* bnew, unsigned long long* emax, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_rk_combine, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (const double2*)b,
                       (const double2*)ks, b5, db, h, eps_abs, eps_rel, (double2*)bnew, emax, n);
    return hipGetLastError();
}

hipError_t launch_observe(const double* b, const double* R, const double* lam, double t, int D, int nu, int nrho,
                          double* a, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_phase, dim3((unsigned)((D + 255) / 256)), dim3(256), 0, s, (const double2*)b, lam, t, D, nu,
                       nrho, (double2*)a, (double2*)out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || nrho == 0) return e;
    const long waves = (long)D * nrho;
    hipLaunchKernelGGL(k_observe, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, (const double2*)a,
                       (const double2*)R, D, nu, nrho, (double2*)out);
    return hipGetLastError();
}

}  // namespace le
