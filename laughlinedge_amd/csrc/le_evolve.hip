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

// y_u[m] = e^{-i lam_m t} (b_u[m] + h sum_r a_r k_r,u[m]); kz (this stage's k) is cleared for
// the tile products that add into it
__global__ void k_stage_prep(const double2* b, const double2* ks, int nstage, const double* coef, double h, double t,
                             const double* lam, double2* y, double2* kz, int D, int nu, unsigned long long* emax)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (emax && g == 0) *emax = 0ull;   // the step's error maximum, raised by k_rk_combine
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
    kz[g] = {0.0, 0.0};
}

// W <- (W + W^H) / 2 for each of nu D x D matrices
__global__ void k_hermitian_part(double2* W, int D, int nu)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)D * D * nu) return;
    const int u = (int)(g / ((long)D * D));
    const long e = g % ((long)D * D);
    const int i = (int)(e / D), j = (int)(e % D);
    if (i > j) return;
    double2* Wu = W + (size_t)u * D * D;
    const double2 a = Wu[(size_t)i * D + j], b = Wu[(size_t)j * D + i];
    const double2 w = {0.5 * (a.x + b.x), 0.5 * (a.y - b.y)};
    Wu[(size_t)i * D + j] = w;
    Wu[(size_t)j * D + i] = {w.x, -w.y};
}

// A 16-byte load marked nontemporal: W is streamed once per stage and is larger than the
// caches, so it should not evict y and k.
__device__ __forceinline__ double2 load_stream(const double2* p)
{
    typedef double v2d __attribute__((ext_vector_type(2)));
    const v2d v = __builtin_nontemporal_load(reinterpret_cast<const v2d*>(p));
    return double2{v.x, v.y};
}

// k_u += -i F e^{i lam t} W_u y_u for Hermitian W_u, reading only its upper 64 x 64 tiles.
// One workgroup per tile (I, J), I <= J, and perturbation u: it adds W_IJ y_J to the rows of
// tile row I and, for I < J, W_IJ^H y_I to the rows of tile row J, so every off-diagonal
// element is read once for two products (half the bytes of a dense product, which bounds
// this kernel: nu D^2 complex values per stage).  The tile passes through LDS in two 32-row
// halves: both halves' coalesced row loads are issued at once (16 per thread), and the second
// half stays in registers while the first is summed.  The row sums take 32 lanes per half,
// each over 8 columns (row stride 65 elements, so 16 lanes' 16-byte reads hit distinct
// banks); the column sums take lane = column, each wave over 8 rows of each half, kept in
// registers across the halves; the partial sums meet in LDS.  39 KB of LDS per workgroup lets
// four share a CU.
__global__ __launch_bounds__(256) void k_stage_hemv(const double2* W, const double2* y, double2* kout, double F,
                                                    double t, const double* lam, int D, int nt)
{
    __shared__ double2 sW[32][65];
    __shared__ double2 yI[64], yJ[64];
    __shared__ double2 part[256];
    const int u = blockIdx.y;
    int I = 0, rest = (int)blockIdx.x;
    while (rest >= nt - I) {
        rest -= nt - I;
        ++I;
    }
    const int J = I + rest;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const double2* Wu = W + (size_t)u * D * D;
    const int col = J * 64 + lane;
    double2 v0[8], v1[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int row = I * 64 + 4 * i + wv;
        v0[i] = (row < D && col < D) ? load_stream(Wu + (size_t)row * D + col) : double2{0.0, 0.0};
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int row = I * 64 + 32 + 4 * i + wv;
        v1[i] = (row < D && col < D) ? load_stream(Wu + (size_t)row * D + col) : double2{0.0, 0.0};
    }
    // -i F e^{i lam t} in wave 0's registers: of column `col` and of row I * 64 + lane (the
    // row whose sum this lane adds: half lane / 32, row lane % 32 of it)
    double2 phI = {0.0, 0.0}, phJ = {0.0, 0.0};
    if (wv == 0) {
        const int ri = I * 64 + lane;
        double s = 0.0, c = 0.0;
        if (ri < D) sincos(lam[ri] * t, &s, &c);
        phI = {F * s, -F * c};
        s = 0.0;
        c = 0.0;
        if (col < D) sincos(lam[col] * t, &s, &c);
        phJ = {F * s, -F * c};
    } else if (wv < 3) {
        // y of the tile's row and column blocks
        const int j = (wv == 2 ? J : I) * 64 + lane;
        (wv == 2 ? yJ : yI)[lane] = j < D ? y[(size_t)u * D + j] : double2{0.0, 0.0};
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) sW[4 * i + wv][lane] = v0[i];
    __syncthreads();
    double cr = 0.0, ci = 0.0;   // column sum over this wave's 8 rows of each half
    for (int hf = 0; hf < 2; ++hf) {
        const int r0 = I * 64 + 32 * hf;
        if (r0 >= D) break;   // workgroup-uniform; the previous pass ended at a barrier
        {
            const int r = tid & 31, c0 = 8 * (tid >> 5);
            double sr = 0.0, si = 0.0;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const double2 w = sW[r][c0 + k], x = yJ[c0 + k];
                sr += w.x * x.x - w.y * x.y;
                si += w.x * x.y + w.y * x.x;
            }
            part[tid] = {sr, si};
        }
        if (I != J) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int r = 8 * wv + k;
                const double2 w = sW[r][lane], x = yI[32 * hf + r];
                cr += w.x * x.x + w.y * x.y;   // conj(w) y_r
                ci += w.x * x.y - w.y * x.x;
            }
        }
        __syncthreads();   // part[] complete; every read of this half in sW done
        if (wv == 0 && (lane >> 5) == hf) {
            const int row = r0 + (lane & 31);
            if (row < D) {
                double xr = 0.0, xi = 0.0;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    xr += part[32 * q + (lane & 31)].x;
                    xi += part[32 * q + (lane & 31)].y;
                }
                const double2 p = phI;
                atomicAdd(&kout[(size_t)u * D + row].x, p.x * xr - p.y * xi);
                atomicAdd(&kout[(size_t)u * D + row].y, p.x * xi + p.y * xr);
            }
        }
        if (hf == 0) {
#pragma unroll
            for (int i = 0; i < 8; ++i) sW[4 * i + wv][lane] = v1[i];
        }
        __syncthreads();   // the second half visible; part[] free again
    }
    if (I == J) return;   // workgroup-uniform
    part[tid] = {cr, ci};
    __syncthreads();
    if (wv == 0 && col < D) {
        double xr = 0.0, xi = 0.0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            xr += part[64 * w + lane].x;
            xi += part[64 * w + lane].y;
        }
        const double2 p = phJ;
        atomicAdd(&kout[(size_t)u * D + col].x, p.x * xr - p.y * xi);
        atomicAdd(&kout[(size_t)u * D + col].y, p.x * xi + p.y * xr);
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
// perturbation u (one sincos per m) and the norms out[u][nrho] = a_u^H a_u.  k_observe:
// wavefronts loop over (row j, chunk of `chunk` columns, a multiple of 64) tasks of R_l; each
// task reads that part of the row once and forms its part of (R_l a_u)_j for eight u at a time,
// times a_u[j]; a workgroup's sums meet in LDS and are added to out[u][l] atomically.
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
                                                 int chunk, double2* out)
{
    constexpr int U = 8;
    __shared__ double red[4][2 * U];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nchunk = (D + chunk - 1) / chunk;
    const long ntask = (long)D * nchunk;   // (row, chunk) pairs of one R_l
    const long nwave = (long)gridDim.x * 4;
    for (int l = 0; l < nrho; ++l) {
        for (int u0 = 0; u0 < nu; u0 += U) {
            double qr[U], qi[U];   // this wavefront's part of out[u][l], in every lane
#pragma unroll
            for (int k = 0; k < U; ++k) qr[k] = qi[k] = 0.0;
            for (long task = (long)blockIdx.x * 4 + wv; task < ntask; task += nwave) {
                const int part = (int)(task % nchunk), j = (int)(task / nchunk);
                const int m1 = min(D, (part + 1) * chunk);
                const double2* rr = R + ((size_t)l * D + j) * D;
                double sr[U], si[U];
#pragma unroll
                for (int k = 0; k < U; ++k) sr[k] = si[k] = 0.0;
                for (int m = part * chunk + lane; m < m1; m += 64) {
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
                    if (u0 + k < nu) {
                        const double2 aj = a[(size_t)(u0 + k) * D + j];
                        const z2 q = zcmul({aj.x, aj.y}, {sr[k], si[k]});
                        qr[k] += q.r;
                        qi[k] += q.i;
                    }
                }
            }
            // the workgroup's four parts meet in LDS; one atomic per (u, Re/Im) and workgroup
            if (lane == 0) {
#pragma unroll
                for (int k = 0; k < U; ++k) {
                    red[wv][2 * k] = qr[k];
                    red[wv][2 * k + 1] = qi[k];
                }
            }
            __syncthreads();
            if (threadIdx.x < 2 * U && u0 + (int)(threadIdx.x >> 1) < nu) {
                const int k = threadIdx.x >> 1;
                const double v = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
                atomicAdd(&out[(size_t)(u0 + k) * (nrho + 1) + l].x + (threadIdx.x & 1), v);
            }
            __syncthreads();   // red[] is reused
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
                             const double* lam, double* y, double* kz, int D, int nu, unsigned long long* emax,
                             hipStream_t s)
{
    const long n = (long)D * nu;
    hipLaunchKernelGGL(k_stage_prep, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (const double2*)b,
                       (const double2*)ks, nstage, coef, h, t, lam, (double2*)y, (double2*)kz, D, nu, emax);
    return hipGetLastError();
}

hipError_t launch_hermitian_part(double* W, int D, int nu, hipStream_t s)
{
    const long n = (long)D * D * nu;
    hipLaunchKernelGGL(k_hermitian_part, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (double2*)W, D, nu);
    return hipGetLastError();
}

hipError_t launch_stage_hemv(const double* W, const double* y, double* kout, double F, double t, const double* lam,
                             int D, int nu, hipStream_t s)
{
    const int nt = (D + 63) / 64;
    if (nu < 1 || nu > 65535) return hipErrorInvalidValue;
    dim3 grid((unsigned)((long)nt * (nt + 1) / 2), (unsigned)nu);
    hipLaunchKernelGGL(k_stage_hemv, grid, dim3(256), 0, s, (const double2*)W, (const double2*)y, (double2*)kout, F,
                       t, lam, D, nt);
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
    // (row, 640-column chunk) tasks of R_l spread over 1024 workgroups of four wavefronts (four
    // per CU): enough wavefronts in flight to cover the row loads' latency, and one atomic add
    // per sum and workgroup instead of one per sum and row (one per row, on the same 2 nu
    // addresses, took 0.21 ms per call at D = 2714, nu = 8; this takes 0.10 ms, 512
    // workgroups 0.13 ms)
    const int chunk = 640;
    hipLaunchKernelGGL(k_observe, dim3(1024u), dim3(256), 0, s, (const double2*)a,
                       (const double2*)R, D, nu, nrho, chunk, (double2*)out);
    return hipGetLastError();
}

}  // namespace le
