// Two-sided Metropolis chain and its features, for the occupation numbers
// <L| a^dagger_l a_l |R> (reference: greenFunctionMC, modules/green_function/green_function.cu:199-373
// and gf_occupation_numbers.cu:16-239).
//
// The reference runs one chain per (L, R) pair on |p_L Psi(z_0..z_{N-1})| |p_R Psi(z_1..z_N)|.
// Here one chain samples |Psi(z_0..z_{N-1})| |Psi(z_1..z_N)| (Psi the Laughlin state) for
// every pair at once; p_L, p_R and the phase enter the accumulated observable:
//   O_ab(l) = < conj(p_a(z_0..z_{N-1})) p_b(z_1..z_N) e^{i phi} phi_l(z_0) conj(phi_l(z_N)) >,
//   e^{i phi} = (q / |q|)^m,  q = prod_{i=1}^{N-1} (z_i - z_N) conj(z_0 - z_i),
//   phi_l(z_0) conj(phi_l(z_N)) = (z_0 conj(z_N) / 2)^l / (2 pi l!) e^{-(|z_0|^2 + |z_N|^2)/4}.
// Move order per sweep: particle 0, particles 1..N-1, particle N (gf_occupation_numbers.cu:53-201).
// log pi = m sum_{0<=i<j<=N-1} log|z_i - z_j| + m sum_{1<=i<j<=N} log|z_i - z_j|
//          - (1/4) sum_{i<N} |z_i|^2 - (1/4) sum_{i>0} |z_i|^2.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "le_device.hpp"
#include "le_kernels.hpp"

namespace le {

template <typename Real>
__device__ __forceinline__ Real lg2(Real x)
{
    if constexpr (sizeof(Real) == 4) return __builtin_amdgcn_logf(fmaxf(x, 1e-30f));
    else return log2(fmax(x, 1e-300));
}

// log of |z' - q|^2 / |z - q|^2 (LogOps units) for the pairs of a bulk move with z_0 and z_N;
// both sweep kernels below evaluate them the same way
template <typename Real, typename V2>
__device__ __forceinline__ Real pair_lr(Real xn, Real yn, V2 zo, V2 q)
{
    const Real ax = xn - q.x, ay = yn - q.y, bx = zo.x - q.x, by = zo.y - q.y;
    return LogOps<Real>::lg(mad(ax, ax, ay * ay)) - LogOps<Real>::lg(mad(bx, bx, by * by));
}

// Bulk particles 1..N-1 sit in LDS slots 0..N-2 (lane l owns slots l + G s); z_0 and z_N
// live in registers of every lane of the group.  Per move, the pair products follow the edge
// sweep (k_sweep): the mover's slot and the pads hold the midpoint of old and proposed
// positions, FP32 takes |z' - q|^2 = |b|^2 + |d|^2 + 2 d.b (b = z - q, d = z' - z) and keeps
// the exponents of 8 products at a time in an integer, FP64 takes both distances and one log
// per 32 products; the group sum is the DPP butterfly.  Random numbers: one Philox block per
// move (counter = (move, sweep, walker)), made G moves at a time, lane l for move i + l; lane
// i mod G hands the move's normals and ln u to the group through __shfl.  With one Philox
// block and one FP64 Box-Muller pair per move in every lane the sweep took 2.9x the edge
// sweep's time at G = 16.
template <int G, typename Real>
__global__ __launch_bounds__(64) void k_sweep2(SweepArgs a)
{
    using V2 = typename V2T<Real>::T;
    extern __shared__ __align__(16) unsigned char smem2[];
    V2* lds = reinterpret_cast<V2*>(smem2);
    constexpr int WPB = 64 / G;
    constexpr int C = LogOps<Real>::chunk;
    const int N = a.N;
    const int NB = N - 1;
    const int S = (NB + G - 1) / G;
    const int NP = S * G;
    const int lane = threadIdx.x % G;
    const int grp = threadIdx.x / G;
    const int walker = blockIdx.x * WPB + grp;
    const bool live = walker < a.walkers;
    V2* P = lds + grp * NP;
    V2* gpos = reinterpret_cast<V2*>(a.pos) + (size_t)(live ? walker : 0) * (N + 1);
    if (live)
        for (int j = lane; j < NB; j += G) P[j] = gpos[1 + j];
    __syncthreads();
    if (!live) return;
    V2 z0 = gpos[0], zN = gpos[N];

    const Real step = (Real)a.step;
    const Real mfac = (Real)a.m * LogOps<Real>::to_ln;
    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
    const bool has_pad = (lane + G * (S - 1)) >= NB;
    const int pad_j = lane + G * (S - 1);
    unsigned accepted = 0;

    for (int sw = 0; sw < a.nsweeps; ++sw) {
        const unsigned long long sidx = a.sweep0 + (a.sweep_base ? *a.sweep_base : 0ull) + (unsigned long long)sw;
        const uint32_t c1 = (uint32_t)sidx;
        const uint32_t c2 = (uint32_t)((sidx >> 32) & 0xffffu);
        Real rg1 = 0, rg2 = 0;
        float rlnu = 0.f;
        uint32_t rub = 0u;
        for (int i = 0; i <= N; ++i) {
            const int t = i % G;
            if (t == 0 && i + lane <= N) {
                uint32_t r[4];
                philox((uint32_t)(i + lane), c1, c2, (uint32_t)walker, k0, k1, r);
                gauss2(r[0], r[1], rg1, rg2);
                if constexpr (sizeof(Real) == 4) rlnu = __builtin_amdgcn_logf(u01f(r[2])) * 0.69314718055994531f;
                rub = r[2];
            }
            const Real g1 = __shfl(rg1, t, G), g2 = __shfl(rg2, t, G);
            const bool bulk = (i > 0 && i < N);
            const V2 zo = (i == 0) ? z0 : ((i == N) ? zN : P[i - 1]);
            const Real xn = mad(step, g1, zo.x), yn = mad(step, g2, zo.y);
            V2 mid;
            mid.x = (Real)0.5 * (zo.x + xn);
            mid.y = (Real)0.5 * (zo.y + yn);
            if (bulk) P[i - 1] = mid;
            if (has_pad) P[pad_j] = mid;
            const Real dx = xn - zo.x, dy = yn - zo.y;
            const Real dd = mad(dx, dx, dy * dy), tdx = dx + dx, tdy = dy + dy;
            Real lsum = 0, pn = 1, pd = 1;
            int ebal = 0;
            for (int s = 0; s < S; ++s) {
                const V2 q = P[lane + G * s];
                const Real bx = zo.x - q.x, by = zo.y - q.y;
                const Real B = mad(bx, bx, by * by);
                Real A;
                if constexpr (sizeof(Real) == 4) {
                    A = mad(tdx, bx, mad(tdy, by, B + dd));
                } else {
                    const Real ax = xn - q.x, ay = yn - q.y;
                    A = mad(ax, ax, ay * ay);
                }
                pn *= A;
                pd *= B;
                if ((s + 1) % C == 0) {
                    if constexpr (sizeof(Real) == 4) {
                        int en, ed;
                        pn = frexpf(pn, &en);
                        pd = frexpf(pd, &ed);
                        ebal += en - ed;
                    } else {
                        lsum += LogOps<Real>::lg(pn) - LogOps<Real>::lg(pd);
                        pn = 1;
                        pd = 1;
                    }
                }
            }
            if constexpr (sizeof(Real) == 4)
                lsum = __builtin_amdgcn_logf(pn) - __builtin_amdgcn_logf(pd) + (float)ebal;
            else
                lsum += LogOps<Real>::lg(pn) - LogOps<Real>::lg(pd);
            lsum = group_sum<G, Real>(lsum);
            const Real dr2 = mad(xn, xn, yn * yn) - mad(zo.x, zo.x, zo.y * zo.y);
            Real logA;
            if (bulk) {
                // bulk pairs count twice (both wavefunctions), the pairs with z_0 and z_N once
                const Real e = pair_lr(xn, yn, zo, z0) + pair_lr(xn, yn, zo, zN);
                logA = mad((Real)-0.5, dr2, mfac * (lsum + (Real)0.5 * e));
            } else {
                logA = mad((Real)-0.25, dr2, (Real)0.5 * mfac * lsum);
            }
            bool acc;
            if constexpr (sizeof(Real) == 4)
                acc = (logA >= 0.f) || (__shfl(rlnu, t, G) < logA);
            else
                acc = (logA >= 0.0) || (log(u01d((uint32_t)__shfl((int)rub, t, G))) < logA);
            V2 nz;
            nz.x = acc ? xn : zo.x;
            nz.y = acc ? yn : zo.y;
            if (bulk) P[i - 1] = nz;
            if (i == 0) z0 = nz;
            if (i == N) zN = nz;
            accepted += acc ? 1u : 0u;
        }
    }
    for (int j = lane; j < NB; j += G) gpos[1 + j] = P[j];
    if (lane == 0) {
        gpos[0] = z0;
        gpos[N] = zN;
        a.accepted[walker] += accepted;
    }
}

// The same chain with each lane's bulk particles in registers: lane l of a group holds bulk
// particles j = l + G s (s < S, j < N - 1) in px[s], py[s]; the moves of z_0 and z_N pair
// with every bulk particle, a bulk move i = 1 + G s + t takes its old position and random
// numbers from lane t by __shfl.  Products, chunking, pads and acceptance are those of
// k_sweep2, so for the same G the chain is bit-identical to it.
template <int G, int S, typename Real>
__global__ __launch_bounds__(64) void k_sweep2_reg(SweepArgs a)
{
    using V2 = typename V2T<Real>::T;
    constexpr int WPB = 64 / G;
    constexpr int C = LogOps<Real>::chunk;
    constexpr int FULL = (S / C) * C;
    const int N = a.N;
    const int NB = N - 1;               // (S - 1) G < NB <= S G, checked by the launcher
    const int lane = threadIdx.x % G;
    const int grp = threadIdx.x / G;
    const int walker = blockIdx.x * WPB + grp;
    if (walker >= a.walkers) return;    // whole groups leave; shuffles stay within a group
    V2* gpos = reinterpret_cast<V2*>(a.pos) + (size_t)walker * (N + 1);
    Real px[S], py[S];
    static_for(std::make_integer_sequence<int, S>{}, [&](auto sc) {
        constexpr int s = decltype(sc)::value;
        const int j = lane + G * s;
        const V2 q = j < NB ? gpos[1 + j] : gpos[1];
        px[s] = q.x;
        py[s] = q.y;
    });
    V2 z0 = gpos[0], zN = gpos[N];
    const bool pad_last = lane + G * (S - 1) >= NB;

    const Real step = (Real)a.step;
    const Real mfac = (Real)a.m * LogOps<Real>::to_ln;
    const unsigned long long sweep_first = a.sweep0 + (a.sweep_base ? *a.sweep_base : 0ull);
    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
    unsigned accepted = 0;

    // group sum of the bulk pair log-ratio of a move from zo to (xn, yn) with midpoint mid
    auto bulk_lsum = [&](const V2 zo, const Real xn, const Real yn, const V2 mid) {
        const Real dx = xn - zo.x, dy = yn - zo.y;
        const Real dd = mad(dx, dx, dy * dy), tdx = dx + dx, tdy = dy + dy;
        Real lsum = 0, pn = 1, pd = 1;
        int ebal = 0;
        static_for(std::make_integer_sequence<int, S>{}, [&](auto uc) {
            constexpr int u = decltype(uc)::value;
            Real qx = px[u], qy = py[u];
            if constexpr (u == S - 1) {
                if (pad_last) {              // pads hold the midpoint
                    qx = mid.x;
                    qy = mid.y;
                }
            }
            const Real bx = zo.x - qx, by = zo.y - qy;
            const Real B = mad(bx, bx, by * by);
            Real A;
            if constexpr (sizeof(Real) == 4) {
                A = mad(tdx, bx, mad(tdy, by, B + dd));
            } else {
                const Real ax = xn - qx, ay = yn - qy;
                A = mad(ax, ax, ay * ay);
            }
            pn *= A;
            pd *= B;
            if constexpr ((u + 1) % C == 0 && u + 1 <= FULL) {
                if constexpr (sizeof(Real) == 4) {
                    int en, ed;
                    pn = frexpf(pn, &en);
                    pd = frexpf(pd, &ed);
                    ebal += en - ed;
                } else {
                    lsum += LogOps<Real>::lg(pn) - LogOps<Real>::lg(pd);
                    pn = 1;
                    pd = 1;
                }
            }
        });
        if constexpr (sizeof(Real) == 4)
            lsum = __builtin_amdgcn_logf(pn) - __builtin_amdgcn_logf(pd) + (float)ebal;
        else
            lsum += LogOps<Real>::lg(pn) - LogOps<Real>::lg(pd);
        return group_sum<G, Real>(lsum);
    };
    // the moves of z_0 (move 0) and z_N (move N): same numbers in every lane
    auto end_move = [&](V2& z, uint32_t mv, uint32_t c1, uint32_t c2) {
        uint32_t r[4];
        philox(mv, c1, c2, (uint32_t)walker, k0, k1, r);
        Real g1, g2;
        gauss2(r[0], r[1], g1, g2);
        const Real xn = mad(step, g1, z.x), yn = mad(step, g2, z.y);
        V2 mid;
        mid.x = (Real)0.5 * (z.x + xn);
        mid.y = (Real)0.5 * (z.y + yn);
        const Real lsum = bulk_lsum(z, xn, yn, mid);
        const Real dr2 = mad(xn, xn, yn * yn) - mad(z.x, z.x, z.y * z.y);
        const Real logA = mad((Real)-0.25, dr2, (Real)0.5 * mfac * lsum);
        bool acc;
        if constexpr (sizeof(Real) == 4)
            acc = (logA >= 0.f) || (__builtin_amdgcn_logf(u01f(r[2])) * 0.69314718055994531f < logA);
        else
            acc = (logA >= 0.0) || (log(u01d(r[2])) < logA);
        if (acc) {
            z.x = xn;
            z.y = yn;
        }
        accepted += acc ? 1u : 0u;
    };

    for (int sw = 0; sw < a.nsweeps; ++sw) {
        const unsigned long long sidx = sweep_first + (unsigned long long)sw;
        const uint32_t c1 = (uint32_t)sidx;
        const uint32_t c2 = (uint32_t)((sidx >> 32) & 0xffffu);
        end_move(z0, 0u, c1, c2);
        static_for(std::make_integer_sequence<int, S>{}, [&](auto sc) {
            constexpr int s = decltype(sc)::value;
            const int base = G * s;
            // opaque counter: keeps the compiler from hoisting the sweep-invariant part of
            // Philox for every slot out of the sweep loop
            uint32_t ctr = (uint32_t)(1 + base + lane);
            asm volatile("" : "+v"(ctr));
            uint32_t r[4];
            philox(ctr, c1, c2, (uint32_t)walker, k0, k1, r);
            Real g1, g2;
            gauss2(r[0], r[1], g1, g2);
            float lnu = 0.f;
            if constexpr (sizeof(Real) == 4) lnu = __builtin_amdgcn_logf(u01f(r[2])) * 0.69314718055994531f;
            const uint32_t ub = r[2];
            const int cnt = min(G, NB - base);
            for (int t = 0; t < cnt; ++t) {
                V2 zo;
                zo.x = __shfl(px[s], t, G);
                zo.y = __shfl(py[s], t, G);
                const Real d1 = __shfl(g1, t, G), d2 = __shfl(g2, t, G);
                const Real xn = mad(step, d1, zo.x), yn = mad(step, d2, zo.y);
                V2 mid;
                mid.x = (Real)0.5 * (zo.x + xn);
                mid.y = (Real)0.5 * (zo.y + yn);
                const bool own = lane == t;
                if (own) {
                    px[s] = mid.x;
                    py[s] = mid.y;
                }
                const Real lsum = bulk_lsum(zo, xn, yn, mid);
                const Real dr2 = mad(xn, xn, yn * yn) - mad(zo.x, zo.x, zo.y * zo.y);
                const Real e = pair_lr(xn, yn, zo, z0) + pair_lr(xn, yn, zo, zN);
                const Real logA = mad((Real)-0.5, dr2, mfac * (lsum + (Real)0.5 * e));
                bool acc;
                if constexpr (sizeof(Real) == 4)
                    acc = (logA >= 0.f) || (__shfl(lnu, t, G) < logA);
                else
                    acc = (logA >= 0.0) || (log(u01d((uint32_t)__shfl((int)ub, t, G))) < logA);
                if (own) {
                    px[s] = acc ? xn : zo.x;
                    py[s] = acc ? yn : zo.y;
                }
                accepted += acc ? 1u : 0u;
            }
        });
        end_move(zN, (uint32_t)N, c1, c2);
    }
    static_for(std::make_integer_sequence<int, S>{}, [&](auto sc) {
        constexpr int s = decltype(sc)::value;
        const int j = lane + G * s;
        if (j < NB) {
            V2 q;
            q.x = px[s];
            q.y = py[s];
            gpos[1 + j] = q;
        }
    });
    if (lane == 0) {
        gpos[0] = z0;
        gpos[N] = zN;
        a.accepted[walker] += accepted;
    }
}

constexpr int kLps2 = 4;   // lanes per sample in k_psums2

// Features of the two-sided chain in three passes.  k_psums2: bulk power sums, 4 lanes per
// sample (bulk particles j = 1 + sub + 4 s): sum_{0<j<N} (z_j / R)^k, k <= KMAX, and the unit
// phase q / |q| with q = prod_{0<j<N} (z_j - z_N) conj(z_0 - z_j), written to the sample's
// scratch row.  k_features2: the monomials (one wavefront per sample writes its row of X).
// k_weights2: the orbital weights, one thread per sample, so that a wavefront writes 64
// consecutive samples of each weight row (weights are stored weight-major for the Gram);
// written by one wavefront per sample they were 2 lcut scattered 4-byte stores per sample,
// and the features took 0.29 ms for 32768 samples at N = 200, lcut = 598.
template <typename Real, int KMAX>
__global__ __launch_bounds__(256) void k_psums2(Features2Args a)
{
    const long gt = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long w = gt / kLps2;
    const int sub = (int)(gt % kLps2);
    if (w >= a.walkers) return;                  // the 4 lanes of a sample leave together
    const int N = a.N;
    auto load = [&](int j, double& x, double& y) {
        if (a.pos_is_double || sizeof(Real) == 8) {
            const double2 q = reinterpret_cast<const double2*>(a.pos)[(size_t)w * (N + 1) + j];
            x = q.x; y = q.y;
        } else {
            const float2 q = reinterpret_cast<const float2*>(a.pos)[(size_t)w * (N + 1) + j];
            x = q.x; y = q.y;
        }
    };
    double x0, y0, xN, yN;
    load(0, x0, y0);
    load(N, xN, yN);
    double acc[2 * KMAX];
#pragma unroll
    for (int q = 0; q < 2 * KMAX; ++q) acc[q] = 0.0;
    double qr = 1.0, qi = 0.0;                  // unit-modulus running product
    for (int j = 1 + sub; j < N; j += kLps2) {
        double x, y;
        load(j, x, y);
        const double ur = x * a.invR, ui = y * a.invR;
        double pr = ur, pi = ui;
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            acc[2 * k] += pr;
            acc[2 * k + 1] += pi;
            const double t = pr * ur - pi * ui;
            pi = pr * ui + pi * ur;
            pr = t;
        }
        const double ar = x - xN, ai = y - yN, br = x0 - x, bi = -(y0 - y);
        double tr = ar * br - ai * bi, ti = ar * bi + ai * br;
        const double inv = 1.0 / sqrt(tr * tr + ti * ti);
        tr *= inv;
        ti *= inv;
        const double nr = qr * tr - qi * ti;
        qi = qr * ti + qi * tr;
        qr = nr;
    }
#pragma unroll
    for (int q = 0; q < 2 * KMAX; ++q) acc[q] = group_sum<kLps2, double>(acc[q]);
#pragma unroll
    for (int off = 1; off < kLps2; off <<= 1) {
        const double orr = __shfl_xor(qr, off, kLps2), oi = __shfl_xor(qi, off, kLps2);
        const double nr = qr * orr - qi * oi;
        qi = qr * oi + qi * orr;
        qr = nr;
    }
    double* ps = a.ps + (size_t)w * a.ps_ld;
#pragma unroll
    for (int q = 0; q < 2 * KMAX; ++q)
        if (q % kLps2 == sub && q < 2 * a.K) ps[q] = acc[q];   // the phase follows at 2 K
    if (sub == 0) {
        ps[2 * a.K] = qr;
        ps[2 * a.K + 1] = qi;
    }
}

template <typename Real, int KMAX>
__global__ __launch_bounds__(256) void k_features2(Features2Args a)
{
    __shared__ double pkL[4][2 * KMAX], pkR[4][2 * KMAX];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int w = blockIdx.x * 4 + wv;
    if (w >= a.walkers) return;
    const int N = a.N;
    if (lane == 0) {
        double x0, y0, xN, yN;
        if (a.pos_is_double || sizeof(Real) == 8) {
            const double2* p = reinterpret_cast<const double2*>(a.pos) + (size_t)w * (N + 1);
            x0 = p[0].x; y0 = p[0].y; xN = p[N].x; yN = p[N].y;
        } else {
            const float2* p = reinterpret_cast<const float2*>(a.pos) + (size_t)w * (N + 1);
            x0 = p[0].x; y0 = p[0].y; xN = p[N].x; yN = p[N].y;
        }
        const double* ps = a.ps + (size_t)w * a.ps_ld;
        // left adds z_0^k, right adds z_N^k
        const double u0r = x0 * a.invR, u0i = y0 * a.invR, uNr = xN * a.invR, uNi = yN * a.invR;
        double lr = u0r, li = u0i, rr = uNr, ri = uNi;
        for (int k = 0; k < a.K; ++k) {
            pkL[wv][2 * k] = ps[2 * k] + lr;
            pkL[wv][2 * k + 1] = ps[2 * k + 1] + li;
            pkR[wv][2 * k] = ps[2 * k] + rr;
            pkR[wv][2 * k + 1] = ps[2 * k + 1] + ri;
            double t = lr * u0r - li * u0i;
            li = lr * u0i + li * u0r;
            lr = t;
            t = rr * uNr - ri * uNi;
            ri = rr * uNi + ri * uNr;
            rr = t;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    const long row = a.row0 + w;
    Real* X = reinterpret_cast<Real*>(a.X) + (size_t)row * a.ldx;
    for (int t = lane; t < a.D; t += 64) {
        const double* pk = a.side[t] ? pkR[wv] : pkL[wv];
        double vr = 1.0, vi = 0.0;
        const uint32_t* f = a.fac + (size_t)t * 8;
#pragma unroll 1
        for (int q = 0; q < 8; ++q) {
            const uint32_t word = f[q];
            if (word == 0u) break;
            const int k = (int)(word & 0xffu), r = (int)(word >> 8);
            const double br = pk[2 * (k - 1)], bi = pk[2 * (k - 1) + 1];
            for (int e = 0; e < r; ++e) {
                const double tr = vr * br - vi * bi;
                vi = vr * bi + vi * br;
                vr = tr;
            }
        }
        const int c = a.col_re[t];
        X[c] = (Real)vr;
        X[c + a.col_dim[t]] = (Real)vi;
    }
}

// w_l = e^{i phi} (z_0 conj(z_N) / 2)^l e^{-(|z_0|^2 + |z_N|^2)/4} / (2 pi l!), e^{i phi} = (q/|q|)^m;
// one thread per sample
template <typename Real>
__global__ __launch_bounds__(256) void k_weights2(Features2Args a)
{
    const long w = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= a.walkers) return;
    const int N = a.N;
    double x0, y0, xN, yN;
    if (a.pos_is_double || sizeof(Real) == 8) {
        const double2* p = reinterpret_cast<const double2*>(a.pos) + (size_t)w * (N + 1);
        x0 = p[0].x; y0 = p[0].y; xN = p[N].x; yN = p[N].y;
    } else {
        const float2* p = reinterpret_cast<const float2*>(a.pos) + (size_t)w * (N + 1);
        x0 = p[0].x; y0 = p[0].y; xN = p[N].x; yN = p[N].y;
    }
    const double* ps = a.ps + (size_t)w * a.ps_ld;
    const double qr = ps[2 * a.K], qi = ps[2 * a.K + 1];
    double er = 1.0, ei = 0.0;
    for (int e = 0; e < a.m; ++e) {
        const double t = er * qr - ei * qi;
        ei = er * qi + ei * qr;
        er = t;
    }
    const double xr = 0.5 * (x0 * xN + y0 * yN), xi = 0.5 * (y0 * xN - x0 * yN);
    const double lx = 0.5 * log(xr * xr + xi * xi), ax = atan2(xi, xr);
    const double g0 = -0.25 * (x0 * x0 + y0 * y0 + xN * xN + yN * yN);
    const long row = a.row0 + w;
    Real* W = reinterpret_cast<Real*>(a.W);
    if constexpr (sizeof(Real) == 4) {
        // FP32 weights: log-magnitude and angle (in revolutions, reduced to [0, 1)) in FP64,
        // then the hardware exp, sin and cos (relative error ~1e-6 per weight)
        const double arev = ax * 0.15915494309189535;
        const float fer = (float)er, fei = (float)ei;
        for (int l = 0; l < a.lcut; ++l) {
            const float mag = __expf((float)(g0 + l * lx - a.ln2pifact[l]));
            double tr = l * arev;
            tr -= floor(tr);
            const float s = __builtin_amdgcn_sinf((float)tr), c = __builtin_amdgcn_cosf((float)tr);
            W[(size_t)(2 * l) * a.wld + row] = mag * (c * fer - s * fei);
            W[(size_t)(2 * l + 1) * a.wld + row] = mag * (c * fei + s * fer);
        }
    } else {
        for (int l = 0; l < a.lcut; ++l) {
            const double mag = exp(g0 + l * lx - a.ln2pifact[l]);
            double s, c;
            sincos(l * ax, &s, &c);
            W[(size_t)(2 * l) * a.wld + row] = mag * (c * er - s * ei);
            W[(size_t)(2 * l + 1) * a.wld + row] = mag * (c * ei + s * er);
        }
    }
}

size_t sweep2_lds_bytes(int prec, int G, int N)
{
    const int S = (N - 1 + G - 1) / G;
    const size_t el = (prec == kFP64) ? sizeof(double2) : sizeof(float2);
    return (size_t)(64 / G) * (size_t)S * G * el;
}

template <int G, typename Real>
static hipError_t sweep2_g(const SweepArgs& a, hipStream_t s)
{
    const size_t lds = sweep2_lds_bytes(sizeof(Real) == 8 ? kFP64 : kFP32, G, a.N);
    constexpr int WPB = 64 / G;
    dim3 grid((unsigned)((a.walkers + WPB - 1) / WPB));
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)k_sweep2<G, Real>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_sweep2<G, Real>), grid, dim3(64), lds, s, a);
    return hipGetLastError();
}

template <typename Real>
static hipError_t sweep2_dispatch(const SweepArgs& a, int G, hipStream_t s)
{
    switch (G) {
    case 2: return sweep2_g<2, Real>(a, s);
    case 4: return sweep2_g<4, Real>(a, s);
    case 8: return sweep2_g<8, Real>(a, s);
    case 16: return sweep2_g<16, Real>(a, s);
    case 32: return sweep2_g<32, Real>(a, s);
    case 64: return sweep2_g<64, Real>(a, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_sweep2(const SweepArgs& a, int prec, int G, hipStream_t s)
{
    if (a.nsweeps <= 0 || a.walkers <= 0) return hipSuccess;
    if (a.N < 2) return hipErrorInvalidValue;
    return prec == kFP64 ? sweep2_dispatch<double>(a, G, s) : sweep2_dispatch<float>(a, G, s);
}

bool sweep2_reg_supported(int G, int N)
{
    const int S = (N - 1 + G - 1) / G;   // N - 1 bulk particles
    return N >= 2 && (G == 8 || G == 16) && (S == 13 || S == 25);
}

template <int G, int S, typename Real>
static hipError_t sweep2_reg_gs(const SweepArgs& a, hipStream_t s)
{
    constexpr int WPB = 64 / G;
    dim3 grid((unsigned)((a.walkers + WPB - 1) / WPB));
    hipLaunchKernelGGL((k_sweep2_reg<G, S, Real>), grid, dim3(64), 0, s, a);
    return hipGetLastError();
}

template <typename Real>
static hipError_t sweep2_reg_dispatch(const SweepArgs& a, int G, hipStream_t s)
{
    const int S = (a.N - 1 + G - 1) / G;
    if (G == 8 && S == 13) return sweep2_reg_gs<8, 13, Real>(a, s);
    if (G == 8 && S == 25) return sweep2_reg_gs<8, 25, Real>(a, s);
    if (G == 16 && S == 13) return sweep2_reg_gs<16, 13, Real>(a, s);
    if (G == 16 && S == 25) return sweep2_reg_gs<16, 25, Real>(a, s);
    return hipErrorInvalidValue;
}

hipError_t launch_sweep2_reg(const SweepArgs& a, int prec, int G, hipStream_t s)
{
    if (a.nsweeps <= 0 || a.walkers <= 0) return hipSuccess;
    if (!sweep2_reg_supported(G, a.N)) return hipErrorInvalidValue;
    return prec == kFP64 ? sweep2_reg_dispatch<double>(a, G, s) : sweep2_reg_dispatch<float>(a, G, s);
}

template <typename Real>
static hipError_t features2_t(const Features2Args& a, hipStream_t s)
{
    if (!a.ps || a.ps_ld < 2 * a.K + 2) return hipErrorInvalidValue;
    dim3 grid_ps((unsigned)(((long)a.walkers * kLps2 + 255) / 256));
    dim3 grid((unsigned)((a.walkers + 3) / 4));
    dim3 grid_w((unsigned)((a.walkers + 255) / 256));
    if (a.K <= 8) {
        hipLaunchKernelGGL((k_psums2<Real, 8>), grid_ps, dim3(256), 0, s, a);
        hipLaunchKernelGGL((k_features2<Real, 8>), grid, dim3(256), 0, s, a);
    } else if (a.K <= 16) {
        hipLaunchKernelGGL((k_psums2<Real, 16>), grid_ps, dim3(256), 0, s, a);
        hipLaunchKernelGGL((k_features2<Real, 16>), grid, dim3(256), 0, s, a);
    } else if (a.K <= 24) {
        hipLaunchKernelGGL((k_psums2<Real, 24>), grid_ps, dim3(256), 0, s, a);
        hipLaunchKernelGGL((k_features2<Real, 24>), grid, dim3(256), 0, s, a);
    } else {
        return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL((k_weights2<Real>), grid_w, dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_features2(const Features2Args& a, int prec, hipStream_t s)
{
    if (a.walkers <= 0) return hipSuccess;
    return prec == kFP64 ? features2_t<double>(a, s) : features2_t<float>(a, s);
}

}  // namespace le
