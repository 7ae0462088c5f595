// Edge-state Monte-Carlo kernels for gfx950 (MI355X).
//
//  k_init_positions  uniform placement in the disc of radius R (Philox tag 1 stream)
//  k_sweep<G, Real>  Metropolis sweeps on |Psi_L|^2; G lanes share one walker whose
//                    positions live in LDS; the O(N) log-ratio of a single-particle move
//                    is split over the G lanes and combined with a butterfly reduction.
//  k_psums, k_monomials
//                    FP64 power sums p_k of z/R and confinement V (4 lanes per sample),
//                    then partition monomials p_lambda (one wavefront per sample), written
//                    as a real-embedded row [Re p(sector) | Im p(sector)] of the sample
//                    matrix X.
//  k_gram<Real>      weighted Gram tiles C = X^T X and X^T diag(V) X on MFMA
//                    (16x16x4 f32 / f64), upper block-triangle tiles only, FP32 tiles
//                    flushed into FP64 registers every 256 samples, FP64 atomics out.
//  k_extract         FP64 real-embedded tiles -> complex Hermitian S, H per sector.
//  k_geneig_vals     FP64 generalized eigensolver, one workgroup per sector: Cholesky,
//                    Householder tridiagonalization, bisection;
//  k_eig_invit, k_eig_back
//                    eigenvectors: inverse iteration on the tridiagonal form, then the
//                    reflectors and L^-H.
//
// Reference behaviour these kernels reproduce (Nardin94/LaughlinEdge):
//  Metropolis move and ratio ........ modules/edge_mc/em_markovchain.cu:282-328
//  power sums / monomials ........... modules/edge_mc/em_markovchain.cu:120-183
//  confinement ...................... modules/edge_mc/em_markovchain.cu:92-99
//  generalized eigenproblem ......... modules/eigensolver/generalized_eigenproblem.cpp:45-52
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "le_device.hpp"
#include <utility>

#include "le_basis.hpp"
#include "le_kernels.hpp"

namespace le {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// initial positions
// ---------------------------------------------------------------------------
template <typename Real>
__global__ void k_init_positions(typename V2T<Real>::T* pos, int N, int walkers, double R,
                                 unsigned long long seed, int walker_id0)
{
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)N * walkers) return;
    int w = (int)(g / N), i = (int)(g % N);
    uint32_t r[4];
    philox((uint32_t)i, 0u, 1u << 16, (uint32_t)(w + walker_id0), (uint32_t)seed, (uint32_t)(seed >> 32), r);
    double rad = R * sqrt(u01d(r[0]));
    double s, c;
    sincos(6.283185307179586 * u01d(r[1]), &s, &c);
    pos[g].x = (Real)(rad * c);
    pos[g].y = (Real)(rad * s);
}

// ---------------------------------------------------------------------------
// Metropolis sweeps.  Each group of G lanes owns one walker; lane l handles particles
// j = l + G s (s < S = ceil(N / G)).  While particle i is moved its LDS slot holds the
// midpoint of old and proposed positions, so the i-i pair contributes equal factors
// to numerator and denominator; pad slots j >= N hold the same midpoint.  Products of
// eight squared distances are taken to log2 at a time (FP32 range), FP64 takes one log.
// ---------------------------------------------------------------------------

// Importance factor of the row-wise estimator (em_markovchain.cu:185-220, em_spectrum.cu:49-96):
// the chain samples |p_f Psi_L|^2, so a move adds log|p_f(z')|^2 - log|p_f(z)|^2 to log A.
// p_f = prod_q p_{k_q}^{r_q} over the row state's factors; the running p_k are updated by
// -u^k + u'^k with u = z / R (powers by repeated multiplication, k = 1..kmax), and every product
// is rounded on its own (no contraction), as the FP64 CPU restatement computes them.
struct RowState {
    double2 pk[kMaxFactors];
    int k[kMaxFactors], r[kMaxFactors];
    int kmax;
    double lm;               // log |p_f|^2 at the current positions
};

__device__ __forceinline__ double2 cmul_exact(double2 a, double2 b)
{
#pragma clang fp contract(off)
    double2 c;
    c.x = a.x * b.x - a.y * b.y;
    c.y = a.x * b.y + a.y * b.x;
    return c;
}

__device__ __forceinline__ double log_abs2_row(const RowState& st, const double2 (&pk)[kMaxFactors])
{
#pragma clang fp contract(off)
    double2 v = make_double2(1.0, 0.0);
#pragma unroll
    for (int q = 0; q < kMaxFactors; ++q)
        for (int t = 0; t < st.r[q]; ++t) v = cmul_exact(v, pk[q]);
    return log(v.x * v.x + v.y * v.y);
}

// proposal z -> z' of a row chain: the updated p_k (prop) and log|p_f(z')|^2
__device__ __forceinline__ double row_propose(const RowState& st, double xo, double yo, double xn, double yn, double R,
                                              double2 (&prop)[kMaxFactors])
{
#pragma clang fp contract(off)
    const double2 uo = make_double2(xo / R, yo / R), un = make_double2(xn / R, yn / R);
    double2 qo = uo, qn = un;
#pragma unroll
    for (int q = 0; q < kMaxFactors; ++q) prop[q] = st.pk[q];
    for (int k = 1; k <= st.kmax; ++k) {
#pragma unroll
        for (int q = 0; q < kMaxFactors; ++q)
            if (st.k[q] == k) {
                prop[q].x = (st.pk[q].x - qo.x) + qn.x;
                prop[q].y = (st.pk[q].y - qo.y) + qn.y;
            }
        qo = cmul_exact(qo, uo);
        qn = cmul_exact(qn, un);
    }
    return log_abs2_row(st, prop);
}

// Mixture density w |Psi_L|^2 (MixParams): lane l of a walker's G lanes keeps the running p_k
// for k = l + 1 + G j (j < M).  A move z -> z' changes them by u'^k - u^k (u = z / R): the lanes
// take u^(l+1) from an inclusive prefix product over the group (log2 G shuffle steps) and step
// on by u^G.  Each lane sums its terms exp(a_k ln|p_k|^2 + b_k); the group sum (the same in every
// lane: butterfly over commutative additions) plus c0 is w, and log A gains ln w' - ln w.  FP32
// evaluates a term as exp2(a_k log2|p_k|^2 + b_k / ln 2) with fused multiply-adds; FP64 rounds
// every product on its own, as the FP64 CPU restatement does.  The running p_k of a launch start
// from launch_mix_init or from the p_k the features of the walkers' last sample wrote.
template <int G, typename Real, int KM = kMixMax>
struct MixLane {
    static constexpr int M = (KM + G - 1) / G;
    Real pr[M], pi[M];       // running p_k
    Real ak[M], bk[M];       // exponent and offset of each term (FP32: offset in log2 units)
    Real c0;
    int kmix;
    Real lw;                 // ln w at the current positions
};

__device__ __forceinline__ float mix_term(float a, float b, float re, float im)
{
    return __builtin_amdgcn_exp2f(a * __builtin_amdgcn_logf(re * re + im * im) + b);
}

__device__ __forceinline__ double mix_term(double a, double b, double re, double im)
{
#pragma clang fp contract(off)
    const double q = re * re + im * im;
    return q > 0.0 ? exp(a * log(q) + b) : 0.0;
}

// (ar, ai) *= (br, bi): FP32 with fused multiply-adds, FP64 with every product rounded
__device__ __forceinline__ void mix_cmul(float& ar, float& ai, float br, float bi)
{
    const float r = ar * br - ai * bi;
    ai = ar * bi + ai * br;
    ar = r;
}

__device__ __forceinline__ void mix_cmul(double& ar, double& ai, double br, double bi)
{
#pragma clang fp contract(off)
    const double r = ar * br - ai * bi;
    ai = ar * bi + ai * br;
    ar = r;
}

template <typename Real>
__device__ __forceinline__ Real mix_ln(Real w)
{
    if constexpr (sizeof(Real) == 4)
        return __builtin_amdgcn_logf(w) * 0.69314718055994531f;
    else
        return log(w);
}

template <int G, typename Real, int KM>
__device__ __forceinline__ void mix_load(MixLane<G, Real, KM>& m, const SweepArgs& a, int walker, int lane)
{
    constexpr int M = MixLane<G, Real, KM>::M;
    m.kmix = a.mix.kmix;
    m.c0 = (Real)a.mix.c0;
    Real local = 0;
#pragma unroll
    for (int j = 0; j < M; ++j) {
        const int k = lane + 1 + G * j;
        if (k <= m.kmix) {
            const double2 p = a.mix_pk[(size_t)walker * kMixMax + (k - 1)];
            m.pr[j] = (Real)p.x;
            m.pi[j] = (Real)p.y;
            m.ak[j] = (Real)a.mix.a[k - 1];
            m.bk[j] = sizeof(Real) == 4 ? (Real)(a.mix.b[k - 1] * 1.4426950408889634) : (Real)a.mix.b[k - 1];
            local += mix_term(m.ak[j], m.bk[j], m.pr[j], m.pi[j]);
        } else {
            m.pr[j] = m.pi[j] = m.ak[j] = m.bk[j] = 0;
        }
    }
    m.lw = mix_ln<Real>(group_sum<G, Real>(local) + m.c0);
}

// ln w at z' when the walker's particle moves from zo to zn; qr, qi receive the proposed p_k
template <int G, typename Real, int KM>
__device__ __forceinline__ Real mix_propose(const MixLane<G, Real, KM>& m, Real xo, Real yo, Real xn, Real yn,
                                            double R, int lane, Real (&qr)[MixLane<G, Real, KM>::M],
                                            Real (&qi)[MixLane<G, Real, KM>::M])
{
    constexpr int M = MixLane<G, Real, KM>::M;
    Real aor, aoi, anr, ani;
    if constexpr (sizeof(Real) == 4) {
        const float ir = (float)(1.0 / R);
        aor = xo * ir; aoi = yo * ir; anr = xn * ir; ani = yn * ir;
    } else {
        aor = xo / R; aoi = yo / R; anr = xn / R; ani = yn / R;
    }
    Real por = aor, poi = aoi, pnr = anr, pni = ani;
    // (lane exchanges through ds_bpermute: a DPP row-shift scan with u^G by squaring measured 15%
    // slower in the VALU-bound register sweep)
#pragma unroll
    for (int d = 1; d < G; d <<= 1) {
        const Real tor = __shfl_up(por, d, G), toi = __shfl_up(poi, d, G);
        const Real tnr = __shfl_up(pnr, d, G), tni = __shfl_up(pni, d, G);
        if (lane >= d) {
            mix_cmul(por, poi, tor, toi);
            mix_cmul(pnr, pni, tnr, tni);
        }
    }
    Real gor = 0, goi = 0, gnr = 0, gni = 0;
    if constexpr (M > 1) {
        gor = __shfl(por, G - 1, G); goi = __shfl(poi, G - 1, G);
        gnr = __shfl(pnr, G - 1, G); gni = __shfl(pni, G - 1, G);
    }
    Real local = 0;
#pragma unroll
    for (int j = 0; j < M; ++j) {
        if (j > 0) {
            mix_cmul(por, poi, gor, goi);
            mix_cmul(pnr, pni, gnr, gni);
        }
        const int k = lane + 1 + G * j;
        if (k <= m.kmix) {
            qr[j] = (m.pr[j] - por) + pnr;
            qi[j] = (m.pi[j] - poi) + pni;
            local += mix_term(m.ak[j], m.bk[j], qr[j], qi[j]);
        } else {
            qr[j] = qi[j] = 0;
        }
    }
    return mix_ln<Real>(group_sum<G, Real>(local) + m.c0);
}

template <int G, typename Real, bool ROW = false, bool MIX = false>
__global__ __launch_bounds__(64) void k_sweep(SweepArgs a)
{
    using V2 = typename V2T<Real>::T;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    V2* lds = reinterpret_cast<V2*>(smem_raw);
    constexpr int WPB = 64 / G;
    const int N = a.N;
    const int S = (N + G - 1) / G;
    const int NP = S * G;
    const int lane = threadIdx.x % G;
    const int grp = threadIdx.x / G;
    const int walker = blockIdx.x * WPB + grp;
    const bool live = walker < a.walkers;
    V2* P = lds + grp * NP;
    // per-group random numbers of the current batch of G moves: (n1, n2, u bits, ln u); the
    // FP32 path takes ln u = log2(u) ln 2 here, once per move, instead of in every lane
    struct Rnd {
        Real n1, n2;
        uint32_t ub;
        float lnu;
    };
    Rnd* rnd = reinterpret_cast<Rnd*>(lds + WPB * NP) + grp * G;
    V2* gpos = reinterpret_cast<V2*>(a.pos) + (size_t)(live ? walker : 0) * N;
    if (live)
        for (int j = lane; j < N; j += G) P[j] = gpos[j];
    __syncthreads();
    if (!live) return;

    const Real step = (Real)a.step;
    const Real mfac = (Real)a.m * LogOps<Real>::to_ln;
    const unsigned long long sweep_first = a.sweep0 + (a.sweep_base ? *a.sweep_base : 0ull);
    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
    const bool has_pad = (lane + G * (S - 1)) >= N;
    const int pad_j = lane + G * (S - 1);
    unsigned accepted = 0;
    const uint32_t wid = (uint32_t)(walker + a.walker_id0);
    RowState row;
    if constexpr (ROW) {
        // every lane of the walker holds the same row state and takes the same decisions
        const int f = walker / a.row_walkers;
        row.kmax = 0;
#pragma unroll
        for (int q = 0; q < kMaxFactors; ++q) {
            const uint32_t w = a.row_fac[(size_t)f * kMaxFactors + q];
            row.k[q] = (int)(w & 0xffu);
            row.r[q] = row.k[q] ? (int)((w >> 8) & 0xffu) : 0;
            row.kmax = max(row.kmax, row.k[q]);
            row.pk[q] = a.row_pk[(size_t)walker * kMaxFactors + q];
        }
        row.lm = log_abs2_row(row, row.pk);
    }
    MixLane<G, Real> mix;
    if constexpr (MIX) mix_load(mix, a, walker, lane);

    for (int sw = 0; sw < a.nsweeps; ++sw) {
        const unsigned long long sidx = sweep_first + (unsigned long long)sw;
        const uint32_t c1 = (uint32_t)sidx;
        const uint32_t c2 = (uint32_t)((sidx >> 32) & 0xffffu);
        for (int base = 0; base < N; base += G) {
            uint32_t r[4];
            philox((uint32_t)(base + lane), c1, c2, wid, k0, k1, r);
            Real g1, g2;
            gauss2(r[0], r[1], g1, g2);
            float lnu = 0.f;
            if constexpr (sizeof(Real) == 4) lnu = __builtin_amdgcn_logf(u01f(r[2])) * 0.69314718055994531f;
            rnd[lane] = Rnd{g1, g2, r[2], lnu};
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const int cnt = min(G, N - base);
            // the next move's old position and random numbers are read ahead: move t
            // writes only slot base + t and the pads, so slot base + t + 1 is final
            V2 zo = P[base];
            Rnd rn = rnd[0];
            for (int t = 0; t < cnt; ++t) {
                const int i = base + t;
                const Real d1 = rn.n1, d2 = rn.n2;
                const uint32_t ub = rn.ub;
                const float lnu_move = rn.lnu;
                const Real xn = mad(step, d1, zo.x), yn = mad(step, d2, zo.y);
                V2 mid;
                mid.x = (Real)0.5 * (zo.x + xn);
                mid.y = (Real)0.5 * (zo.y + yn);
                P[i] = mid;
                if (has_pad) P[pad_j] = mid;
                V2 zo_next = zo;
                Rnd rn_next = rn;
                if (t + 1 < cnt) {
                    zo_next = P[i + 1];
                    rn_next = rnd[t + 1];
                }
                Real lsum = 0;
                Real pn = 1, pd = 1;
                int ebal = 0;
                // FP32: |z' - q|^2 = |b|^2 + |d|^2 + 2 d.b with b = z - q, d = z' - z (9 VALU per
                // pair instead of 10); FP64 takes both distances directly
                const Real dx = xn - zo.x, dy = yn - zo.y;
                const Real dd = mad(dx, dx, dy * dy), tdx = dx + dx, tdy = dy + dy;
                auto pair = [&](const V2 q) {
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
                };
                constexpr int C = LogOps<Real>::chunk;
                int s = 0;
                for (; s + C <= S; s += C) {
                    constexpr int CL = C < 8 ? C : 8;   // slots whose loads issue together
#pragma unroll
                    for (int u0 = 0; u0 < C; u0 += CL) {
                        V2 q[CL];
#pragma unroll
                        for (int u = 0; u < CL; ++u) q[u] = P[lane + G * (s + u0 + u)];
#pragma unroll
                        for (int u = 0; u < CL; ++u) pair(q[u]);
                    }
                    if constexpr (sizeof(Real) == 4) {
                        // keep the products in range by moving their exponents to an integer
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
                // remainder (< C slots): blocks of 4, then 2 and 1, whose loads issue together
                for (; s + 4 <= S; s += 4) {
                    V2 q[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) q[u] = P[lane + G * (s + u)];
#pragma unroll
                    for (int u = 0; u < 4; ++u) pair(q[u]);
                }
#pragma unroll
                for (int blk = 2; blk >= 1; blk >>= 1) {
                    if (S - s >= blk) {
                        V2 q[2];
#pragma unroll
                        for (int u = 0; u < blk; ++u) q[u] = P[lane + G * (s + u)];
#pragma unroll
                        for (int u = 0; u < blk; ++u) pair(q[u]);
                        s += blk;
                    }
                }
                if constexpr (sizeof(Real) == 4)
                    lsum = __builtin_amdgcn_logf(pn) - __builtin_amdgcn_logf(pd) + (float)ebal;
                else
                    lsum += LogOps<Real>::lg(pn) - LogOps<Real>::lg(pd);
                lsum = group_sum<G, Real>(lsum);
                Real logA = mad((Real)-0.5, mad(xn, xn, yn * yn) - mad(zo.x, zo.x, zo.y * zo.y), mfac * lsum);
                double2 prop[kMaxFactors];
                double lm_new = 0.0;
                if constexpr (ROW) {
                    lm_new = row_propose(row, (double)zo.x, (double)zo.y, (double)xn, (double)yn, a.R, prop);
                    logA = (Real)((double)logA + (lm_new - row.lm));
                }
                Real mqr[MixLane<G, Real>::M], mqi[MixLane<G, Real>::M];
                Real lw_new = 0;
                if constexpr (MIX) {
                    lw_new = mix_propose(mix, zo.x, zo.y, xn, yn, a.R, lane, mqr, mqi);
                    logA += lw_new - mix.lw;
                }
                bool acc;
                if constexpr (sizeof(Real) == 4) {
                    acc = (logA >= 0.f) || (lnu_move < logA);
                } else {
                    acc = (logA >= 0.0) || (log(u01d(ub)) < logA);
                }
                if constexpr (ROW) {
                    if (acc) {
#pragma unroll
                        for (int q = 0; q < kMaxFactors; ++q) row.pk[q] = prop[q];
                        row.lm = lm_new;
                    }
                }
                if constexpr (MIX) {
                    if (acc) {
#pragma unroll
                        for (int j = 0; j < MixLane<G, Real>::M; ++j) {
                            mix.pr[j] = mqr[j];
                            mix.pi[j] = mqi[j];
                        }
                        mix.lw = lw_new;
                    }
                }
                V2 nz;
                nz.x = acc ? xn : zo.x;
                nz.y = acc ? yn : zo.y;
                P[i] = nz;
                accepted += acc ? 1u : 0u;
                zo = zo_next;
                rn = rn_next;
            }
            __builtin_amdgcn_wave_barrier();   // rnd is rewritten by the next batch
        }
    }
    for (int j = lane; j < N; j += G) gpos[j] = P[j];
    if (lane == 0) a.accepted[walker] += accepted;
    if constexpr (ROW) {
        if (lane == 0)
#pragma unroll
            for (int q = 0; q < kMaxFactors; ++q) a.row_pk[(size_t)walker * kMaxFactors + q] = row.pk[q];
    }
}

// ---------------------------------------------------------------------------
// The same sweeps with each lane's particles in registers instead of LDS: lane l of a group
// holds particles j = l + G s, s < S, in px[s], py[s] (S = ceil(N / G) fixed at compile time).
// Moves run in the order i = G s + t; with the loop over s unrolled, the slot of the moving
// particle is a compile-time index in its owner lane t, which holds the midpoint there during
// the pair loop.  The old position and the random numbers of move i come from lane t by
// __shfl.  Products, chunking, pad slots and the acceptance follow k_sweep, so for the same G
// the chain is bit-identical to it.  The register file holds more walkers per CU than the LDS,
// so a launch needs fewer rounds of resident workgroups.
// ---------------------------------------------------------------------------

// register budget per lane: positions (2 S values, doubled in FP64) plus ~40 for the move
template <int G, int S, typename Real>
constexpr int reg_sweep_waves()
{
    return (2 * S * (int)(sizeof(Real) / 4) + 40) <= 64 ? 8 : (2 * S * (int)(sizeof(Real) / 4) + 40) <= 84 ? 6
         : (2 * S * (int)(sizeof(Real) / 4) + 40) <= 100 ? 5 : (2 * S * (int)(sizeof(Real) / 4) + 40) <= 128 ? 4 : 3;
}

// The sweeps of the 64 / G walkers of wavefront unit `unit` (walkers unit * 64 / G ..), lane
// `tid` (0..63) of the wavefront: the body of k_sweep_reg, shared with the fused step kernel.
// EXACT: S = ceil(N / G).  Otherwise the walkers use the first SA = ceil(N / G) <= S slots
// and the rest are skipped (wave-uniform tests): the moves, pair order, chunks and pads are
// those of SA slots, so the chain is still the LDS kernel's.
template <int G, int S, typename Real, int MIXK = 0, bool EXACT = true>   // MIXK: 0, or the largest mixture k held
__device__ __forceinline__ void sweep_reg_walkers(const SweepArgs& a, long unit, int tid)
{
    constexpr bool MIX = MIXK > 0;
    constexpr int KM = MIX ? MIXK : 1;
    using V2 = typename V2T<Real>::T;
    constexpr int WPB = 64 / G;
    constexpr int C = LogOps<Real>::chunk;
    constexpr int FULL = (S / C) * C;   // slots covered by whole chunks
    const int N = a.N;                  // (SA - 1) G < N <= SA G, SA <= S, checked by the launcher
    const int SA = EXACT ? S : (N + G - 1) / G;
    const int FULLA = EXACT ? FULL : (SA / C) * C;
    const int lane = tid % G;
    const int grp = tid / G;
    const long wl = unit * WPB + grp;
    if (wl >= a.walkers) return;        // whole groups leave; shuffles stay within a group
    const int walker = (int)wl;
    V2* gpos = reinterpret_cast<V2*>(a.pos) + (size_t)walker * N;
    Real px[S], py[S];
    static_for(std::make_integer_sequence<int, S>{}, [&](auto sc) {
        constexpr int s = decltype(sc)::value;
        const int j = lane + G * s;
        const V2 q = j < N ? gpos[j] : gpos[0];
        px[s] = q.x;
        py[s] = q.y;
    });
    const bool pad_last = lane + G * (SA - 1) >= N;   // the last slot used holds no particle

    const Real step = (Real)a.step;
    const Real mfac = (Real)a.m * LogOps<Real>::to_ln;
    const unsigned long long sweep_first = a.sweep0 + (a.sweep_base ? *a.sweep_base : 0ull);
    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
    unsigned accepted = 0;
    MixLane<G, Real, KM> mix;
    if constexpr (MIX) mix_load(mix, a, walker, lane);

    for (int sw = 0; sw < a.nsweeps; ++sw) {
        const unsigned long long sidx = sweep_first + (unsigned long long)sw;
        const uint32_t c1 = (uint32_t)sidx;
        const uint32_t c2 = (uint32_t)((sidx >> 32) & 0xffffu);
        static_for(std::make_integer_sequence<int, S>{}, [&](auto sc) {
            constexpr int s = decltype(sc)::value;
            if (!EXACT && s >= SA) return;
            const int base = G * s;
            // opaque counter: keeps the compiler from hoisting the sweep-invariant part of
            // Philox for every slot out of the sweep loop (S x 3 live registers)
            uint32_t ctr = (uint32_t)(base + lane);
            asm volatile("" : "+v"(ctr));
            uint32_t r[4];
            philox(ctr, c1, c2, (uint32_t)walker, k0, k1, r);
            Real g1, g2;
            gauss2(r[0], r[1], g1, g2);
            float lnu = 0.f;
            if constexpr (sizeof(Real) == 4) lnu = __builtin_amdgcn_logf(u01f(r[2])) * 0.69314718055994531f;
            const uint32_t ub = r[2];
            const int cnt = min(G, N - base);
            for (int t = 0; t < cnt; ++t) {
                const Real zox = __shfl(px[s], t, G), zoy = __shfl(py[s], t, G);
                const Real d1 = __shfl(g1, t, G), d2 = __shfl(g2, t, G);
                const Real xn = mad(step, d1, zox), yn = mad(step, d2, zoy);
                V2 mid;
                mid.x = (Real)0.5 * (zox + xn);
                mid.y = (Real)0.5 * (zoy + yn);
                const bool own = lane == t;
                if (own) {
                    px[s] = mid.x;
                    py[s] = mid.y;
                }
                Real lsum = 0;
                Real pn = 1, pd = 1;
                int ebal = 0;
                const Real dx = xn - zox, dy = yn - zoy;
                const Real dd = mad(dx, dx, dy * dy), tdx = dx + dx, tdy = dy + dy;
                static_for(std::make_integer_sequence<int, S>{}, [&](auto uc) {
                    constexpr int u = decltype(uc)::value;
                    if (!EXACT && u >= SA) return;
                    Real qx = px[u], qy = py[u];
                    if constexpr (EXACT) {
                        if constexpr (u == S - 1) {
                            if (pad_last) {          // pads hold the midpoint
                                qx = mid.x;
                                qy = mid.y;
                            }
                        }
                    } else if (u == SA - 1 && pad_last) {
                        qx = mid.x;
                        qy = mid.y;
                    }
                    const Real bx = zox - qx, by = zoy - qy;
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
                        if (!EXACT && u + 1 > FULLA) return;
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
                lsum = group_sum<G, Real>(lsum);
                Real logA = mad((Real)-0.5, mad(xn, xn, yn * yn) - mad(zox, zox, zoy * zoy), mfac * lsum);
                Real mqr[MixLane<G, Real, KM>::M], mqi[MixLane<G, Real, KM>::M];
                Real lw_new = 0;
                if constexpr (MIX) {
                    lw_new = mix_propose(mix, zox, zoy, xn, yn, a.R, lane, mqr, mqi);
                    logA += lw_new - mix.lw;
                }
                bool acc;
                if constexpr (sizeof(Real) == 4) {
                    acc = (logA >= 0.f) || (__shfl(lnu, t, G) < logA);
                } else {
                    acc = (logA >= 0.0) || (log(u01d((uint32_t)__shfl((int)ub, t, G))) < logA);
                }
                if (own) {
                    px[s] = acc ? xn : zox;
                    py[s] = acc ? yn : zoy;
                }
                if constexpr (MIX) {
                    if (acc) {
#pragma unroll
                        for (int j = 0; j < MixLane<G, Real, KM>::M; ++j) {
                            mix.pr[j] = mqr[j];
                            mix.pi[j] = mqi[j];
                        }
                        mix.lw = lw_new;
                    }
                }
                accepted += acc ? 1u : 0u;
            }
        });
    }
    static_for(std::make_integer_sequence<int, S>{}, [&](auto sc) {
        constexpr int s = decltype(sc)::value;
        const int j = lane + G * s;
        if (j < N) {
            V2 q;
            q.x = px[s];
            q.y = py[s];
            gpos[j] = q;
        }
    });
    if (lane == 0) a.accepted[walker] += accepted;
}

template <int G, int S, typename Real, int MIXK = 0, bool EXACT = true>
__global__ __launch_bounds__(64) void k_sweep_reg(SweepArgs a)
{
    sweep_reg_walkers<G, S, Real, MIXK, EXACT>(a, blockIdx.x, threadIdx.x);
}

// ---------------------------------------------------------------------------
// Exact symmetry moves of |Psi_L|^2, one wavefront per walker.  With Z = sum z_i / N and
// w_i = z_i - Z, sum |z_i|^2 = N |Z|^2 + sum |w_i|^2 and the Jastrow factor depends on the w_i
// only, so |Psi_L|^2 = exp(-N |Z|^2 / 2) x f(w), and three group moves leave it invariant:
//  - centre of mass: Z is an independent Gaussian of variance 1/N per component, redrawn from
//    it (heat bath);
//  - dilation of the relative coordinates, w -> lambda w: f(lambda w) = lambda^(2 L0) f(w)
//    exp(-(lambda^2 - 1) W / 2) with W = sum |w_i|^2 and L0 = m N (N - 1) / 2.  The generalized
//    Gibbs step for a group (Liu and Sabatti, JASA 95 (2000) 1224) draws lambda from
//    pi(lambda w) lambda^(2N - 2) d lambda / lambda, i.e. t = lambda^2 W / 2 ~ Gamma(L0 + N - 1);
//    the Gamma variate comes from Marsaglia and Tsang's rejection method (ACM TOMS 26 (2000)
//    363), at most 32 attempts (each accepts with probability > 0.95 for shape >= 1; if all
//    fail, lambda = 1);
//  - rotation z -> e^{i phi} z with uniform phi.
// Together: z_i -> e^{i phi} (Z' + lambda (z_i - Z)), accepted with probability 1.
// Random numbers: Philox block (walker, sweep_lo, sweep_hi16 | tag 4 << 16, 0xfffffff0) for Z'
// and phi, and (walker, sweep_lo, sweep_hi16 | tag 5 << 16, attempt) for the Gamma variate.
// All lanes hold the same butterfly sums and draw the same numbers, so they agree on lambda.
// ---------------------------------------------------------------------------
template <typename Real>
__global__ __launch_bounds__(64) void k_symmetry(SweepArgs a)
{
    using V2 = typename V2T<Real>::T;
    const int walker = blockIdx.x;
    if (walker >= a.walkers) return;
    const int lane = threadIdx.x;
    const int N = a.N;
    V2* P = reinterpret_cast<V2*>(a.pos) + (size_t)walker * N;
    double sx = 0.0, sy = 0.0, sr = 0.0;
    for (int j = lane; j < N; j += 64) {
        const V2 q = P[j];
        const double x = (double)q.x, y = (double)q.y;
        sx += x;
        sy += y;
        sr += x * x + y * y;
    }
    sx = group_sum<64, double>(sx);
    sy = group_sum<64, double>(sy);
    sr = group_sum<64, double>(sr);
    const double cx = sx / N, cy = sy / N;
    const double W = sr - (sx * cx + sy * cy);
    const unsigned long long sidx = a.sweep0 + (a.sweep_base ? *a.sweep_base : 0ull);
    const uint32_t c1 = (uint32_t)sidx, c2 = (uint32_t)((sidx >> 32) & 0xffffu);
    uint32_t r[4];
    philox((uint32_t)walker, c1, c2 | (4u << 16), 0xfffffff0u, (uint32_t)a.seed, (uint32_t)(a.seed >> 32), r);
    double g1, g2;
    gauss2(r[0], r[1], g1, g2);
    const double sd = 1.0 / sqrt((double)N);
    const double zx = g1 * sd, zy = g2 * sd;                        // Z'
    double sphi, cphi;
    sincos(6.283185307179586 * u01d(r[2]), &sphi, &cphi);
    double lam = 1.0;
    if (a.dilate && N >= 2 && W > 0.0) {
        const double shape = 0.5 * a.m * (double)N * (N - 1) + (N - 1);
        const double d = shape - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
        for (uint32_t att = 0; att < 32; ++att) {
            philox((uint32_t)walker, c1, c2 | (5u << 16), att, (uint32_t)a.seed, (uint32_t)(a.seed >> 32), r);
            double x, unused;
            gauss2(r[0], r[1], x, unused);
            double v = 1.0 + c * x;
            if (v <= 0.0) continue;
            v = v * v * v;
            const double u = u01d(r[2]);
            if (log(u) < 0.5 * x * x + d * (1.0 - v + log(v))) {
                lam = sqrt(2.0 * d * v / W);
                break;
            }
        }
    }
    for (int j = lane; j < N; j += 64) {
        const V2 q = P[j];
        const double x = zx + lam * ((double)q.x - cx), y = zy + lam * ((double)q.y - cy);
        V2 o;
        o.x = (Real)(cphi * x - sphi * y);
        o.y = (Real)(sphi * x + cphi * y);
        P[j] = o;
    }
}

// Exact moves of the mixture density w |Psi_L|^2, 16 lanes per walker (4 walkers per wavefront).
// Translating every particle by d leaves the Jastrow factor unchanged and multiplies |Psi_L|^2
// by exp(-(2 d.sum z + N |d|^2) / 2), so a random-walk proposal z -> z + d, d Gaussian of variance
// 4 / N per component (twice the spread of the centre of mass under |Psi_L|^2), is accepted with
// that factor times w(z + d) / w(z); then a rotation z -> e^{i phi} z with uniform phi, which
// leaves every |p_k| and so w unchanged.  The walkers then cross between the bulk and the |p_1|
// tail terms of w in a few moves instead of diffusing there by single-particle steps.  Two
// passes over the particles (lane l takes j = l + 16 t) sum the FP64 power sums of z and of
// z + d (one butterfly per p_k); lane k evaluates term k + 1 of w, so each lane takes one log
// and one exp per configuration.  Every lane of a walker holds the same sums and takes the same
// decision.  Random numbers: Philox block (walker, sweep_lo, sweep_hi16 | tag 6 << 16,
// 0xfffffff1).
constexpr int kMixLanes = 16;

template <typename Real, int KM>
__device__ __forceinline__ double mix_ln_w_pass(const SweepArgs& a, const typename V2T<Real>::T* P, int sub, double dx,
                                                double dy, double* sx, double* sy)
{
    double pr[KM], pi[KM];
#pragma unroll
    for (int k = 0; k < KM; ++k) pr[k] = pi[k] = 0.0;
    double ax = 0.0, ay = 0.0;
    for (int j = sub; j < a.N; j += kMixLanes) {
        const auto q = P[j];
        const double x = (double)q.x + dx, y = (double)q.y + dy;
        ax += x;
        ay += y;
        const double ur = x / a.R, ui = y / a.R;
        double qr = ur, qi = ui;
#pragma unroll
        for (int k = 0; k < KM; ++k) {
            pr[k] += qr;
            pi[k] += qi;
            const double o = qr * ur - qi * ui;
            qi = qr * ui + qi * ur;
            qr = o;
        }
    }
    if (sx) {
        *sx = group_sum<kMixLanes, double>(ax);
        *sy = group_sum<kMixLanes, double>(ay);
    }
    // lane sub keeps p_{sub + 1} (and p_{sub + 17} when KM > 16) and evaluates its terms
    double t = 0.0;
#pragma unroll
    for (int base = 0; base < KM; base += kMixLanes) {
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int k = base; k < base + kMixLanes && k < KM; ++k) {
            const double sr = group_sum<kMixLanes, double>(pr[k]), si = group_sum<kMixLanes, double>(pi[k]);
            if (k - base == sub) {
                re = sr;
                im = si;
            }
        }
        const int k = base + sub;
        const double q = re * re + im * im;
        if (k < a.mix.kmix && q > 0.0) t += exp(a.mix.a[k] * log(q) + a.mix.b[k]);
    }
    return log(a.mix.c0 + group_sum<kMixLanes, double>(t));
}

template <typename Real, int KM>
__global__ __launch_bounds__(64) void k_symmetry_mix(SweepArgs a)
{
    using V2 = typename V2T<Real>::T;
    constexpr int WPB = 64 / kMixLanes;
    const int walker = blockIdx.x * WPB + (int)threadIdx.x / kMixLanes;
    const int sub = (int)threadIdx.x % kMixLanes;
    if (walker >= a.walkers) return;            // the 16 lanes of a walker leave together
    const int N = a.N;
    V2* P = reinterpret_cast<V2*>(a.pos) + (size_t)walker * N;
    const unsigned long long sidx = a.sweep0 + (a.sweep_base ? *a.sweep_base : 0ull);
    const uint32_t c1 = (uint32_t)sidx, c2 = (uint32_t)((sidx >> 32) & 0xffffu);
    uint32_t r[4];
    philox((uint32_t)walker, c1, c2 | (6u << 16), 0xfffffff1u, (uint32_t)a.seed, (uint32_t)(a.seed >> 32), r);
    double g1, g2;
    gauss2(r[0], r[1], g1, g2);
    const double sd = 2.0 / sqrt((double)N);
    const double dx = g1 * sd, dy = g2 * sd;
    double sx, sy;
    const double lw0 = mix_ln_w_pass<Real, KM>(a, P, sub, 0.0, 0.0, &sx, &sy);
    const double lw1 = mix_ln_w_pass<Real, KM>(a, P, sub, dx, dy, nullptr, nullptr);
    // N (|Z + d|^2 - |Z|^2) = 2 d.sum z + N |d|^2
    const double logA = -0.5 * (2.0 * (dx * sx + dy * sy) + N * (dx * dx + dy * dy)) + (lw1 - lw0);
    const bool acc = logA >= 0.0 || log(u01d(r[2])) < logA;
    double sphi, cphi;
    sincos(6.283185307179586 * u01d(r[3]), &sphi, &cphi);
    for (int j = sub; j < N; j += kMixLanes) {
        const V2 q = P[j];
        const double x = (double)q.x + (acc ? dx : 0.0), y = (double)q.y + (acc ? dy : 0.0);
        V2 o;
        o.x = (Real)(cphi * x - sphi * y);
        o.y = (Real)(sphi * x + cphi * y);
        P[j] = o;
    }
}

hipError_t launch_symmetry_mix(const SweepArgs& a, int prec, hipStream_t s)
{
    if (a.walkers <= 0) return hipSuccess;
    if (a.mix.kmix < 1 || a.mix.kmix > kMixMax) return hipErrorInvalidValue;
    constexpr int WPB = 64 / kMixLanes;
    dim3 grid((unsigned)((a.walkers + WPB - 1) / WPB));
    if (a.mix.kmix <= 16) {
        if (prec == kFP64) hipLaunchKernelGGL((k_symmetry_mix<double, 16>), grid, dim3(64), 0, s, a);
        else hipLaunchKernelGGL((k_symmetry_mix<float, 16>), grid, dim3(64), 0, s, a);
    } else {
        if (prec == kFP64) hipLaunchKernelGGL((k_symmetry_mix<double, kMixMax>), grid, dim3(64), 0, s, a);
        else hipLaunchKernelGGL((k_symmetry_mix<float, kMixMax>), grid, dim3(64), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_symmetry(const SweepArgs& a, int prec, hipStream_t s)
{
    if (a.walkers <= 0) return hipSuccess;
    if (prec == kFP64)
        hipLaunchKernelGGL(k_symmetry<double>, dim3((unsigned)a.walkers), dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL(k_symmetry<float>, dim3((unsigned)a.walkers), dim3(64), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Features
// ---------------------------------------------------------------------------
__device__ __forceinline__ double potential_r2(const Potential& p, double r2)
{
    double v = 0.5 * p.w * r2 + p.eps * r2 * r2;
    if (p.strength != 0.0) v += p.strength * pow(sqrt(r2) / p.R, (double)p.expo);
    return v;
}

// Features in two passes.  k_psums: 4 lanes per sample (particles j = sub + 4 s) accumulate the
// FP64 power sums p_k = sum (z / R)^k, k <= KMAX, the potential and the harmonic sums, meet
// in a 2-step DPP butterfly, and write p_k to a per-sample scratch row, the weights V and
// cos/sin sums to the weight rows, and the trace.  k_monomials: one wavefront per sample
// builds the table p_k^r (k r <= K) by FP64 binary powering, stores it in LDS at the output
// precision and writes every monomial (at most 8 table factors following a packed plan) to the
// real-embedded sample row.  The FP32 form multiplies the factors in FP32: its row is FP32
// anyway, and FP64 products made the kernel FP64-VALU bound (N = 200, dL 1..16: 0.083 ms per
// 32768 samples, ~700 FP64 instructions per sample and wave).  Splitting keeps each kernel's
// live registers small, so both run at several waves per SIMD.
constexpr int kMaxPow = 128;
constexpr int kMaxStatesLds = 4096;
constexpr int kLps = 4;          // lanes per sample in k_psums

// NH: harmonic pairs held (0, or at least a.nh); MIX: mixture weights (a.mix.kmix > 0)
// Lane `sub` of the kLps lanes of sample w (w < a.walkers; the lanes of a sample are
// consecutive and aligned): the body of k_psums, shared with the fused step kernel.
template <typename Real, int KMAX, int NH, bool MIX = false>
__device__ __forceinline__ void psums_sample(const FeatureArgs& a, long w, int sub)
{
    constexpr bool HARM = NH > 0;
    constexpr int NQ = 2 * KMAX + 1 + 2 * NH;   // Re/Im p_1..p_K, V, cos/sin sums
    constexpr int QV = 2 * KMAX;
    constexpr int QH = 2 * KMAX + 1;
    const int nh = HARM ? a.nh : 0;
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
    auto load = [&](int j, double& x, double& y) {
        if (a.pos_is_double || sizeof(Real) == 8) {
            const double2 q = reinterpret_cast<const double2*>(a.pos)[(size_t)w * a.N + j];
            x = q.x; y = q.y;
        } else {
            const float2 q = reinterpret_cast<const float2*>(a.pos)[(size_t)w * a.N + j];
            x = q.x; y = q.y;
        }
    };
    // e^{i L theta} = (z / |z|)^L by binary powering, added to the harmonic sums
    auto harmonics = [&](double x, double y, double r2) {
        if constexpr (HARM) {
            const double inv = r2 > 0.0 ? 1.0 / sqrt(r2) : 0.0;
            const double er = r2 > 0.0 ? x * inv : 1.0, ei = y * inv;
#pragma unroll
            for (int h = 0; h < NH; ++h) {
                if (h < nh) {
                    unsigned e = (unsigned)a.harm[h];
                    double br = er, bi = ei, pr = 1.0, pi = 0.0;
                    while (e) {
                        if (e & 1u) {
                            const double t = pr * br - pi * bi;
                            pi = pr * bi + pi * br;
                            pr = t;
                        }
                        e >>= 1;
                        const double t = br * br - bi * bi;
                        bi = 2.0 * br * bi;
                        br = t;
                    }
                    acc[QH + 2 * h] += pr;
                    acc[QH + 2 * h + 1] += pi;
                }
            }
        }
    };
    // Without harmonics, two particles j, j + kLps per pass: their power chains are
    // independent, so the FP64 multiplies of one fill the latency of the other's.  Each sum
    // still adds the particles' terms in the order j = sub, sub + kLps, ..., so the results are
    // those of one particle per pass, bit for bit.
    int j = sub;
    for (; !HARM && j + kLps < a.N; j += 2 * kLps) {
        double xa, ya, xb, yb;
        load(j, xa, ya);
        load(j + kLps, xb, yb);
        const double r2a = xa * xa + ya * ya, r2b = xb * xb + yb * yb;
        acc[QV] += potential_r2(a.pot, r2a);
        acc[QV] += potential_r2(a.pot, r2b);
        const double ura = xa * a.invR, uia = ya * a.invR;
        const double urb = xb * a.invR, uib = yb * a.invR;
        double qra = ura, qia = uia, qrb = urb, qib = uib;
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            acc[2 * k] += qra;
            acc[2 * k] += qrb;
            acc[2 * k + 1] += qia;
            acc[2 * k + 1] += qib;
            const double ta = qra * ura - qia * uia;
            qia = qra * uia + qia * ura;
            qra = ta;
            const double tb = qrb * urb - qib * uib;
            qib = qrb * uib + qib * urb;
            qrb = tb;
        }
        harmonics(xa, ya, r2a);
        harmonics(xb, yb, r2b);
    }
    for (; j < a.N; j += kLps) {
        double x, y;
        load(j, x, y);
        const double r2 = x * x + y * y;
        acc[QV] += potential_r2(a.pot, r2);
        const double ur = x * a.invR, ui = y * a.invR;
        double qr = ur, qi = ui;
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            acc[2 * k] += qr;
            acc[2 * k + 1] += qi;
            const double t = qr * ur - qi * ui;
            qi = qr * ui + qi * ur;
            qr = t;
        }
        harmonics(x, y, r2);
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = group_sum<kLps, double>(acc[q]);
    double* ps = a.ps + (size_t)w * a.ps_ld;
    // only Re/Im p_1..p_K: a row holds ps_ld >= 2 K values (KMAX rounds K up to 8, 16 or 24)
#pragma unroll
    for (int q = 0; q < 2 * KMAX; ++q)
        if (q % kLps == sub && q < 2 * a.K) ps[q] = acc[q];
    const long row = a.row0 + w;
    Real* W = reinterpret_cast<Real*>(a.W);
    if constexpr (MIX) {
        // mixture estimator: w = c0 + sum_k exp(a_k ln|p_k|^2 + b_k), lane sub taking k = sub + 1 + 4 r
        double t = 0.0;
#pragma unroll
        for (int r = 0; r < KMAX / kLps; ++r) {
            double re = acc[2 * kLps * r], im = acc[2 * kLps * r + 1];
#pragma unroll
            for (int c = 1; c < kLps; ++c)
                if (sub == c) {
                    re = acc[2 * (kLps * r + c)];
                    im = acc[2 * (kLps * r + c) + 1];
                }
            const int k = kLps * r + sub;
            if (k < a.mix.kmix) {
#pragma clang fp contract(off)
                const double q = re * re + im * im;
                if (q > 0.0) t += exp(a.mix.a[k] * log(q) + a.mix.b[k]);
            }
        }
        t = group_sum<kLps, double>(t);
        const double iw = 1.0 / (a.mix.c0 + t);
        if (a.mix_pk_out) {
#pragma unroll
            for (int k = 0; k < KMAX && k < kMixMax; ++k)
                if (k % kLps == sub && k < a.mix.kmix)
                    a.mix_pk_out[(size_t)w * kMixMax + k] = make_double2(acc[2 * k], acc[2 * k + 1]);
        }
        if (sub == 0) {
            W[row] = (Real)(acc[QV] * iw);
            W[(size_t)a.mix_row * a.wld + row] = (Real)iw;
            a.mix_acc[row] += iw;
        }
    } else if (sub == 0) {
        W[row] = (Real)acc[QV];
    }
    if constexpr (HARM) {
#pragma unroll
        for (int q = 0; q < 2 * NH; ++q)
            if (q % kLps == sub && q < 2 * nh) W[(size_t)(1 + q) * a.wld + row] = (Real)acc[QH + q];
    }
    if (a.trace && w < a.trace_walkers && sub == 0) {
        const long long tb = a.trace_base ? *a.trace_base : 0ll;
        float* tr = a.trace + ((size_t)(a.trace_row0 + tb) + (size_t)w) * 9;
#pragma unroll
        for (int q = 0; q < 8; ++q) tr[q] = (q / 2 < a.K) ? (float)acc[q] : 0.f;
        tr[8] = (float)acc[QV];
    }
}

template <typename Real, int KMAX, int NH, bool MIX = false>
__global__ __launch_bounds__(256) void k_psums(FeatureArgs a)
{
    const long gt = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long w = gt / kLps;
    if (w >= a.walkers) return;                  // the 4 lanes of a sample leave together
    psums_sample<Real, KMAX, NH, MIX>(a, w, (int)(gt % kLps));
}

// Monomials of sample w by one wavefront (lane 0..63): pkw (32 double2) and ptw (kMaxPow
// entries) are the wavefront's LDS scratch; the product plans come from LDS (plan_lds) or from
// global memory.  The body of k_monomials, shared with the fused step kernel.
template <typename Real>
__device__ __forceinline__ void monomials_sample(const FeatureArgs& a, long w, int lane, double2* pkw,
                                                 typename V2T<Real>::T* ptw, const unsigned long long* plan,
                                                 const uint32_t* cols, bool plan_lds)
{
    using C2 = typename V2T<Real>::T;
    const int D = a.D;
    const double* ps = a.ps + (size_t)w * a.ps_ld;
    if (lane < a.K) pkw[lane] = make_double2(ps[2 * lane], ps[2 * lane + 1]);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // power table: entry e = p_k^r, one lane per entry (binary powering)
    for (int e = lane; e < a.npw; e += 64) {
        const int k = a.pw_k[e];
        unsigned r = (unsigned)a.pw_r[e];
        const double2 b0 = pkw[k - 1];
        double br = b0.x, bi = b0.y, pr = 1.0, pi = 0.0;
        while (r) {
            if (r & 1u) {
                const double t = pr * br - pi * bi;
                pi = pr * bi + pi * br;
                pr = t;
            }
            r >>= 1;
            if (r) {
                const double t = br * br - bi * bi;
                bi = 2.0 * br * bi;
                br = t;
            }
        }
        ptw[e] = C2{(Real)pr, (Real)pi};
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const long row = a.row0 + w;
    Real* X = reinterpret_cast<Real*>(a.X) + (size_t)row * a.ldx;
#pragma unroll 2
    for (int t = lane; t < D; t += 64) {
        const unsigned long long pl = plan_lds ? plan[t] : a.prod[t];
        const uint32_t cw = plan_lds ? cols[t] : ((uint32_t)a.col_re[t] | ((uint32_t)a.col_dim[t] << 16));
        const int cnt = (int)(pl & 0xfull);
        Real vr = 1, vi = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            if (q < cnt) {
                const C2 f = ptw[(pl >> (4 + 7 * q)) & 0x7full];
                const Real tr = vr * f.x - vi * f.y;
                vi = vr * f.y + vi * f.x;
                vr = tr;
            }
        }
        const int c = (int)(cw & 0xffffu);
        X[c] = vr;
        X[c + (int)(cw >> 16)] = vi;
    }
    __builtin_amdgcn_wave_barrier();   // pk / ptab are reused by the next sample
}

template <typename Real>
__global__ __launch_bounds__(256) void k_monomials(FeatureArgs a)
{
    using C2 = typename V2T<Real>::T;                                // float2 or double2
    extern __shared__ __align__(16) unsigned char smem_mono[];
    double2* pk = reinterpret_cast<double2*>(smem_mono);             // [4][32]
    C2* ptab = reinterpret_cast<C2*>(pk + 4 * 32);                   // [4][kMaxPow]
    unsigned long long* plan = reinterpret_cast<unsigned long long*>(ptab + 4 * kMaxPow);
    const int D = a.D;
    const bool plan_lds = D <= kMaxStatesLds;
    uint32_t* cols = reinterpret_cast<uint32_t*>(plan + (plan_lds ? D : 0));
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    if (plan_lds)
        for (int t = threadIdx.x; t < D; t += blockDim.x) {
            plan[t] = a.prod[t];
            cols[t] = (uint32_t)a.col_re[t] | ((uint32_t)a.col_dim[t] << 16);
        }
    __syncthreads();
    double2* pkw = pk + wv * 32;
    C2* ptw = ptab + wv * kMaxPow;
    const int waves = gridDim.x * 4;
    for (int w = blockIdx.x * 4 + wv; w < a.walkers; w += waves)
        monomials_sample<Real>(a, w, lane, pkw, ptw, plan, cols, plan_lds);
}

// ---------------------------------------------------------------------------
// Weighted Gram tiles on MFMA.  Workgroup = 4 waves = one 64x64 output tile over
// columns [i0, i0+64) x [j0, j0+64); wave (wi, wj) owns a 32x32 quadrant as 2x2
// 16x16x4 MFMA tiles.  Samples are the reduction dimension, staged 32 at a time.
// MFMA 16x16x4 operands: A[i = l%16][k = l/16], B[k = l/16][j = l%16]; the result
// register r of lane l holds C[4 (l/16) + r][l%16] for f32 and C[4 r + l/16][l%16] for f64.
// ---------------------------------------------------------------------------
template <typename Real> struct Mfma;
template <> struct Mfma<float> {
    using acc_t = floatx4;
    __device__ static acc_t op(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
};
template <> struct Mfma<double> {
    using acc_t = doublex4;
    __device__ static acc_t op(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
};

// Loads through a pointer known to address global memory count only in vmcnt.  A load through a
// generic (flat) pointer counts in lgkmcnt too, so the next wait for an LDS read would also wait
// for it, and a prefetch issued before the LDS reads of a step would stall them.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* as_global(const T* p)
{
    return (const __attribute__((address_space(1))) T*)p;
}

__device__ __forceinline__ void load4(const float* p, float* o)
{
    const float4 a = reinterpret_cast<const float4*>(p)[0];
    o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
}
__device__ __forceinline__ void load4(const double* p, double* o)
{
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const double2 a = reinterpret_cast<const double2*>(p)[e];
        o[2 * e] = a.x;
        o[2 * e + 1] = a.y;
    }
}

// Workgroup -> (tile, row split).  The dispatcher hands consecutive workgroups to the 8 XCDs in
// turn (workgroup i to XCD i mod 8), and each XCD has its own L2.
// - xcd_spread = 8 (or any value outside 1, 2, 4): consecutive workgroups take consecutive tiles
//   of one split, so a split's tiles spread over all 8 XCDs and every XCD fetches most of each
//   sample row (FETCH_SIZE: 5.9 times the sample matrix at N = 200, dL 1..16), mostly from the
//   Infinity Cache.
// - xcd_spread = K in {1, 2, 4}: the 8 XCDs form 8 / K groups of K; split s runs on group
//   s mod (8 / K), its tile t on XCD t mod K of that group.  Each row is then fetched by K L2s,
//   while the workgroups of one XCD share fewer, hotter L2 lines.
__device__ __forceinline__ bool gram_block(const GramArgs& g, int& tile, int& split)
{
    const int id = blockIdx.x;
    const int K = g.xcd_spread;
    if (K != 1 && K != 2 && K != 4) {
        tile = id % g.ntiles;
        split = id / g.ntiles;
        return split < g.splits;
    }
    const int x = id & 7, j = id >> 3;
    const int per = (g.ntiles + K - 1) / K;
    const int sl = j / per, tq = j - sl * per;
    tile = (x % K) + K * tq;
    split = (x / K) + (8 / K) * sl;
    return tile < g.ntiles && split < g.splits;
}

long gram_grid(int ntiles, int splits, int xcd_spread)
{
    const int K = xcd_spread;
    if (K != 1 && K != 2 && K != 4) return (long)ntiles * splits;
    const long groups = 8 / K;
    return 8L * ((splits + groups - 1) / groups) * ((ntiles + K - 1) / K);
}

constexpr int kKC = 32;          // samples staged per step
constexpr int kLdsStride = 80;   // padded LDS row (elements)
constexpr int kFlushSteps = 8;   // FP32: flush into FP64 every 8 x 32 samples

// Gram tiles C_q += X_i^T diag(w_q) X_j of one 64x64 upper tile (i0, j0) over one row split,
// for the NW weights at once (w_0 = 1 for S, w_1 = V for H); 8 waves per tile.  Wave w takes
// the two 16x16 blocks the tile lists for it (one row block, two column blocks; only blocks
// holding requested entries are listed) with v_mfma_f32_16x16x4_f32, or _f64 in the FP64
// form: one row operand per weight, one column operand per block.  Waves with no blocks only
// stage; the MFMA loop is straight-line code behind one wave-uniform branch (a branch per
// block made the compiler move every MFMA out of line and wait on each LDS read).
// Samples are staged 32 at a time through LDS (row stride 80 elements): threads
// 0..255 load the row side's 32 x 64 block, threads 256..511 the column side's (idle on
// diagonal tiles), 8 elements each, one step ahead in registers while the MFMAs run.  In the
// FP32 form the row side is also stored times each weight other than 1, so the MFMA loop issues
// no multiplies (VALU slots it would take from a sampler sweep running beside it).  FP32
// accumulators are flushed into FP64 registers every 256 samples; the workgroup's FP64
// result is added to the global accumulator with FP64 atomics.  Small per-wave blocks keep
// the FP32 form at 80 VGPRs (6 waves per SIMD), which hides the staging barriers and the load
// latency: N = 200, dL 1..16, 131072 rows of whole tiles ran at 0.87 of the FP32 MFMA peak in
// isolation, against 0.62 for 4 waves of 32 x 32 blocks (2 waves per SIMD).  The FP64 form
// targets 5 waves per SIMD: at 6 its two column operands spilled inside the loop.
//
// gram_tile is the work of one workgroup (all of its threads call it): tile `tile_idx` over row
// split `split` for job `job_idx`; k_gram maps its workgroups onto it, the fused step kernel
// takes (tile, split) items from a queue.  TC = 64: 64x64 tiles, 512 threads.  TC = 32: narrow
// 64x32 tiles (rows [i0, i0 + 64), columns [j0, j0 + 32), eight 16x16 blocks, two per wave),
// 256 threads, each staging 8 row-side and 4 column-side values per step; when the column side
// lies inside the row side (j0 - i0 in [0, 32]) it is read from the row side's LDS rows.
template <typename Real, int NW, int TC = 64>
__device__ __forceinline__ void gram_tile(const GramArgs& g, int tile_idx, int split, int job_idx)
{
    static_assert(TC == 64 || TC == 32, "64- or 32-column tiles");
    constexpr bool kNarrow = TC == 32;
    using acc_t = typename Mfma<Real>::acc_t;
    constexpr bool kF32 = sizeof(Real) == 4;
    // FP32: the row side is also staged times each weight other than 1; FP64 multiplies in the
    // loop instead (its weighted copies would double the LDS and cut the resident workgroups)
    constexpr bool kPre = kF32;
    __shared__ __align__(16) Real xi[kKC][kLdsStride];
    __shared__ __align__(16) Real xj[kKC][kLdsStride];
    __shared__ __align__(16) Real xw[kPre ? NW : 1][kPre ? kKC : 1][kPre ? kLdsStride : 1];
    __shared__ Real vs[kPre ? 1 : NW][kKC];

    const Tile tile = g.tiles[tile_idx];
    const long r_begin = (long)split * g.rows_per_split;
    long r_end = r_begin + g.rows_per_split;
    if (r_end > g.rows) r_end = g.rows;
    // narrow tiles: column-side offset inside the row side's LDS rows, when it is there
    const int coff = kNarrow ? tile.j0 - tile.i0 : 0;
    const bool same = kNarrow ? (coff >= 0 && coff <= 32) : tile.i0 == tile.j0;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int l16 = lane & 15, k4 = lane >> 4;
    // this wave's two 16x16 blocks: rows ro .. +15, columns co[a] .. +15 within the tile
    const bool active = wave < tile.nwave;
    const uint32_t wb = active ? g.tiles[tile_idx].wblk[wave] : 0u;
    const int ro = (int)(wb >> 2 & 3) * 16;
    const int co[2] = {(int)(wb & 3) * 16, (int)(wb >> 8 & 3) * 16};
    const bool wr[2] = {true, !(wb >> 12 & 1)};

    acc_t acc[NW][2];
    double sh[NW][2][4];
#pragma unroll
    for (int q = 0; q < NW; ++q)
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            acc[q][a] = acc_t{0, 0, 0, 0};
#pragma unroll
            for (int r = 0; r < 4; ++r) sh[q][a][r] = 0.0;
        }

    const Real* X = reinterpret_cast<const Real*>(g.X);
    const GramJob job = g.jobs ? g.jobs[job_idx] : GramJob{g.w0, g.w1, g.C0, g.C1};
    const Real* wp[2] = {reinterpret_cast<const Real*>(job.w0), reinterpret_cast<const Real*>(job.w1)};
    // FP32 operand A of weight q: the row side itself when the weight is 1, else its weighted copy
    const Real(*aq[2])[kLdsStride];
#pragma unroll
    for (int q = 0; q < 2; ++q)
        aq[q] = (q < NW && wp[q]) ? reinterpret_cast<const Real(*)[kLdsStride]>(&xw[q < NW ? q : 0][0][0]) : xi;
    const int side = kNarrow ? 0 : tid >> 8;   // 0: row-side block, 1: column-side block
    const int rr = (tid & 255) >> 3;           // 0..31
    // columns cc .. cc + 3 and cc + 32 .. cc + 35: each store instruction of 8 lanes then fills
    // 128 contiguous bytes (all 32 banks once); 8 consecutive columns per lane hit every bank
    // twice (7.3e7 conflict cycles per launch at N = 200, dL 1..16)
    const int cc = (tid & 7) * 4;              // 0..28
    const bool loader = !(side && same);
    const long col0 = side ? tile.j0 : tile.i0;
    Real(*dst)[kLdsStride] = side ? xj : xi;
    Real buf[8] = {}, bufc[4] = {}, bw[NW] = {};
    // The prefetch loads are unconditional (a row past r_end loads row r_end - 1, and its values
    // are zeroed when staged: bin, bwin), the weights are read through global pointers, and no
    // register that a load fills is written on another branch.  The compiler then waits for
    // them only where they are staged, one step later.  A conditional zero, or a weight read
    // through a generic pointer, made it wait for the whole prefetch before the MFMA loop
    // (4% of the launch at N = 200, dL 1..16).
    bool bin = false, bwin = false;
    auto fetch = [&](long r0) {
        const long row = r0 + rr;
        bin = row < r_end;
        const long rc = bin ? row : r_end - 1;
        if (loader) {
            load4(X + (size_t)rc * g.ldx + col0 + cc, buf);
            load4(X + (size_t)rc * g.ldx + col0 + cc + 32, buf + 4);
        }
        if constexpr (kNarrow) {
            if (!same) load4(X + (size_t)rc * g.ldx + tile.j0 + cc, bufc);
        }
        if constexpr (kPre) {
            if (!side) {
#pragma unroll
                for (int q = 0; q < NW; ++q)
                    if (wp[q]) bw[q] = as_global(wp[q])[rc];
            }
        } else if (tid < NW * kKC) {
            const int q = tid / kKC, k = tid % kKC;
            const long row2 = r0 + k;
            bwin = row2 < r_end;
            if (wp[q]) bw[0] = as_global(wp[q])[bwin ? row2 : r_end - 1];
        }
    };
    if (r_begin < r_end) fetch(r_begin);
    int steps = 0;
    for (long r0 = r_begin; r0 < r_end; r0 += kKC) {
        __syncthreads();   // previous step's readers are done
        if (!bin) {
#pragma unroll
            for (int e = 0; e < 8; ++e) buf[e] = 0;
#pragma unroll
            for (int e = 0; e < 4; ++e) bufc[e] = 0;
        }
        if (loader) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                dst[rr][cc + e] = buf[e];
                dst[rr][cc + 32 + e] = buf[4 + e];
            }
        }
        if constexpr (kNarrow) {
            if (!same) {
#pragma unroll
                for (int e = 0; e < 4; ++e) xj[rr][cc + e] = bufc[e];
            }
        }
        if constexpr (kPre) {
            if (!side) {
#pragma unroll
                for (int q = 0; q < NW; ++q)
                    if (wp[q]) {
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            xw[q][rr][cc + e] = buf[e] * bw[q];
                            xw[q][rr][cc + 32 + e] = buf[4 + e] * bw[q];
                        }
                    }
            }
        } else if (tid < NW * kKC) {
            vs[tid / kKC][tid % kKC] = bwin ? (wp[tid / kKC] ? bw[0] : (Real)1) : (Real)0;
        }
        __syncthreads();
        if (r0 + kKC < r_end) fetch(r0 + kKC);
        const Real(*yj)[kLdsStride] = same ? xi : xj;
        if (active) {
#pragma unroll
            for (int kk = 0; kk < kKC; kk += 4) {
                const int kr = kk + k4;
                Real bv[2];
#pragma unroll
                for (int a = 0; a < 2; ++a) bv[a] = yj[kr][co[a] + l16 + (kNarrow && same ? coff : 0)];
#pragma unroll
                for (int q = 0; q < NW; ++q) {
                    Real av;
                    if constexpr (kPre) av = aq[q][kr][ro + l16];
                    else av = xi[kr][ro + l16] * vs[q][kr];
#pragma unroll
                    for (int a = 0; a < 2; ++a) acc[q][a] = Mfma<Real>::op(av, bv[a], acc[q][a]);
                }
            }
        }
        if (kF32 && ++steps == kFlushSteps) {
            steps = 0;
#pragma unroll
            for (int q = 0; q < NW; ++q)
#pragma unroll
                for (int a = 0; a < 2; ++a) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) sh[q][a][r] += (double)acc[q][a][r];
                    acc[q][a] = acc_t{0, 0, 0, 0};
                }
        }
    }
    double* Cs[2] = {job.C0, job.C1};
#pragma unroll
    for (int q = 0; q < NW; ++q)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                if (!active || !wr[a]) continue;
                const double v = sh[q][a][r] + (double)acc[q][a][r];
                // f32 16x16x4: reg r of lane l is row 4 (l / 16) + r; f64 16x16x4: row 4 r + l / 16
                const int orow = kF32 ? (4 * k4 + r) : (4 * r + k4);
                const long row = tile.i0 + ro + orow;
                const long col = tile.j0 + co[a] + l16;
                atomicAdd(&Cs[q][row * g.ldc + col], v);
            }
}

// 64-column tiles: 6 (FP32) or 5 (FP64) waves per SIMD; 32-column tiles stage 12 values per
// thread and hold 40 KB of LDS per 4 waves: 4 (FP32) or 3 (FP64)
template <typename Real, int NW, int TC = 64>
__global__ __launch_bounds__(TC == 64 ? 512 : 256)
__attribute__((amdgpu_waves_per_eu(TC == 64 ? (sizeof(Real) == 4 ? 6 : 5) : (sizeof(Real) == 4 ? 4 : 3)))) void k_gram(GramArgs g)
{
    int tile_idx, split;
    if (!gram_block(g, tile_idx, split)) return;
    gram_tile<Real, NW, TC>(g, tile_idx, split, blockIdx.y);
}

// ---------------------------------------------------------------------------
// BF16 split Gram tiles.  Each staged FP32 value x (times the weight on the row side) is
// split into NS bf16 terms, t0 = bf16(x), t1 = bf16(x - t0), t2 = bf16(x - t0 - t1) (round to
// nearest even; every residual is exact in FP32), and a tile accumulates the products
// t_a.t_b' with a + b < NS on v_mfma_f32_16x16x32_bf16.
// - NS = 2 (hi/lo, three products): each product carries a relative error of order 2^-16
//   (the residual x - hi - lo and the dropped lo.lo'); on heavy-tailed monomials it averages
//   down slowly: N = 200, dL 1..16, 8192 samples give 5e-6 relative (Frobenius) against FP64,
//   3.8e-6 with lo.lo' added.
// - NS = 3 (hi/mid/lo, six products): the dropped terms are of order 2^-24, that of the FP32
//   operands themselves.
// One MFMA covers the 32 staged
// samples.  Lane l supplies row (or column) l % 16 with the 8 samples 8 (l / 16) .. +7 of
// both operands; as both use the same lane-to-sample map, the sum over samples is the same
// whatever order the hardware takes them in.  The result layout is that of the FP32
// 16x16x4 MFMA (register r of lane l = row 4 (l / 16) + r, column l % 16).
// LDS keeps [column][sample] with a row of 40 bf16 (80 bytes), so the 16-byte fragment
// reads and the staging writes of 16 lanes hit distinct banks.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
constexpr int kLdsB = 40;

template <int NS>
__device__ __forceinline__ void split8(const float* x, bf16x8 (&t)[NS])
{
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        float r = x[e];
#pragma unroll
        for (int a = 0; a < NS; ++a) {
            const __bf16 h = (__bf16)r;
            t[a][e] = h;
            r -= (float)h;
        }
    }
}

template <int NW, int NS>
__global__ __launch_bounds__(512) void k_gram_bf16(GramArgs g)
{
    __shared__ __align__(16) __bf16 ai[NW][NS][64][kLdsB];   // row side: weight q, split term
    __shared__ __align__(16) __bf16 bj[NS][64][kLdsB];       // column side: split term

    int tile_idx, split;
    if (!gram_block(g, tile_idx, split)) return;
    const Tile tile = g.tiles[tile_idx];
    const long r_begin = (long)split * g.rows_per_split;
    long r_end = r_begin + g.rows_per_split;
    if (r_end > g.rows) r_end = g.rows;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int l16 = lane & 15, k4 = lane >> 4;
    // this wave's two 16x16 blocks: rows ro .. +15, columns co[a] .. +15 within the tile
    const bool active = wave < tile.nwave;
    const uint32_t wb = active ? g.tiles[tile_idx].wblk[wave] : 0u;
    const int ro = (int)(wb >> 2 & 3) * 16;
    const int co[2] = {(int)(wb & 3) * 16, (int)(wb >> 8 & 3) * 16};
    const bool wr[2] = {true, !(wb >> 12 & 1)};

    floatx4 acc[NW][2];
    double sh[NW][2][4];
#pragma unroll
    for (int q = 0; q < NW; ++q)
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            acc[q][a] = floatx4{0, 0, 0, 0};
#pragma unroll
            for (int r = 0; r < 4; ++r) sh[q][a][r] = 0.0;
        }

    const float* X = reinterpret_cast<const float*>(g.X);
    const GramJob job = g.jobs ? g.jobs[blockIdx.y] : GramJob{g.w0, g.w1, g.C0, g.C1};
    const float* wp[2] = {reinterpret_cast<const float*>(job.w0), reinterpret_cast<const float*>(job.w1)};
    // staging: waves 0..3 take the row side (times each weight), waves 4..7 the column side;
    // thread = (column lane, 8 consecutive samples 8 (wave % 4) .. +7), loaded into registers
    // one step ahead while the MFMAs run on the current LDS tile
    const int side = wave >> 2;
    const int sc = lane, sk = (wave & 3) * 8;
    const long col = (side ? tile.j0 : tile.i0) + sc;
    float x[8], wv[NW][8];
    auto fetch = [&](long r0) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const long row = r0 + sk + e;
            const bool in = row < r_end;
            x[e] = in ? X[(size_t)row * g.ldx + col] : 0.f;
            if (!side) {
#pragma unroll
                for (int q = 0; q < NW; ++q) wv[q][e] = (in && wp[q]) ? wp[q][row] : 1.f;
            }
        }
    };
    if (r_begin < r_end) fetch(r_begin);
    int steps = 0;
    for (long r0 = r_begin; r0 < r_end; r0 += kKC) {
        bf16x8 t[NS];
        __syncthreads();   // the previous step's fragments have been read
        if (!side) {
#pragma unroll
            for (int q = 0; q < NW; ++q) {
                float y[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) y[e] = x[e] * wv[q][e];
                split8<NS>(y, t);
#pragma unroll
                for (int a = 0; a < NS; ++a) *reinterpret_cast<bf16x8*>(&ai[q][a][sc][sk]) = t[a];
            }
        } else {
            split8<NS>(x, t);
#pragma unroll
            for (int a = 0; a < NS; ++a) *reinterpret_cast<bf16x8*>(&bj[a][sc][sk]) = t[a];
        }
        __syncthreads();
        if (r0 + kKC < r_end) fetch(r0 + kKC);

        if (active) {
            // column-side fragments once for both weights
            bf16x8 bv[2][NS];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int v = 0; v < NS; ++v)
                    bv[a][v] = *reinterpret_cast<const bf16x8*>(&bj[v][co[a] + l16][8 * k4]);
#pragma unroll
            for (int q = 0; q < NW; ++q) {
                const int rr = ro + l16;
                bf16x8 av[NS];
#pragma unroll
                for (int u = 0; u < NS; ++u) av[u] = *reinterpret_cast<const bf16x8*>(&ai[q][u][rr][8 * k4]);
#pragma unroll
                for (int a = 0; a < 2; ++a) {
                    floatx4 d = acc[q][a];
                    // smallest products first
#pragma unroll
                    for (int v = NS - 1; v >= 0; --v)
#pragma unroll
                        for (int u = NS - 1 - v; u >= 0; --u)
                            d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[u], bv[a][v], d, 0, 0, 0);
                    acc[q][a] = d;
                }
            }
        }
        if (++steps == kFlushSteps) {
            steps = 0;
#pragma unroll
            for (int q = 0; q < NW; ++q)
#pragma unroll
                for (int a = 0; a < 2; ++a) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) sh[q][a][r] += (double)acc[q][a][r];
                    acc[q][a] = floatx4{0, 0, 0, 0};
                }
        }
    }
    double* Cs[2] = {job.C0, job.C1};
#pragma unroll
    for (int q = 0; q < NW; ++q)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                if (!active || !wr[a]) continue;
                const double v = sh[q][a][r] + (double)acc[q][a][r];
                const long row = tile.i0 + ro + 4 * k4 + r;
                const long cl = tile.j0 + co[a] + l16;
                atomicAdd(&Cs[q][row * g.ldc + cl], v);
            }
}

// ---------------------------------------------------------------------------
// Many weights on one small block (the occupation numbers: one weight per orbital and real
// or imaginary part, all on the block between the left and right copies of one sector).  As
// a product over samples, C[q][(i, j)] = sum_s W[q][s] P[(i, j)][s] with the pair products
// P = X[s][r0 + i] X[s][c0 + j]: weights on the MFMA rows, pairs on the columns.  One
// 512-thread workgroup per 64-weight x 64-pair tile and row split; wave w owns weights
// 32 (w / 4) .. +31 (two 16x16 blocks) and pairs 16 (w % 4) .. +15.  Per 32-sample step the
// workgroup stages the weights ([sample][weight]) and the 64 + 64 columns of X it reads in
// LDS; a lane forms its pair product when it reads the MFMA column operand.  Every weight
// and X value is fetched once per pair tile, where one Gram launch per weight pair would
// fetch a 64-column X tile per weight pair.  FP32 accumulators are flushed into FP64
// registers every 256 samples, and the result is added to the weights' FP64 accumulators.
// ---------------------------------------------------------------------------
template <typename Real>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(sizeof(Real) == 4 ? 6 : 4))) void k_gram_weights(WeightGramArgs g)
{
    using acc_t = typename Mfma<Real>::acc_t;
    constexpr bool kF32 = sizeof(Real) == 4;
    __shared__ __align__(16) Real wt[kKC][kLdsStride];   // [sample][weight]
    __shared__ __align__(16) Real xr[kKC][kLdsStride];   // [sample][row column i]
    __shared__ __align__(16) Real xc[kKC][kLdsStride];   // [sample][column column j]

    const int P = g.nr * g.nc;
    const int nwt = (g.nweights + 63) / 64;
    const int tiles = nwt * ((P + 63) / 64);
    const int tile = (int)(blockIdx.x % (unsigned)tiles);
    const int split = (int)(blockIdx.x / (unsigned)tiles);
    if (split >= g.splits) return;
    const int q0 = (tile % nwt) * 64, p0 = (tile / nwt) * 64;
    const long r_begin = (long)split * g.rows_per_split;
    long r_end = r_begin + g.rows_per_split;
    if (r_end > g.rows) r_end = g.rows;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wi = wave >> 2, wj = wave & 3;
    const int l16 = lane & 15, k4 = lane >> 4;
    const bool active = q0 + 32 * wi < g.nweights && p0 + 16 * wj < P;   // wave-uniform
    const int pj = p0 + 16 * wj + l16;                                   // this lane's pair
    const bool pv = pj < P;
    const int pi_ = pv ? pj / g.nc : 0, pc = pv ? pj % g.nc : 0;

    acc_t acc[2];
    double sh[2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        acc[a] = acc_t{0, 0, 0, 0};
#pragma unroll
        for (int r = 0; r < 4; ++r) sh[a][r] = 0.0;
    }
    const Real* X = reinterpret_cast<const Real*>(g.X);
    const Real* W = reinterpret_cast<const Real*>(g.W);
    // staging: weight q0 + (tid >> 3), samples 4 (tid & 7) .. +3; X sample tid >> 4, columns
    // 8 (tid & 7) .. +7 of the row side (tid & 8 == 0) or of the column side
    const int sq = tid >> 3, ss = (tid & 7) * 4;
    const int xs = tid >> 4, xcol = (tid & 7) * 8;
    const bool xside = (tid & 8) != 0;
    const int xn = xside ? g.nc : g.nr;
    const long xbase = xside ? g.c0 : g.r0;
    const bool wq_ok = q0 + sq < g.nweights;
    const Real* wrow = W + (size_t)(wq_ok ? q0 + sq : 0) * g.wld;
    Real wb[4], xb[8];
    auto fetch = [&](long r0) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const long row = r0 + ss + e;
            wb[e] = (wq_ok && row < r_end) ? wrow[row] : (Real)0;
        }
        const long row = r0 + xs;
        const Real* xp = X + (size_t)(row < r_end ? row : 0) * g.ldx + xbase;
#pragma unroll
        for (int e = 0; e < 8; ++e) xb[e] = (row < r_end && xcol + e < xn) ? xp[xcol + e] : (Real)0;
    };
    if (r_begin < r_end) fetch(r_begin);
    int steps = 0;
    for (long r0 = r_begin; r0 < r_end; r0 += kKC) {
        __syncthreads();   // previous step's readers are done
#pragma unroll
        for (int e = 0; e < 4; ++e) wt[ss + e][sq] = wb[e];
        Real(*xd)[kLdsStride] = xside ? xc : xr;
#pragma unroll
        for (int e = 0; e < 8; ++e) xd[xs][xcol + e] = xb[e];
        __syncthreads();
        if (r0 + kKC < r_end) fetch(r0 + kKC);
        if (active) {
#pragma unroll
            for (int kk = 0; kk < kKC; kk += 4) {
                const int kr = kk + k4;
                const Real bv = pv ? xr[kr][pi_] * xc[kr][pc] : (Real)0;
#pragma unroll
                for (int a = 0; a < 2; ++a) acc[a] = Mfma<Real>::op(wt[kr][wi * 32 + a * 16 + l16], bv, acc[a]);
            }
        }
        if (kF32 && ++steps == kFlushSteps) {
            steps = 0;
#pragma unroll
            for (int a = 0; a < 2; ++a) {
#pragma unroll
                for (int r = 0; r < 4; ++r) sh[a][r] += (double)acc[a][r];
                acc[a] = acc_t{0, 0, 0, 0};
            }
        }
    }
    if (!active || !pv) return;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            // f32 16x16x4: reg r of lane l is row 4 (l / 16) + r; f64 16x16x4: row 4 r + l / 16
            const int q = q0 + wi * 32 + a * 16 + (kF32 ? (4 * k4 + r) : (4 * r + k4));
            if (q >= g.nweights) continue;
            const double v = sh[a][r] + (double)acc[a][r];
            double* C = (q & 1) ? g.jobs[q >> 1].C1 : g.jobs[q >> 1].C0;
            atomicAdd(&C[(long)(g.r0 + pi_) * g.ldc + g.c0 + pc], v);
        }
}

// ---------------------------------------------------------------------------
// FP64 tiles -> complex sector blocks: S_ab = (C[a][b] + C[d+a][d+b]) + i (C[a][d+b] - C[d+a][b])
// (an entry p > q is read as its mirror q < p, so only the upper triangle is ever read and
// the extracted S and H are exactly Hermitian)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double cget(const double* C, long ldc, long p, long q)
{
    return (p <= q) ? C[p * ldc + q] : C[q * ldc + p];
}

__global__ void k_extract(ExtractArgs e)
{
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= e.total) return;
    int lo = 0, hi = e.nitems - 1;
    while (lo < hi) {   // last item with out_off <= g
        const int mid = (lo + hi + 1) >> 1;
        if (e.items[mid].out_off <= g) lo = mid; else hi = mid - 1;
    }
    const ExtractItem it = e.items[lo];
    const long loc = g - it.out_off;
    const int a = (int)(loc / it.d2), b = (int)(loc % it.d2);
    const long pa = it.c1 + a, qa = it.c1 + it.d1 + a;
    const long pb = it.c2 + b, qb = it.c2 + it.d2 + b;
    const long ldc = e.ldc;
    const double re = cget(it.C, ldc, pa, pb) + cget(it.C, ldc, qa, qb);
    const double im = cget(it.C, ldc, pa, qb) - cget(it.C, ldc, qa, pb);
    e.out[2 * g] = re * e.scale;
    e.out[2 * g + 1] = im * e.scale;
}

__global__ void k_set_bases(unsigned long long* b, unsigned long long sweep_base, long long trace_base)
{
    if (threadIdx.x == 0) {
        b[0] = sweep_base;
        b[1] = (unsigned long long)trace_base;
    }
}

__global__ void k_zero(double* p, long n)
{
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; g < n; g += stride) p[g] = 0.0;
}

// ---------------------------------------------------------------------------
// FP64 generalized eigensolver, one workgroup (256 threads) per problem.
// ---------------------------------------------------------------------------
struct cd {
    double r, i;
};
__device__ __forceinline__ cd cmul(cd a, cd b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ __forceinline__ cd cmulc(cd a, cd b) { return {a.r * b.r + a.i * b.i, a.i * b.r - a.r * b.i}; }  // a conj(b)
__device__ __forceinline__ cd cconj(cd a) { return {a.r, -a.i}; }
__device__ __forceinline__ cd csub(cd a, cd b) { return {a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ cd cadd(cd a, cd b) { return {a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ cd cscale(cd a, double s) { return {a.r * s, a.i * s}; }

constexpr int kEigThreads = 1024;
constexpr int kEigMaxPairs = 320;     // d <= 640

__device__ double block_sum(double v, double* red)
{
    const int tid = threadIdx.x;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < kEigThreads / 64; ++w) t += red[w];
    return t;
}

// Cholesky S = L L^H (lower), right-looking: column j is scaled, then the trailing lower
// triangle takes the rank-1 update with column j staged in LDS (coalesced row reads).
// A holds S on entry (full); on return its lower triangle holds L.  false if S is not
// positive definite (uniform across the block).
__device__ bool chol_lower(cd* A, int d, cd* col, double* sh)
{
    const int tid = threadIdx.x;
    for (int j = 0; j < d; ++j) {
        __syncthreads();
        if (tid == 0) {
            const double a = A[(size_t)j * d + j].r;
            *sh = a > 0.0 ? sqrt(a) : -1.0;
        }
        __syncthreads();
        const double ljj = *sh;
        if (!(ljj > 0.0)) return false;
        const int n = d - j - 1;
        const double inv = 1.0 / ljj;
        for (int i = tid; i < n; i += kEigThreads) {
            const cd v = cscale(A[(size_t)(j + 1 + i) * d + j], inv);
            A[(size_t)(j + 1 + i) * d + j] = v;
            col[i] = v;
        }
        if (tid == 0) A[(size_t)j * d + j] = {ljj, 0.0};
        __syncthreads();
        if (n == 0) break;
        const int rows = kEigThreads / n;          // rows per pass (n < kEigThreads)
        const int r0 = tid / n, k = tid % n;
        if (r0 < rows)
            for (int i = r0; i < n; i += rows)
                if (k <= i) {
                    cd* a = &A[(size_t)(j + 1 + i) * d + (j + 1 + k)];
                    *a = csub(*a, cmulc(col[i], col[k]));
                }
    }
    __syncthreads();
    return true;
}

// C = L^-1 H L^-H from S and H: Cholesky into L, T = L^-1 H (one thread per column),
// W = L^-1 T^H, C = (W^H + W) / 2 made exactly Hermitian.
__device__ bool reduce_to_standard(const cd* S, const cd* H, cd* L, cd* C, cd* W, cd* T, int d, cd* col, double* sh)
{
    const int tid = threadIdx.x;
    for (long x = tid; x < (long)d * d; x += kEigThreads) L[x] = S[x];
    if (!chol_lower(L, d, col, sh)) return false;
    for (int c = tid; c < d; c += kEigThreads)
        for (int i = 0; i < d; ++i) {
            cd t = H[(size_t)i * d + c];
            for (int k = 0; k < i; ++k) t = csub(t, cmul(L[(size_t)i * d + k], T[(size_t)k * d + c]));
            T[(size_t)i * d + c] = cscale(t, 1.0 / L[(size_t)i * d + i].r);
        }
    __syncthreads();
    for (int r = tid; r < d; r += kEigThreads)
        for (int i = 0; i < d; ++i) {
            cd t = cconj(T[(size_t)r * d + i]);
            for (int k = 0; k < i; ++k) t = csub(t, cmul(L[(size_t)i * d + k], W[(size_t)k * d + r]));
            W[(size_t)i * d + r] = cscale(t, 1.0 / L[(size_t)i * d + i].r);
        }
    __syncthreads();
    for (long x = tid; x < (long)d * d; x += kEigThreads) {
        const int i = (int)(x / d), j = (int)(x % d);
        cd c = cscale(cadd(cconj(W[(size_t)j * d + i]), W[(size_t)i * d + j]), 0.5);
        if (i == j) c.i = 0.0;
        C[x] = c;
    }
    __syncthreads();
    return true;
}

// Eigenvalues only: reduce to C = L^-1 H L^-H, then Householder tridiagonalization of the
// Hermitian C (H_k = I - tau v v^H zeroes column k below the subdiagonal; the trailing block
// takes B <- B - v w^H - w v^H with p = tau B v, w = p - (tau/2)(v^H p) v, its columns read
// coalesced through B[i][j] = conj(B[j][i])), then the real symmetric tridiagonal with
// diagonal Re C_ii and off-diagonal |beta_k| (a diagonal unitary similarity removes the
// phases) is solved by Sturm-count bisection, one thread per eigenvalue, ascending.
// Per-problem scratch beyond L, C, W, T (4 d^2 complex): E, 8 d^2 + 8 d doubles holding the
// real tridiagonal (a, b), the reflector factors tau, the off-diagonal phases, the real
// tridiagonal eigenvectors X [d][d] and the pivoted LU factors of T - lambda I (5 x [d][d]).
__device__ __forceinline__ double* eig_extra(double* workd, const GenEigProblem& pr)
{
    return workd + 2 * (pr.work_off + 4L * pr.d * pr.d);
}

__global__ __launch_bounds__(kEigThreads) void k_geneig_vals(const double* Sd, const double* Hd, double* workd,
                                                            const GenEigProblem* probs, double* evals, int* status,
                                                            int keep)
{
    __shared__ double red[kEigThreads / 64];
    __shared__ cd vv[2 * kEigMaxPairs], ww[2 * kEigMaxPairs];
    __shared__ double da[2 * kEigMaxPairs], db[2 * kEigMaxPairs];
    __shared__ double sh;

    const GenEigProblem pr = probs[blockIdx.x];
    const int d = pr.d;
    const int tid = threadIdx.x;
    const cd* S = reinterpret_cast<const cd*>(Sd) + pr.in_off;
    const cd* H = reinterpret_cast<const cd*>(Hd) + pr.in_off;
    cd* L = reinterpret_cast<cd*>(workd) + pr.work_off;
    cd* C = L + (size_t)d * d;
    cd* W = C + (size_t)d * d;
    cd* T = W + (size_t)d * d;
    double* ev = evals + pr.eval_off;

    if (!reduce_to_standard(S, H, L, C, W, T, d, vv, &sh)) {
        if (tid == 0) status[blockIdx.x] = 1;
        return;
    }
    for (int k = 0; k + 2 < d; ++k) {
        const int n = d - k - 1;
        double part = 0.0;
        for (int i = tid; i < n; i += kEigThreads) {
            const cd x = C[(size_t)(k + 1 + i) * d + k];
            vv[i] = x;
            part += x.r * x.r + x.i * x.i;
        }
        const double alpha2 = block_sum(part, red);   // synchronises
        const double alpha = sqrt(alpha2);
        const cd x0 = vv[0];
        const double ax0 = sqrt(x0.r * x0.r + x0.i * x0.i);
        const cd ph = ax0 > 0.0 ? cd{x0.r / ax0, x0.i / ax0} : cd{1.0, 0.0};
        if (tid == 0) db[k] = alpha;                 // |beta_k|, beta_k = -ph alpha
        __syncthreads();                              // every thread has read vv[0]
        if (!(alpha > 0.0)) {                         // column already reduced (uniform)
            if (keep && tid == 0) {
                double* E = eig_extra(workd, pr);
                E[2 * d + k] = 0.0;
                E[3 * d + 2 * k] = 1.0;
                E[3 * d + 2 * k + 1] = 0.0;
            }
            continue;
        }
        const double tau = 1.0 / (alpha * (alpha + ax0));
        if (tid == 0) vv[0] = cadd(x0, cscale(ph, alpha));
        __syncthreads();
        if (keep) {   // H_k = I - tau v v^H: v into column k below the diagonal, beta_k = -ph alpha
            for (int i = tid; i < n; i += kEigThreads) C[(size_t)(k + 1 + i) * d + k] = vv[i];
            if (tid == 0) {
                double* E = eig_extra(workd, pr);
                E[2 * d + k] = tau;
                E[3 * d + 2 * k] = -ph.r;
                E[3 * d + 2 * k + 1] = -ph.i;
            }
        }
        // p = tau B v (thread i: column i of B, coalesced across threads)
        for (int i = tid; i < n; i += kEigThreads) {
            cd acc = {0.0, 0.0};
            const cd* colp = &C[(size_t)(k + 1) * d + (k + 1 + i)];
            for (int j = 0; j < n; ++j) {
                const cd b = colp[(size_t)j * d];
                const cd v = vv[j];
                acc.r += b.r * v.r + b.i * v.i;      // conj(b) v
                acc.i += b.r * v.i - b.i * v.r;
            }
            ww[i] = cscale(acc, tau);
        }
        __syncthreads();
        double vp = 0.0;
        for (int i = tid; i < n; i += kEigThreads) vp += vv[i].r * ww[i].r + vv[i].i * ww[i].i;   // Re(v^H p)
        const double K = 0.5 * tau * block_sum(vp, red);
        for (int i = tid; i < n; i += kEigThreads) ww[i] = csub(ww[i], cscale(vv[i], K));
        __syncthreads();
        const int rows = kEigThreads / n;
        const int r0 = tid / n, j = tid % n;
        if (r0 < rows)
            for (int i = r0; i < n; i += rows) {
                cd* a = &C[(size_t)(k + 1 + i) * d + (k + 1 + j)];
                const cd u = cadd(cmulc(vv[i], ww[j]), cmulc(ww[i], vv[j]));
                *a = csub(*a, u);
            }
        __syncthreads();
    }
    if (d >= 2) {
        __syncthreads();
        if (tid == 0) {
            const cd e = C[(size_t)(d - 1) * d + (d - 2)];
            const double ae = sqrt(e.r * e.r + e.i * e.i);
            db[d - 2] = ae;
            if (keep) {
                double* E = eig_extra(workd, pr);
                E[2 * d + (d - 2)] = 0.0;
                E[3 * d + 2 * (d - 2)] = ae > 0.0 ? e.r / ae : 1.0;
                E[3 * d + 2 * (d - 2) + 1] = ae > 0.0 ? e.i / ae : 0.0;
            }
        }
    }
    for (int i = tid; i < d; i += kEigThreads) da[i] = C[(size_t)i * d + i].r;
    __syncthreads();
    if (keep) {
        double* E = eig_extra(workd, pr);
        for (int i = tid; i < d; i += kEigThreads) {
            E[i] = da[i];
            E[d + i] = db[i];
        }
    }
    // Gershgorin interval and pivot floor
    double lo = 1e300, hi = -1e300, bmax = 0.0;
    for (int i = 0; i < d; ++i) {
        const double r = (i > 0 ? db[i - 1] : 0.0) + (i + 1 < d ? db[i] : 0.0);
        lo = fmin(lo, da[i] - r);
        hi = fmax(hi, da[i] + r);
        if (i + 1 < d) bmax = fmax(bmax, db[i]);
    }
    const double span = fmax(fabs(lo), fabs(hi));
    const double pivmin = 1e-290 * fmax(1.0, bmax * bmax);
    for (int m = tid; m < d; m += kEigThreads) {
        double l = lo - 1e-12 * span - 1e-300, h = hi + 1e-12 * span + 1e-300;
        for (int it = 0; it < 200 && h - l > 2.2e-16 * fmax(fabs(l), fabs(h)) + 1e-300; ++it) {
            const double x = 0.5 * (l + h);
            int cnt = 0;
            double q = da[0] - x;
            if (fabs(q) < pivmin) q = -pivmin;
            cnt += q < 0.0;
            for (int i = 1; i < d; ++i) {
                q = da[i] - x - db[i - 1] * db[i - 1] / q;
                if (fabs(q) < pivmin) q = -pivmin;
                cnt += q < 0.0;
            }
            if (cnt > m) h = x;
            else l = x;
        }
        ev[m] = 0.5 * (l + h);
    }
    if (tid == 0) status[blockIdx.x] = 0;
}

// Eigenvectors of the real tridiagonal (a, b) by inverse iteration (one thread per cluster of
// eigenvalues closer than 1e-7 |T|; members after the first are shifted apart by 10 eps |T| and
// made orthogonal to the earlier members by modified Gram-Schmidt in every iteration, as in
// LAPACK's stein).  (T - lambda I) is factored with partial pivoting (LAPACK gttrf layout:
// multipliers dl, diagonal dd, super-diagonals du, du2, pivot flags); zero pivots become
// eps |T|.  Each vector gets 3 solves from a fixed pseudo-random start and is normalised.
__global__ __launch_bounds__(kEigThreads) void k_eig_invit(double* workd, const GenEigProblem* probs,
                                                           const double* evals, const int* status)
{
    __shared__ double a[2 * kEigMaxPairs], b[2 * kEigMaxPairs];
    __shared__ int cstart[2 * kEigMaxPairs + 1];
    __shared__ int ncl;
    __shared__ double tnorm;
    const GenEigProblem pr = probs[blockIdx.x];
    if (status[blockIdx.x] != 0) return;
    const int d = pr.d;
    const int tid = threadIdx.x;
    double* E = eig_extra(workd, pr);
    const double* ev = evals + pr.eval_off;
    for (int i = tid; i < d; i += kEigThreads) {
        a[i] = E[i];
        b[i] = E[d + i];
    }
    __syncthreads();
    if (tid == 0) {
        double an = 0.0, bn = 0.0;
        for (int i = 0; i < d; ++i) {
            an = fmax(an, fabs(a[i]));
            if (i + 1 < d) bn = fmax(bn, b[i]);
        }
        tnorm = fmax(an + 2.0 * bn, 1e-300);
        int c = 0;
        cstart[0] = 0;
        for (int j = 1; j < d; ++j)
            if (ev[j] - ev[j - 1] > 1e-7 * tnorm) cstart[++c] = j;
        cstart[c + 1] = d;
        ncl = c + 1;
    }
    __syncthreads();
    const double eps = 2.220446049250313e-16;
    const double tiny = eps * tnorm;
    double* X = E + 5 * d;
    double* dl = X + (size_t)d * d;
    double* dd = dl + (size_t)d * d;
    double* du = dd + (size_t)d * d;
    double* du2 = du + (size_t)d * d;
    double* pv = du2 + (size_t)d * d;
    auto at = [d](double* base, int i, int j) -> double& { return base[(size_t)i * d + j]; };
    for (int cl = tid; cl < ncl; cl += kEigThreads) {
        double lam_prev = 0.0;
        for (int j = cstart[cl]; j < cstart[cl + 1]; ++j) {
            double lam = ev[j];
            if (j > cstart[cl]) lam = fmax(lam, lam_prev + 10.0 * tiny);
            lam_prev = lam;
            // factor T - lam I
            for (int i = 0; i < d; ++i) {
                at(dd, i, j) = a[i] - lam;
                if (i + 1 < d) {
                    at(dl, i, j) = b[i];
                    at(du, i, j) = b[i];
                }
            }
            for (int i = 0; i + 1 < d; ++i) {
                const double di = at(dd, i, j), li = at(dl, i, j);
                if (fabs(di) >= fabs(li)) {
                    const double piv = di != 0.0 ? di : tiny;
                    at(dd, i, j) = piv;
                    const double l = li / piv;
                    at(dl, i, j) = l;
                    at(dd, i + 1, j) -= l * at(du, i, j);
                    if (i + 2 < d) at(du2, i, j) = 0.0;
                    at(pv, i, j) = 0.0;
                } else {
                    const double l = di / li;
                    at(dd, i, j) = li;
                    at(dl, i, j) = l;
                    const double t = at(du, i, j);
                    at(du, i, j) = at(dd, i + 1, j);
                    at(dd, i + 1, j) = t - l * at(dd, i + 1, j);
                    if (i + 2 < d) {
                        at(du2, i, j) = at(du, i + 1, j);
                        at(du, i + 1, j) = -l * at(du, i + 1, j);
                    }
                    at(pv, i, j) = 1.0;
                }
            }
            if (at(dd, d - 1, j) == 0.0) at(dd, d - 1, j) = tiny;
            // start vector
            for (int i = 0; i < d; ++i) {
                uint32_t h = (uint32_t)(i * 2654435761u) ^ (uint32_t)((j + 1) * 40503u);
                h ^= h >> 15;
                h *= 2246822519u;
                h ^= h >> 13;
                at(X, i, j) = 0.5 + (double)(h & 0xffffu) / 65536.0;
            }
            for (int it = 0; it < 3; ++it) {
                // solve L U x = x (pivoted forward, then backward)
                for (int i = 0; i + 1 < d; ++i) {
                    if (at(pv, i, j) == 0.0) {
                        at(X, i + 1, j) -= at(dl, i, j) * at(X, i, j);
                    } else {
                        const double t = at(X, i, j);
                        at(X, i, j) = at(X, i + 1, j);
                        at(X, i + 1, j) = t - at(dl, i, j) * at(X, i, j);
                    }
                }
                at(X, d - 1, j) /= at(dd, d - 1, j);
                if (d >= 2) at(X, d - 2, j) = (at(X, d - 2, j) - at(du, d - 2, j) * at(X, d - 1, j)) / at(dd, d - 2, j);
                for (int i = d - 3; i >= 0; --i)
                    at(X, i, j) = (at(X, i, j) - at(du, i, j) * at(X, i + 1, j) - at(du2, i, j) * at(X, i + 2, j)) /
                                  at(dd, i, j);
                // orthogonalise against the earlier members of the cluster, normalise
                for (int k = cstart[cl]; k < j; ++k) {
                    double dot = 0.0;
                    for (int i = 0; i < d; ++i) dot += at(X, i, k) * at(X, i, j);
                    for (int i = 0; i < d; ++i) at(X, i, j) -= dot * at(X, i, k);
                }
                double nrm = 0.0;
                for (int i = 0; i < d; ++i) nrm += at(X, i, j) * at(X, i, j);
                nrm = nrm > 0.0 ? 1.0 / sqrt(nrm) : 0.0;
                for (int i = 0; i < d; ++i) at(X, i, j) *= nrm;
            }
        }
    }
}

// Back to generalized eigenvectors, one thread per eigenvector j of problem blockIdx.y:
// y = D x (D restores the off-diagonal phases), w = H_0 H_1 ... H_{d-3} y (reflectors applied
// last to first), v = L^-H w, so that v^H S v = 1.
__global__ __launch_bounds__(256) void k_eig_back(double* workd, const GenEigProblem* probs, double* evecsd,
                                                  const int* status)
{
    const GenEigProblem pr = probs[blockIdx.y];
    if (status[blockIdx.y] != 0) return;
    const int d = pr.d;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d) return;
    cd* L = reinterpret_cast<cd*>(workd) + pr.work_off;
    cd* C = L + (size_t)d * d;
    cd* Y = C + 2 * (size_t)d * d;    // the T area
    const double* E = eig_extra(workd, pr);
    const double* tau = E + 2 * d;
    const double* ph = E + 3 * d;
    const double* X = E + 5 * d;
    cd* V = reinterpret_cast<cd*>(evecsd) + pr.in_off;
    cd dk = {1.0, 0.0};
    for (int i = 0; i < d; ++i) {
        const double x = X[(size_t)i * d + j];
        Y[(size_t)i * d + j] = cscale(dk, x);
        if (i + 1 < d) dk = cmul(dk, cd{ph[2 * i], ph[2 * i + 1]});
    }
    for (int k = d - 3; k >= 0; --k) {
        const double t = tau[k];
        if (t == 0.0) continue;
        const int n = d - k - 1;
        cd s = {0.0, 0.0};
        for (int i = 0; i < n; ++i) {
            const cd v = C[(size_t)(k + 1 + i) * d + k];
            const cd y = Y[(size_t)(k + 1 + i) * d + j];
            s.r += v.r * y.r + v.i * y.i;     // conj(v) y
            s.i += v.r * y.i - v.i * y.r;
        }
        s = cscale(s, t);
        for (int i = 0; i < n; ++i) {
            const cd v = C[(size_t)(k + 1 + i) * d + k];
            cd* y = &Y[(size_t)(k + 1 + i) * d + j];
            *y = csub(*y, cmul(v, s));
        }
    }
    for (int i = d - 1; i >= 0; --i) {
        cd t = Y[(size_t)i * d + j];
        for (int k = i + 1; k < d; ++k) t = csub(t, cmul(cconj(L[(size_t)k * d + i]), V[(size_t)k * d + j]));
        V[(size_t)i * d + j] = cscale(t, 1.0 / L[(size_t)i * d + i].r);
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
size_t sweep_lds_bytes(int prec, int G, int N)
{
    const int S = (N + G - 1) / G;
    const size_t el = (prec == kFP64) ? sizeof(double2) : sizeof(float2);
    const size_t rnd = (prec == kFP64) ? 24 : 16;   // per-lane random numbers of a move batch
    return (size_t)(64 / G) * (size_t)S * G * el + 64 * rnd;
}

bool sweep_supported(int G) { return G == 2 || G == 4 || G == 8 || G == 16 || G == 32 || G == 64; }

hipError_t launch_init_positions(void* pos, int prec, int N, int walkers, double R,
                                 unsigned long long seed, hipStream_t s, int walker_id0)
{
    long n = (long)N * walkers;
    dim3 grid((unsigned)((n + 255) / 256));
    if (prec == kFP64)
        hipLaunchKernelGGL(k_init_positions<double>, grid, dim3(256), 0, s, (double2*)pos, N, walkers, R, seed,
                           walker_id0);
    else
        hipLaunchKernelGGL(k_init_positions<float>, grid, dim3(256), 0, s, (float2*)pos, N, walkers, R, seed,
                           walker_id0);
    return hipGetLastError();
}

template <int G, typename Real, bool ROW = false, bool MIX = false>
static hipError_t sweep_g(const SweepArgs& a, hipStream_t s)
{
    const size_t lds = sweep_lds_bytes(sizeof(Real) == 8 ? kFP64 : kFP32, G, a.N);
    constexpr int WPB = 64 / G;
    dim3 grid((unsigned)((a.walkers + WPB - 1) / WPB));
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)k_sweep<G, Real, ROW, MIX>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_sweep<G, Real, ROW, MIX>), grid, dim3(64), lds, s, a);
    return hipGetLastError();
}

template <typename Real, bool ROW = false, bool MIX = false>
static hipError_t sweep_dispatch(const SweepArgs& a, int G, hipStream_t s)
{
    switch (G) {
    case 2: return sweep_g<2, Real, ROW, MIX>(a, s);
    case 4: return sweep_g<4, Real, ROW, MIX>(a, s);
    case 8: return sweep_g<8, Real, ROW, MIX>(a, s);
    case 16: return sweep_g<16, Real, ROW, MIX>(a, s);
    case 32: return sweep_g<32, Real, ROW, MIX>(a, s);
    case 64: return sweep_g<64, Real, ROW, MIX>(a, s);
    default: return hipErrorInvalidValue;
    }
}

// Register-resident sweeps: instantiated for G in {8, 16} and S in {13, 25} slots per lane.
// S = ceil(N / G) exactly (N = 97..104, 193..200 at G = 8; N = 193..208, 385..400 at G = 16)
// runs the exact kernels, for every estimator.  Other N with 2 <= ceil(N / G) <= 25 run the
// S = 13 or S = 25 kernel with the unused slots skipped (plain and row-free chains only; the
// mixture chains keep the exact fits).
template <int G, int S, typename Real>
static hipError_t sweep_reg_gs(const SweepArgs& a, hipStream_t s)
{
    constexpr int WPB = 64 / G;
    dim3 grid((unsigned)((a.walkers + WPB - 1) / WPB));
    if ((a.N + G - 1) / G != S) {
        if (a.mix.kmix > 0) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_sweep_reg<G, S, Real, 0, false>), grid, dim3(64), 0, s, a);
        return hipGetLastError();
    }
    // mixture: hold 16 terms when they suffice (N = 200 at 8 lanes: 124 VGPRs and 4 waves per
    // SIMD, against 131 and 3 with room for 24)
    if (a.mix.kmix > 16)
        hipLaunchKernelGGL((k_sweep_reg<G, S, Real, kMixMax>), grid, dim3(64), 0, s, a);
    else if (a.mix.kmix > 0)
        hipLaunchKernelGGL((k_sweep_reg<G, S, Real, 16>), grid, dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL((k_sweep_reg<G, S, Real>), grid, dim3(64), 0, s, a);
    return hipGetLastError();
}

bool sweep_reg_supported(int G, int N, bool mixture)
{
    const int S = (N + G - 1) / G;
    if (G != 8 && G != 16) return false;
    if (S == 13 || S == 25) return true;
    return !mixture && S >= 2 && S <= 25;
}

template <typename Real>
static hipError_t sweep_reg_dispatch(const SweepArgs& a, int G, hipStream_t s)
{
    const int S = (a.N + G - 1) / G;
    if (G == 8) return S <= 13 ? sweep_reg_gs<8, 13, Real>(a, s) : sweep_reg_gs<8, 25, Real>(a, s);
    if (G == 16) return S <= 13 ? sweep_reg_gs<16, 13, Real>(a, s) : sweep_reg_gs<16, 25, Real>(a, s);
    return hipErrorInvalidValue;
}

hipError_t launch_sweep_reg(const SweepArgs& a, int prec, int G, hipStream_t s)
{
    if (a.nsweeps <= 0 || a.walkers <= 0) return hipSuccess;
    if (!sweep_reg_supported(G, a.N, a.mix.kmix > 0)) return hipErrorInvalidValue;
    if (a.mix.kmix > 0 && (a.mix.kmix > kMixMax || !a.mix_pk)) return hipErrorInvalidValue;
    return prec == kFP64 ? sweep_reg_dispatch<double>(a, G, s) : sweep_reg_dispatch<float>(a, G, s);
}

hipError_t launch_sweep(const SweepArgs& a, int prec, int G, hipStream_t s)
{
    if (a.nsweeps <= 0 || a.walkers <= 0) return hipSuccess;
    if (a.mix.kmix > 0) {
        if (a.mix.kmix > kMixMax || !a.mix_pk) return hipErrorInvalidValue;
        return prec == kFP64 ? sweep_dispatch<double, false, true>(a, G, s) : sweep_dispatch<float, false, true>(a, G, s);
    }
    return prec == kFP64 ? sweep_dispatch<double>(a, G, s) : sweep_dispatch<float>(a, G, s);
}

// p_k, k = 1..kmix, of every walker's positions for the mixture chains: 4 lanes per walker
// (particles j = sub + 4 s), FP64 powers by repeated multiplication, a 2-step butterfly
template <typename Real>
__global__ __launch_bounds__(256) void k_mix_init(SweepArgs a)
{
    using V2 = typename V2T<Real>::T;
    const long gt = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long w = gt / kLps;
    const int sub = (int)(gt % kLps);
    if (w >= a.walkers) return;                  // the 4 lanes of a walker leave together
    const int kmix = a.mix.kmix;
    double pr[kMixMax], pi[kMixMax];
#pragma unroll
    for (int k = 0; k < kMixMax; ++k) pr[k] = pi[k] = 0.0;
    const V2* P = reinterpret_cast<const V2*>(a.pos) + (size_t)w * a.N;
    for (int j = sub; j < a.N; j += kLps) {
        const V2 q = P[j];
        const double ur = (double)q.x / a.R, ui = (double)q.y / a.R;
        double qr = ur, qi = ui;
#pragma unroll
        for (int k = 0; k < kMixMax; ++k) {
            if (k < kmix) {
                pr[k] += qr;
                pi[k] += qi;
                const double t = qr * ur - qi * ui;
                qi = qr * ui + qi * ur;
                qr = t;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kMixMax; ++k) {
        pr[k] = group_sum<kLps, double>(pr[k]);
        pi[k] = group_sum<kLps, double>(pi[k]);
        if (k % kLps == sub && k < kmix) a.mix_pk[(size_t)w * kMixMax + k] = make_double2(pr[k], pi[k]);
    }
}

hipError_t launch_mix_init(const SweepArgs& a, int prec, hipStream_t s)
{
    if (a.walkers <= 0) return hipSuccess;
    if (a.mix.kmix < 1 || a.mix.kmix > kMixMax || !a.mix_pk) return hipErrorInvalidValue;
    const long threads = (long)a.walkers * kLps;
    dim3 grid((unsigned)((threads + 255) / 256));
    if (prec == kFP64)
        hipLaunchKernelGGL(k_mix_init<double>, grid, dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(k_mix_init<float>, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Row-wise estimator (the reference's: em_spectrum.cu:16-137, em_markovchain.cu:185-220).
// ---------------------------------------------------------------------------
hipError_t launch_sweep_row(const SweepArgs& a, int prec, int G, hipStream_t s)
{
    if (a.nsweeps <= 0 || a.walkers <= 0) return hipSuccess;
    if (!a.row_fac || !a.row_pk || a.row_walkers < 1) return hipErrorInvalidValue;
    return prec == kFP64 ? sweep_dispatch<double, true>(a, G, s) : sweep_dispatch<float, true>(a, G, s);
}

// p_k (k = 1..kmax of the chain's row state) of the chain's positions, summed over the
// particles in order with u^k by repeated multiplication (sampling_ek_construct,
// em_markovchain.cu:185-199); one thread per chain
template <typename Real>
__global__ void k_row_init(SweepArgs a)
{
#pragma clang fp contract(off)
    using V2 = typename V2T<Real>::T;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.walkers) return;
    const int f = c / a.row_walkers;
    int k[kMaxFactors], kmax = 0;
#pragma unroll
    for (int q = 0; q < kMaxFactors; ++q) {
        k[q] = (int)(a.row_fac[(size_t)f * kMaxFactors + q] & 0xffu);
        kmax = max(kmax, k[q]);
    }
    double2 pk[kMaxFactors];
#pragma unroll
    for (int q = 0; q < kMaxFactors; ++q) pk[q] = make_double2(0.0, 0.0);
    const V2* P = reinterpret_cast<const V2*>(a.pos) + (size_t)c * a.N;
    for (int j = 0; j < a.N; ++j) {
        const double2 u = make_double2((double)P[j].x / a.R, (double)P[j].y / a.R);
        double2 q = u;
        for (int kk = 1; kk <= kmax; ++kk) {
#pragma unroll
            for (int t = 0; t < kMaxFactors; ++t)
                if (k[t] == kk) {
                    pk[t].x = pk[t].x + q.x;
                    pk[t].y = pk[t].y + q.y;
                }
            q = cmul_exact(q, u);
        }
    }
#pragma unroll
    for (int q = 0; q < kMaxFactors; ++q) a.row_pk[(size_t)c * kMaxFactors + q] = pk[q];
}

hipError_t launch_row_init(const SweepArgs& a, int prec, hipStream_t s)
{
    if (a.walkers <= 0) return hipSuccess;
    if (!a.row_fac || !a.row_pk || a.row_walkers < 1) return hipErrorInvalidValue;
    dim3 grid((unsigned)((a.walkers + 127) / 128));
    if (prec == kFP64) hipLaunchKernelGGL(k_row_init<double>, grid, dim3(128), 0, s, a);
    else hipLaunchKernelGGL(k_row_init<float>, grid, dim3(128), 0, s, a);
    return hipGetLastError();
}

// One workgroup per row state f: thread g' sums, over the row's chains, g/f (1 at g = f),
// V g/f and |g/f|^2 for g = o + g' of f's sector, and adds them to the FP64 accumulators (one
// workgroup owns each row, so no atomics).  Complex division in FP64.
template <typename Real>
__global__ __launch_bounds__(256) void k_row_accum(RowAccumArgs a)
{
    const int f = blockIdx.x;
    if (f >= a.D) return;
    const int s = a.row_sector[f];
    const int o = a.sec_off[s], d = a.sec_off[s + 1] - o;
    const int fa = f - o;
    const Real* X = reinterpret_cast<const Real*>(a.X);
    const Real* V = reinterpret_cast<const Real*>(a.V);
    const int cf = a.col_re[f], df = a.col_dim[f];
    const long c0 = (long)f * a.row_walkers;
    for (int gp = threadIdx.x; gp < d; gp += blockDim.x) {
        const int g = o + gp;
        const int cg = a.col_re[g], dg = a.col_dim[g];
        double sr = 0, si = 0, hr = 0, hi = 0, gg = 0;
        for (int w = 0; w < a.row_walkers; ++w) {
            const Real* x = X + (size_t)(c0 + w) * a.ldx;
            double rr = 1.0, ri = 0.0;
            if (g != f) {
                const double fr = (double)x[cf], fi = (double)x[cf + df];
                const double gr = (double)x[cg], gi = (double)x[cg + dg];
                const double den = fr * fr + fi * fi;
                rr = (gr * fr + gi * fi) / den;
                ri = (gi * fr - gr * fi) / den;
            }
            const double v = (double)V[c0 + w];
            sr += rr;
            si += ri;
            hr += v * rr;
            hi += v * ri;
            gg += rr * rr + ri * ri;
        }
        const size_t ix = (size_t)a.sec_bo[s] + (size_t)fa * d + gp;
        a.acc[ix] += sr;
        a.acc[a.block_elems + ix] += si;
        a.acc[2 * a.block_elems + ix] += hr;
        a.acc[3 * a.block_elems + ix] += hi;
        a.acc[4 * a.block_elems + ix] += gg;
    }
}

hipError_t launch_row_accum(const RowAccumArgs& a, int prec, hipStream_t s)
{
    if (a.D <= 0) return hipSuccess;
    if (prec == kFP64) hipLaunchKernelGGL(k_row_accum<double>, dim3((unsigned)a.D), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(k_row_accum<float>, dim3((unsigned)a.D), dim3(256), 0, s, a);
    return hipGetLastError();
}

template <typename Real, int NH>
static hipError_t features_h(const FeatureArgs& a, hipStream_t s)
{
    const long threads = (long)a.walkers * kLps;
    dim3 grid_ps((unsigned)((threads + 255) / 256));
    const bool mix = NH == 0 && a.mix.kmix > 0;
    if (a.K <= 8) {
        if (mix) hipLaunchKernelGGL((k_psums<Real, 8, 0, true>), grid_ps, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_psums<Real, 8, NH>), grid_ps, dim3(256), 0, s, a);
    } else if (a.K <= 16) {
        if (mix) hipLaunchKernelGGL((k_psums<Real, 16, 0, true>), grid_ps, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_psums<Real, 16, NH>), grid_ps, dim3(256), 0, s, a);
    } else if (a.K <= 24) {
        if (mix) hipLaunchKernelGGL((k_psums<Real, 24, 0, true>), grid_ps, dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_psums<Real, 24, NH>), grid_ps, dim3(256), 0, s, a);
    } else {
        return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // enough waves to fill the chip, each looping over samples
    const long blocks = std::min<long>((a.walkers + 3) / 4, 2048);
    const size_t lds = (size_t)4 * 32 * sizeof(double2) + (size_t)4 * kMaxPow * 2 * sizeof(Real) +
                       (a.D <= kMaxStatesLds ? (size_t)a.D * (sizeof(unsigned long long) + sizeof(uint32_t)) : 0);
    hipLaunchKernelGGL((k_monomials<Real>), dim3((unsigned)blocks), dim3(256), lds, s, a);
    return hipGetLastError();
}

template <typename Real>
static hipError_t features_t(const FeatureArgs& a, hipStream_t s)
{
    if (a.npw > kMaxPow || !a.ps || a.ps_ld < 2 * a.K) return hipErrorInvalidValue;
    // mixture weights: S and H only (no harmonic sums), p_k up to the largest term
    if (a.mix.kmix > 0 && (a.mix.kmix > a.K || a.mix.kmix > kMixMax || !a.mix_acc || a.nh > 0 || a.mix_row < 1))
        return hipErrorInvalidValue;
    // harmonic accumulators are FP64 registers: hold only as many pairs as the call has
    if (a.nh > kMaxHarmonics) return hipErrorInvalidValue;
    if (a.nh == 0) return features_h<Real, 0>(a, s);
    if (a.nh <= 2) return features_h<Real, 2>(a, s);
    return features_h<Real, kMaxHarmonics>(a, s);
}

hipError_t launch_features(const FeatureArgs& a, int prec, hipStream_t s)
{
    if (a.walkers <= 0) return hipSuccess;
    return prec == kFP64 ? features_t<double>(a, s) : features_t<float>(a, s);
}

hipError_t launch_gram(const GramArgs& a, int prec, int splits, hipStream_t s)
{
    if (a.rows <= 0 || a.ntiles <= 0) return hipSuccess;
    if (a.splits != splits) return hipErrorInvalidValue;
    // 32-column tiles: FP32 and FP64 tiles only
    if (a.tile_cols != 0 && a.tile_cols != 64 && (a.tile_cols != 32 || prec == kBF16X3 || prec == kBF16X6))
        return hipErrorInvalidValue;
    if (a.jobs ? (a.njobs < 1 || a.njobs > 65535) : a.njobs > 1) return hipErrorInvalidValue;
    dim3 grid((unsigned)gram_grid(a.ntiles, splits, a.xcd_spread), (unsigned)(a.jobs ? a.njobs : 1));
    if (prec == kBF16X3 || prec == kBF16X6) {
        if (a.rows_per_split % kKC) return hipErrorInvalidValue;
        if (prec == kBF16X3) {
            if (a.nw == 2) hipLaunchKernelGGL((k_gram_bf16<2, 2>), grid, dim3(512), 0, s, a);
            else hipLaunchKernelGGL((k_gram_bf16<1, 2>), grid, dim3(512), 0, s, a);
        } else {
            if (a.nw == 2) hipLaunchKernelGGL((k_gram_bf16<2, 3>), grid, dim3(512), 0, s, a);
            else hipLaunchKernelGGL((k_gram_bf16<1, 3>), grid, dim3(512), 0, s, a);
        }
    } else if (a.tile_cols == 32) {
        if (prec == kFP64) {
            if (a.nw == 2) hipLaunchKernelGGL((k_gram<double, 2, 32>), grid, dim3(256), 0, s, a);
            else hipLaunchKernelGGL((k_gram<double, 1, 32>), grid, dim3(256), 0, s, a);
        } else {
            if (a.nw == 2) hipLaunchKernelGGL((k_gram<float, 2, 32>), grid, dim3(256), 0, s, a);
            else hipLaunchKernelGGL((k_gram<float, 1, 32>), grid, dim3(256), 0, s, a);
        }
    } else if (prec == kFP64) {
        if (a.nw == 2) hipLaunchKernelGGL((k_gram<double, 2>), grid, dim3(512), 0, s, a);
        else hipLaunchKernelGGL((k_gram<double, 1>), grid, dim3(512), 0, s, a);
    } else {
        if (a.nw == 2) hipLaunchKernelGGL((k_gram<float, 2>), grid, dim3(512), 0, s, a);
        else hipLaunchKernelGGL((k_gram<float, 1>), grid, dim3(512), 0, s, a);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Fused sampling step: one launch per kept sample holds the sweeps, power sums and monomials of
// every walker (into the current sample's rows) and the S/H Gram tiles of the previous
// sample's rows, as work items: sweep units (the register sweeps of 8 wavefronts' walkers,
// then their FP64 power sums and monomials) and Gram items (one 64x64 tile over one row split).
// Each workgroup (512 threads, at most 128 VGPRs so that two share a CU at 4 wavefronts per
// SIMD) takes the next item of its queue from an atomic counter, so items run in queue order
// whatever order the dispatcher starts workgroups in, and the dispatcher backfills a CU as soon
// as a workgroup ends.  A CU then holds sweep workgroups, bound by the VALU, and Gram
// workgroups, bound by the matrix cores, side by side; on gfx950 the two overlap (a probe of
// FMA-only and MFMA-only workgroups in one launch ran in the time of the FMA part alone).
// Within a queue, sweep and Gram items alternate while both remain, so that the long sweep
// items start early and short Gram items fill the end.  (A persistent loop over items made
// the compiler keep values live across the loop: 166 VGPRs spilled at the 128 cap.)
// The chains, power sums and monomials are those of the separate kernels (same device code,
// same random numbers; the power sums in the quads of k_psums); the Gram sums differ from a
// separate Gram launch only in where the row splits fall.
// ---------------------------------------------------------------------------
constexpr int kStepPlanMax = 992;    // product plans held in LDS (larger D reads them from memory; 64 KB of LDS in all)

// Power sums of the fused kernel: 8 lanes per sample, lanes 0..3 of the 8 holding Re/Im p_k for
// k < KMID and V, lanes 4..7 the k >= KMID, each quad over the particles j = sub + 4 s as in
// k_psums (same powers, same sums, same quad reduction: the same values), with half the
// accumulators of a lane (k_psums holds all of them: 152 VGPRs at K = 16).
template <int KMAX>
__device__ __forceinline__ void step_psums(const FeatureArgs& a, long w, int sub8)
{
    constexpr int KMID = KMAX / 2;
    constexpr int NQ = KMAX + 1;                 // the larger half: Re/Im of KMID powers, and V
    const bool hi = sub8 >= kLps;
    const int sub = sub8 & (kLps - 1);
    const int k0 = hi ? KMID : 0;
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
    for (int j = sub; j < a.N; j += kLps) {
        const float2 p = reinterpret_cast<const float2*>(a.pos)[(size_t)w * a.N + j];
        const double x = p.x, y = p.y;
        const double r2 = x * x + y * y;
        if (!hi) acc[KMAX] += potential_r2(a.pot, r2);
        const double ur = x * a.invR, ui = y * a.invR;
        double qr = ur, qi = ui;
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            if (k < KMID) {
                if (!hi) {
                    acc[2 * k] += qr;
                    acc[2 * k + 1] += qi;
                }
            } else if (hi) {
                acc[2 * (k - KMID)] += qr;
                acc[2 * (k - KMID) + 1] += qi;
            }
            const double t = qr * ur - qi * ui;
            qi = qr * ui + qi * ur;
            qr = t;
        }
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = group_sum<kLps, double>(acc[q]);
    double* ps = a.ps + (size_t)w * a.ps_ld;
#pragma unroll
    for (int q = 0; q < KMAX; ++q) {
        const int qq = q + 2 * k0;
        if (q % kLps == sub && qq < 2 * a.K) ps[qq] = acc[q];
    }
    if (hi) return;
    const long row = a.row0 + w;
    if (sub == 0) reinterpret_cast<float*>(a.W)[row] = (float)acc[KMAX];
    if (a.trace && w < a.trace_walkers && sub == 0) {
        const long long tb = a.trace_base ? *a.trace_base : 0ll;
        float* tr = a.trace + ((size_t)(a.trace_row0 + tb) + (size_t)w) * 9;
#pragma unroll
        for (int q = 0; q < 8; ++q) tr[q] = (q / 2 < a.K) ? (float)acc[q] : 0.f;
        tr[8] = (float)acc[KMAX];
    }
}

// Items of queue q: sweep units u = q + 8 j (u < units) and Gram items (tile, split) with
// split = q + 8 j' (split < splits); in queue order, item i < 2 m (m = the smaller count) is a
// sweep item for even i, a Gram item for odd i, and the rest are of the larger kind.
struct StepQueue {
    int ns, ng, m, total;
};

__device__ __host__ inline StepQueue step_queue(int q, int units, long grows, int splits, int ntiles)
{
    StepQueue r;
    r.ns = units > q ? (units - q + 7) / 8 : 0;
    r.ng = (grows > 0 && splits > q) ? ((splits - q + 7) / 8) * ntiles : 0;
    r.m = r.ns < r.ng ? r.ns : r.ng;
    r.total = r.ns + r.ng;
    return r;
}

template <int G, int S, int KMAX>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void k_step(StepArgs st)
{
    static_assert(G == 8, "sweep units are 8 wavefronts of 8-lane walkers");
    __shared__ __align__(16) double2 mpk[8][32];
    __shared__ __align__(16) float2 mtab[8][kMaxPow];
    __shared__ unsigned long long mplan[kStepPlanMax];
    __shared__ uint32_t mcols[kStepPlanMax];
    __shared__ int item_s;
    constexpr int NWU = 8 * (64 / G);            // walkers per sweep unit
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const FeatureArgs& f = st.f;
    const int q = blockIdx.x & 7;
    const int nwg_q = ((int)gridDim.x - q + 7) / 8;   // workgroups of queue q (one fetch each)
    const StepQueue Q = step_queue(q, st.units, st.g.rows, st.g.splits, st.g.ntiles);
    if (tid == 0) item_s = (int)atomicAdd(&st.queue[q], 1u);
    __syncthreads();
    const int i = __builtin_amdgcn_readfirstlane(item_s);
    if (i < Q.total) {
        bool sweep;
        int j;
        if (i < 2 * Q.m) {
            sweep = !(i & 1);
            j = i >> 1;
        } else {
            sweep = Q.ns > Q.ng;
            j = Q.m + (i - 2 * Q.m);
        }
        if (sweep) {
            const long u = q + 8L * j;
            const bool plan_lds = f.D <= kStepPlanMax;
            if (plan_lds)
                for (int t = tid; t < f.D; t += 512) {
                    mplan[t] = f.prod[t];
                    mcols[t] = (uint32_t)f.col_re[t] | ((uint32_t)f.col_dim[t] << 16);
                }
            sweep_reg_walkers<G, S, float>(st.sw, u * 8 + wave, lane);
            __syncthreads();   // the unit's positions are written
            const long w0 = u * NWU;
            {
                const long w = w0 + tid / 8;   // 8 lanes per sample, 64 samples per unit
                if (w < f.walkers) step_psums<KMAX>(f, w, tid % 8);
            }
            __syncthreads();   // their power sums are written
            for (int k = wave; k < NWU; k += 8) {
                const long w = w0 + k;
                if (w < f.walkers) monomials_sample<float>(f, w, lane, mpk[wave], mtab[wave], mplan, mcols, plan_lds);
            }
        } else {
            gram_tile<float, 2, 64>(st.g, j % st.g.ntiles, q + 8 * (j / st.g.ntiles), 0);
        }
    }
    // the last workgroup of a queue to finish clears it for the next launch
    if (tid == 0 && atomicAdd(&st.queue[8 + q], 1u) == (unsigned)(nwg_q - 1)) {
        atomicExch(&st.queue[q], 0u);
        atomicExch(&st.queue[8 + q], 0u);
    }
}

bool step_supported(int G, int N, int K)
{
    const int S = (N + G - 1) / G;
    return G == 8 && (S == 13 || S == 25) && K >= 1 && K <= 16;
}

hipError_t launch_step(const StepArgs& a, int G, hipStream_t s)
{
    if (!step_supported(G, a.sw.N, a.f.K) || !a.queue) return hipErrorInvalidValue;
    if (a.sw.nsweeps < 1 || a.sw.walkers != a.f.walkers || a.sw.mix.kmix > 0 || a.f.mix.kmix > 0 || a.f.nh > 0 ||
        a.f.npw > kMaxPow || !a.f.ps || a.f.ps_ld < 2 * a.f.K || a.f.pos != a.sw.pos || a.f.pos_is_double)
        return hipErrorInvalidValue;
    const long units = ((long)a.sw.walkers + 63) / 64;
    if (a.units != units) return hipErrorInvalidValue;
    if (a.g.rows > 0 && (a.g.nw != 2 || a.g.jobs || a.g.ntiles < 1 || a.g.rows_per_split < 1 ||
                         a.g.rows_per_split % kKC || (long)a.g.splits * a.g.rows_per_split < a.g.rows))
        return hipErrorInvalidValue;
    // one workgroup per item of the longest queue, for each of the 8 queues
    int most = 0;
    for (int q = 0; q < 8; ++q) most = std::max(most, step_queue(q, a.units, a.g.rows, a.g.splits, a.g.ntiles).total);
    if (most == 0) return hipSuccess;
    const dim3 grid((unsigned)(8 * most));
    const int S = (a.sw.N + G - 1) / G;
    const bool k8 = a.f.K <= 8;
    if (S == 25) {
        if (k8) hipLaunchKernelGGL((k_step<8, 25, 8>), grid, dim3(512), 0, s, a);
        else hipLaunchKernelGGL((k_step<8, 25, 16>), grid, dim3(512), 0, s, a);
    } else {
        if (k8) hipLaunchKernelGGL((k_step<8, 13, 8>), grid, dim3(512), 0, s, a);
        else hipLaunchKernelGGL((k_step<8, 13, 16>), grid, dim3(512), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_gram_weights(const WeightGramArgs& a, int prec, hipStream_t s)
{
    if (a.rows <= 0 || a.nweights <= 0 || a.nr <= 0 || a.nc <= 0) return hipSuccess;
    if (a.nr > 64 || a.nc > 64 || a.rows_per_split <= 0 || a.rows_per_split % kKC || a.splits < 1 ||
        (long)a.splits * a.rows_per_split < a.rows || !a.jobs)
        return hipErrorInvalidValue;
    const long tiles = (long)((a.nweights + 63) / 64) * ((a.nr * a.nc + 63) / 64);
    dim3 grid((unsigned)(tiles * a.splits));
    if (prec == kFP64) hipLaunchKernelGGL((k_gram_weights<double>), grid, dim3(512), 0, s, a);
    else if (prec == kFP32) hipLaunchKernelGGL((k_gram_weights<float>), grid, dim3(512), 0, s, a);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_extract(const ExtractArgs& a, hipStream_t s)
{
    if (a.total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_extract, dim3((unsigned)((a.total + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_geneig(const double* S, const double* H, double* work, const GenEigProblem* probs,
                         int nprob, int dmax, double* evals, double* evecs, int* status, hipStream_t s)
{
    if (nprob <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_geneig_vals, dim3((unsigned)nprob), dim3(kEigThreads), 0, s, S, H, work, probs, evals,
                       status, evecs ? 1 : 0);
    if (!evecs) return hipGetLastError();
    hipLaunchKernelGGL(k_eig_invit, dim3((unsigned)nprob), dim3(kEigThreads), 0, s, work, probs, evals, status);
    hipLaunchKernelGGL(k_eig_back, dim3((unsigned)((dmax + 255) / 256), (unsigned)nprob), dim3(256), 0, s, work, probs,
                       evecs, status);
    return hipGetLastError();
}

hipError_t launch_set_bases(unsigned long long* bases, unsigned long long sweep_base, long long trace_base,
                            hipStream_t s)
{
    hipLaunchKernelGGL(k_set_bases, dim3(1), dim3(64), 0, s, bases, sweep_base, trace_base);
    return hipGetLastError();
}

hipError_t launch_zero(double* p, long n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    long blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

}  // namespace le
