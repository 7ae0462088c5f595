/*
 * le_oracle.c -- CPU FP64 restatement of the LaughlinEdge Monte-Carlo edge-state
 * path.  TEST INFRASTRUCTURE ONLY: this library is the checker that the HIP
 * product path is compared against, and the CPU baseline that bench.py times.
 * Nothing in the product (laughlinedge_amd/) links, loads or calls it.
 *
 * Reference behaviour restated here (file:line into Nardin94/LaughlinEdge):
 *   - Metropolis single-particle sweep on |p_f Psi_L|^2 ........ modules/edge_mc/em_markovchain.cu:266-334
 *     (identical loop in modules/edge_mc/em_spectrum.cu:49-96)
 *   - power sums e_k = sum_i (s z_i)^k ........................ modules/edge_mc/em_markovchain.cu:120-154
 *   - monomials p_lambda = prod_k e_k^{r_k} .................... modules/edge_mc/em_markovchain.cu:156-183
 *   - confinement sum_i V(r_i) ................................. modules/edge_mc/em_markovchain.cu:92-99
 *   - row-wise accumulation g/f, V g/f, |g/f|^2 ................ modules/edge_mc/em_spectrum.cu:98-115
 *   - block means, /sqrt(gg) normalisation, (X+X^H)/2 .......... modules/edge_mc/em_spectrum.cu:318-335,382-390
 *   - generalized eigenproblem H v = E M v (Cholesky + Hermitian eig, ascending)
 *                                                                 modules/eigensolver/generalized_eigenproblem.cpp:45-52
 *   - DSF weights |(Vc - i Vs)[0,1:] . v_j|^2 .................. modules/eigensolver/generalized_eigenproblem.cpp:54-95
 *   - cos/sin(dL theta) excitation sums ........................ modules/edge_mc/em_markovchain.cu:101-117
 *
 * The plain |Psi_L|^2 estimator (S = <conj(p_l) p_m>, H = <conj(p_l) V p_m>) and the mixture
 * estimator (chains on w |Psi_L|^2, S and H reweighted by 1 / w) have no counterpart in the
 * reference; they are estimators the GPU path implements and are restated here from their
 * definitions.
 *
 * Random numbers: Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), the same
 * counter layout as the GPU sampler, Box-Muller normals in FP64.
 *
 * What pins each function (tests/test_oracle.py; the reference ships no tests or vectors):
 *   or_philox4x32_10 ........ Random123 known-answer vectors
 *   or_potential ............ numpy sum of V(r_i) in the fixed-accumulation test
 *   or_accumulate_fixed ..... independent numpy evaluation; exact small-N expansion
 *                             (oracle/exact_lll.py); H = w (L0 + dL + N) S at eps = 0
 *   or_accumulate_weighted .. independent numpy evaluation
 *   or_features ............. through or_accumulate_fixed (same routine)
 *   or_plain_run ............ <sum r^2> = 2 (L0 + N); eigenvalues vs exact N = 3
 *   or_plain_positions ...... the same chain code as or_plain_run
 *   or_rowwise_sector ....... exact normalised metric and spectrum at N = 3
 *   or_mixture_run .......... equal to or_plain_run when every tail term vanishes; exact S, H
 *                             and spectra at N = 3, also with the centre-of-mass moves
 *   or_mixture_weights ...... the AM-GM bound |p_lambda|^2 <= C w on configurations deep in
 *                             the tails
 *   or_geneig, or_dsf_weights scipy.linalg.eigh; residual, S-orthonormality
 *   or_evolve ............... closed form with F = 0; dense matrix exponential for a
 *                             piecewise-constant drive
 * Reference-produced numbers themselves are parity unpinned: the reference cannot be
 * built here and publishes no outputs.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* ------------------------------------------------------------------------ */
/* Philox-4x32-10                                                            */
/* ------------------------------------------------------------------------ */
void or_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Stream layout shared with the GPU sampler:
 *   key = {seed_lo, seed_hi}
 *   ctr = {particle (or slot), sweep_lo, (sweep_hi & 0xffff) | (tag << 16), walker}
 * tag 0: Metropolis move of `particle` in sweep `sweep`
 * tag 1: initial placement of `particle` */
static void rng4(uint64_t seed, uint32_t walker, uint64_t sweep, uint32_t particle, uint32_t tag,
                 uint32_t out[4])
{
    uint32_t ctr[4] = {particle, (uint32_t)sweep, (uint32_t)((sweep >> 32) & 0xffffu) | (tag << 16), walker};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    or_philox4x32_10(ctr, key, out);
}

/* uniform in (0,1]: (x + 1) * 2^-32 */
static double u01(uint32_t x) { return ((double)x + 1.0) * 2.3283064365386963e-10; }

static void box_muller(uint32_t a, uint32_t b, double *n1, double *n2)
{
    double r = sqrt(-2.0 * log(u01(a)));
    double t = 6.283185307179586 * u01(b);
    *n1 = r * cos(t);
    *n2 = r * sin(t);
}

/* ------------------------------------------------------------------------ */
/* System description                                                        */
/* ------------------------------------------------------------------------ */
typedef struct or_system {
    int N;            /* particles */
    int m;            /* inverse filling */
    double w, eps;    /* V(r) = 0.5 w r^2 + eps r^4 + strength (r/R)^expo */
    double strength;
    int expo;
    double R;         /* sqrt(2 m N); power sums use z / R */
} or_system;

/* Basis: D states; state t has nf[t] factors (k, r) starting at fo[t]; k = 0 is the
 * empty partition (p = 1).  Sectors are contiguous: sector s covers [so[s], so[s+1]). */
typedef struct or_basis {
    int D;
    int K;                 /* largest k */
    const int *fo;         /* D+1 */
    const int *fk;
    const int *fr;
    int nsec;
    const int *so;         /* nsec+1 */
} or_basis;

double or_potential(const or_system *sys, double r2)
{
    double v = 0.5 * sys->w * r2 + sys->eps * r2 * r2;
    if (sys->strength != 0.0)
        v += sys->strength * pow(sqrt(r2) / sys->R, (double)sys->expo);
    return v;
}

/* power sums p_k, k = 0..K (p_0 = N) of u = z / R */
static void power_sums(int N, double R, const double *pos, int K, cplx *pk)
{
    for (int k = 0; k <= K; ++k) pk[k] = 0;
    pk[0] = N;
    for (int i = 0; i < N; ++i) {
        cplx u = (pos[2 * i] + I * pos[2 * i + 1]) / R;
        cplx q = u;
        for (int k = 1; k <= K; ++k) { pk[k] += q; q *= u; }
    }
}

static cplx monomial(const or_basis *b, int t, const cplx *pk)
{
    cplx v = 1;
    for (int f = b->fo[t]; f < b->fo[t + 1]; ++f) {
        int k = b->fk[f], r = b->fr[f];
        if (k == 0) continue;
        for (int q = 0; q < r; ++q) v *= pk[k];
    }
    return v;
}

static int sector_block_offset(const or_basis *b, int s)
{
    int off = 0;
    for (int q = 0; q < s; ++q) { int d = b->so[q + 1] - b->so[q]; off += d * d; }
    return off;
}

int or_block_size(const or_basis *b) { return sector_block_offset(b, b->nsec); }

/* ------------------------------------------------------------------------ */
/* Fixed-sample plain estimator: S, H block-diagonal sums (not divided by n). */
/* pos: n samples x N particles x (x, y), FP64.                               */
/* Output layout: per sector s, a d x d complex row-major block (interleaved  */
/* re/im), blocks concatenated in sector order.  S_ab = sum conj(p_a) p_b.    */
/* ------------------------------------------------------------------------ */
int or_accumulate_fixed(const or_system *sys, const or_basis *b, const double *pos, long n,
                        double *S, double *H, double *Vsum)
{
    int nb = or_block_size(b);
    memset(S, 0, sizeof(double) * 2 * nb);
    memset(H, 0, sizeof(double) * 2 * nb);
    double vs = 0;
#pragma omp parallel
    {
        double *Sl = calloc(2 * (size_t)nb, sizeof(double));
        double *Hl = calloc(2 * (size_t)nb, sizeof(double));
        cplx *pk = malloc(sizeof(cplx) * (b->K + 1));
        cplx *pl = malloc(sizeof(cplx) * b->D);
        double vl = 0;
#pragma omp for schedule(static)
        for (long s = 0; s < n; ++s) {
            const double *p = pos + (size_t)s * sys->N * 2;
            power_sums(sys->N, sys->R, p, b->K, pk);
            double V = 0;
            for (int i = 0; i < sys->N; ++i) V += or_potential(sys, p[2 * i] * p[2 * i] + p[2 * i + 1] * p[2 * i + 1]);
            vl += V;
            for (int t = 0; t < b->D; ++t) pl[t] = monomial(b, t, pk);
            for (int sec = 0; sec < b->nsec; ++sec) {
                int o = b->so[sec], d = b->so[sec + 1] - o;
                double *Sb = Sl + 2 * (size_t)sector_block_offset(b, sec);
                double *Hb = Hl + 2 * (size_t)sector_block_offset(b, sec);
                for (int a = 0; a < d; ++a) {
                    cplx ca = conj(pl[o + a]);
                    for (int c = 0; c < d; ++c) {
                        cplx x = ca * pl[o + c];
                        Sb[2 * (a * d + c)] += creal(x);
                        Sb[2 * (a * d + c) + 1] += cimag(x);
                        Hb[2 * (a * d + c)] += V * creal(x);
                        Hb[2 * (a * d + c) + 1] += V * cimag(x);
                    }
                }
            }
        }
#pragma omp critical
        {
            for (size_t q = 0; q < 2 * (size_t)nb; ++q) { S[q] += Sl[q]; H[q] += Hl[q]; }
            vs += vl;
        }
        free(Sl); free(Hl); free(pk); free(pl);
    }
    if (Vsum) *Vsum = vs;
    return 0;
}

/* Weighted cross block between sectors s1 and s2: out_ab = sum_s w_s conj(p_a) p_b
 * (d1 x d2 complex, interleaved, not divided by n).  The weights w are per sample. */
int or_accumulate_weighted(const or_system *sys, const or_basis *b, const double *pos, long n,
                           const double *w, int s1, int s2, double *out)
{
    int o1 = b->so[s1], d1 = b->so[s1 + 1] - o1;
    int o2 = b->so[s2], d2 = b->so[s2 + 1] - o2;
    memset(out, 0, sizeof(double) * 2 * (size_t)d1 * d2);
#pragma omp parallel
    {
        double *loc = calloc(2 * (size_t)d1 * d2, sizeof(double));
        cplx *pk = malloc(sizeof(cplx) * (b->K + 1));
        cplx *pl = malloc(sizeof(cplx) * b->D);
#pragma omp for schedule(static)
        for (long s = 0; s < n; ++s) {
            const double *p = pos + (size_t)s * sys->N * 2;
            power_sums(sys->N, sys->R, p, b->K, pk);
            for (int t = 0; t < d1; ++t) pl[o1 + t] = monomial(b, o1 + t, pk);
            for (int t = 0; t < d2; ++t) pl[o2 + t] = monomial(b, o2 + t, pk);
            for (int a = 0; a < d1; ++a) {
                cplx ca = conj(pl[o1 + a]) * w[s];
                for (int c = 0; c < d2; ++c) {
                    cplx x = ca * pl[o2 + c];
                    loc[2 * ((size_t)a * d2 + c)] += creal(x);
                    loc[2 * ((size_t)a * d2 + c) + 1] += cimag(x);
                }
            }
        }
#pragma omp critical
        for (size_t q = 0; q < 2 * (size_t)d1 * d2; ++q) out[q] += loc[q];
        free(loc); free(pk); free(pl);
    }
    return 0;
}

/* Features of one configuration: p_lambda for all states (complex, interleaved) and V. */
int or_features(const or_system *sys, const or_basis *b, const double *pos, double *plam, double *V)
{
    cplx *pk = malloc(sizeof(cplx) * (b->K + 1));
    power_sums(sys->N, sys->R, pos, b->K, pk);
    for (int t = 0; t < b->D; ++t) {
        cplx v = monomial(b, t, pk);
        plam[2 * t] = creal(v); plam[2 * t + 1] = cimag(v);
    }
    double v = 0;
    for (int i = 0; i < sys->N; ++i) v += or_potential(sys, pos[2 * i] * pos[2 * i] + pos[2 * i + 1] * pos[2 * i + 1]);
    *V = v;
    free(pk);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Metropolis on |p_f Psi_L|^2 (p_f = 1 gives the plain |Psi_L|^2 chain).     */
/* ------------------------------------------------------------------------ */
/* Mixture importance density w(z) |Psi_L|^2 with
 *   w = c0 + sum_{k=1..kmix} exp(a_k ln|p_k|^2 + b_k)
 * (an estimator of this build with no counterpart in the reference).  With a_k = L/k for the
 * largest sector L, every |p_lambda|^2 of every sector up to L is at most a constant times w
 * (weighted AM-GM over the parts of lambda), so the reweighted integrands p_l* p_m / w are
 * bounded; S = sum(p_l* p_m / w) / sum(1 / w) over the chain. */
typedef struct or_mixture {
    int kmix;
    double c0;
    double a[32];
    double b[32];
} or_mixture;

typedef struct chain {
    int N, m;
    double *x, *y;          /* positions */
    const or_basis *b;
    int f;                  /* sampling state, -1 for the plain chain */
    double R;
    const or_mixture *mx;   /* mixture density (f = -1), or NULL */
} chain;

static double mix_log_w(const or_mixture *mx, const cplx *pk)
{
    double w = mx->c0;
    for (int k = 1; k <= mx->kmix; ++k) {
        double q = creal(pk[k]) * creal(pk[k]) + cimag(pk[k]) * cimag(pk[k]);
        if (q > 0.0) w += exp(mx->a[k - 1] * log(q) + mx->b[k - 1]);
    }
    return log(w);
}

/* power sums of the chain's current positions (used to start the incremental update) */
static void chain_power_sums(const chain *c, cplx *pk)
{
    int K = c->mx ? c->mx->kmix : (c->b ? c->b->K : 0);
    for (int k = 0; k <= K; ++k) pk[k] = 0;
    pk[0] = c->N;
    for (int j = 0; j < c->N; ++j) {
        cplx u = (c->x[j] + I * c->y[j]) / c->R;
        cplx q = u;
        for (int k = 1; k <= K; ++k) { pk[k] += q; q *= u; }
    }
}

static double log_abs2(cplx v) { return log(creal(v) * creal(v) + cimag(v) * cimag(v)); }

/* One sweep: particle i = 0..N-1 in order; returns accepted moves.
 * log A = log|p_f'|^2 - log|p_f|^2 + m sum_{j!=i} log(|z'-z_j|^2/|z-z_j|^2) - (|z'|^2-|z|^2)/2
 * (em_markovchain.cu:282-328: the ratio of |p_f Psi_L|^2 with Psi_L = prod (z_i-z_j)^m e^{-|z|^2/4});
 * the power sums of the sampling state are updated per move by -u_old^k + u_new^k
 * (em_markovchain.cu:201-206). */
static int sweep(chain *c, uint64_t seed, uint32_t walker, uint64_t sweep_idx, double step, cplx *cur_pk)
{
    int acc = 0;
    int K = (c->f >= 0) ? c->b->K : (c->mx ? c->mx->kmix : 0);
    cplx prop[64];
    for (int i = 0; i < c->N; ++i) {
        uint32_t r[4];
        rng4(seed, walker, sweep_idx, (uint32_t)i, 0, r);
        double n1, n2;
        box_muller(r[0], r[1], &n1, &n2);
        double xo = c->x[i], yo = c->y[i];
        double xn = xo + step * n1, yn = yo + step * n2;
        double lr = 0.0;
        for (int j = 0; j < c->N; ++j) {
            if (j == i) continue;
            double ax = xn - c->x[j], ay = yn - c->y[j];
            double bx = xo - c->x[j], by = yo - c->y[j];
            lr += log((ax * ax + ay * ay) / (bx * bx + by * by));
        }
        double logA = c->m * lr - 0.5 * ((xn * xn + yn * yn) - (xo * xo + yo * yo));
        if (c->f >= 0) {
            cplx uo = (xo + I * yo) / c->R, un = (xn + I * yn) / c->R;
            cplx qo = uo, qn = un;
            prop[0] = cur_pk[0];
            for (int k = 1; k <= K; ++k) { prop[k] = cur_pk[k] - qo + qn; qo *= uo; qn *= un; }
            logA += log_abs2(monomial(c->b, c->f, prop)) - log_abs2(monomial(c->b, c->f, cur_pk));
        } else if (c->mx) {
            cplx uo = (xo + I * yo) / c->R, un = (xn + I * yn) / c->R;
            cplx qo = uo, qn = un;
            prop[0] = cur_pk[0];
            for (int k = 1; k <= K; ++k) { prop[k] = cur_pk[k] - qo + qn; qo *= uo; qn *= un; }
            logA += mix_log_w(c->mx, prop) - mix_log_w(c->mx, cur_pk);
        }
        double u = u01(r[2]);
        if (logA >= 0.0 || log(u) < logA) {
            c->x[i] = xn; c->y[i] = yn;
            if (c->f >= 0 || c->mx) for (int k = 0; k <= K; ++k) cur_pk[k] = prop[k];
            ++acc;
        }
    }
    return acc;
}

/* initial positions: uniform in the disc of radius R (tag 1 stream) */
static void init_positions(chain *c, uint64_t seed, uint32_t walker)
{
    for (int i = 0; i < c->N; ++i) {
        uint32_t r[4];
        rng4(seed, walker, 0, (uint32_t)i, 1, r);
        double rad = c->R * sqrt(u01(r[0]));
        double th = 6.283185307179586 * u01(r[1]);
        c->x[i] = rad * cos(th);
        c->y[i] = rad * sin(th);
    }
}

/* ------------------------------------------------------------------------ */
/* Plain |Psi_L|^2 chains: S, H sums over walkers x samples (not divided).    */
/* Optional: final positions, per-sample trace of (p_1..p_4, V) for walker 0..*/
/* ------------------------------------------------------------------------ */
int or_plain_run(const or_system *sys, const or_basis *b, int walkers, int burnin, int samples,
                 int thin, double step, uint64_t seed, int threads,
                 double *S, double *H, double *acceptance, double *final_pos, double *r2_mean)
{
    int nb = or_block_size(b);
    memset(S, 0, sizeof(double) * 2 * nb);
    memset(H, 0, sizeof(double) * 2 * nb);
    long acc_total = 0;
    double r2_total = 0.0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
#pragma omp parallel
    {
        double *Sl = calloc(2 * (size_t)nb, sizeof(double));
        double *Hl = calloc(2 * (size_t)nb, sizeof(double));
        double *buf = malloc(sizeof(double) * 2 * sys->N);
        double *pl = malloc(sizeof(double) * 2 * b->D);
        long acc_l = 0;
        double r2_l = 0.0;
#pragma omp for schedule(dynamic, 1)
        for (int wk = 0; wk < walkers; ++wk) {
            chain c = {sys->N, sys->m, malloc(sizeof(double) * sys->N), malloc(sizeof(double) * sys->N), b, -1, sys->R, NULL};
            cplx pk[64];
            init_positions(&c, seed, (uint32_t)wk);
            uint64_t sw = 0;
            for (int t = 0; t < burnin; ++t) sweep(&c, seed, (uint32_t)wk, sw++, step, pk);
            for (int s = 0; s < samples; ++s) {
                for (int t = 0; t < thin; ++t) acc_l += sweep(&c, seed, (uint32_t)wk, sw++, step, pk);
                for (int i = 0; i < sys->N; ++i) { buf[2 * i] = c.x[i]; buf[2 * i + 1] = c.y[i]; }
                double V;
                or_features(sys, b, buf, pl, &V);
                for (int i = 0; i < sys->N; ++i) r2_l += c.x[i] * c.x[i] + c.y[i] * c.y[i];
                for (int sec = 0; sec < b->nsec; ++sec) {
                    int o = b->so[sec], d = b->so[sec + 1] - o;
                    size_t bo = 2 * (size_t)sector_block_offset(b, sec);
                    for (int a = 0; a < d; ++a) {
                        double ar = pl[2 * (o + a)], ai = -pl[2 * (o + a) + 1];
                        for (int q = 0; q < d; ++q) {
                            double br = pl[2 * (o + q)], bi = pl[2 * (o + q) + 1];
                            double xr = ar * br - ai * bi, xi = ar * bi + ai * br;
                            Sl[bo + 2 * (a * d + q)] += xr;
                            Sl[bo + 2 * (a * d + q) + 1] += xi;
                            Hl[bo + 2 * (a * d + q)] += V * xr;
                            Hl[bo + 2 * (a * d + q) + 1] += V * xi;
                        }
                    }
                }
            }
            if (final_pos)
                for (int i = 0; i < sys->N; ++i) {
                    final_pos[((size_t)wk * sys->N + i) * 2] = c.x[i];
                    final_pos[((size_t)wk * sys->N + i) * 2 + 1] = c.y[i];
                }
            free(c.x); free(c.y);
        }
#pragma omp critical
        {
            for (size_t q = 0; q < 2 * (size_t)nb; ++q) { S[q] += Sl[q]; H[q] += Hl[q]; }
            acc_total += acc_l;
            r2_total += r2_l;
        }
        free(Sl); free(Hl); free(buf); free(pl);
    }
    if (acceptance) *acceptance = (double)acc_total / ((double)walkers * samples * thin * sys->N);
    if (r2_mean) *r2_mean = r2_total / ((double)walkers * samples);
    return 0;
}

/* Plain chain positions only: write `samples` kept configurations of each walker. */
int or_plain_positions(const or_system *sys, int walkers, int burnin, int samples, int thin,
                       double step, uint64_t seed, double *out, double *acceptance)
{
    long acc = 0;
#pragma omp parallel for reduction(+:acc) schedule(dynamic, 1)
    for (int wk = 0; wk < walkers; ++wk) {
        chain c = {sys->N, sys->m, malloc(sizeof(double) * sys->N), malloc(sizeof(double) * sys->N), NULL, -1, sys->R, NULL};
        cplx pk[64];
        init_positions(&c, seed, (uint32_t)wk);
        uint64_t sw = 0;
        for (int t = 0; t < burnin; ++t) sweep(&c, seed, (uint32_t)wk, sw++, step, pk);
        for (int s = 0; s < samples; ++s) {
            for (int t = 0; t < thin; ++t) acc += sweep(&c, seed, (uint32_t)wk, sw++, step, pk);
            double *o = out + (((size_t)wk * samples + s) * sys->N) * 2;
            for (int i = 0; i < sys->N; ++i) { o[2 * i] = c.x[i]; o[2 * i + 1] = c.y[i]; }
        }
        free(c.x); free(c.y);
    }
    if (acceptance) *acceptance = (double)acc / ((double)walkers * samples * thin * sys->N);
    return 0;
}

/* w(z) of n configurations (n x N x 2) for the mixture density mx */
int or_mixture_weights(const or_system *sys, const or_mixture *mx, const double *pos, long n, double *w)
{
    if (mx->kmix < 1 || mx->kmix > 31) return -1;
    cplx pk[32];
    for (long s = 0; s < n; ++s) {
        power_sums(sys->N, sys->R, pos + (size_t)s * sys->N * 2, mx->kmix, pk);
        w[s] = exp(mix_log_w(mx, pk));
    }
    return 0;
}

/* Exact move of the mixture density: every particle translated by d (Gaussian, variance 4 / N
 * per component; |Psi_L|^2 changes by exp(-N (|Z + d|^2 - |Z|^2) / 2), Z the centre of mass),
 * accepted with that factor times w(z + d) / w(z), then a rotation by a uniform angle (w and
 * |Psi_L|^2 unchanged).  Philox block (walker, sweep_lo, sweep_hi16 | 6 << 16, 0xfffffff1):
 * normals from words 0, 1, the acceptance from word 2, the angle from word 3. */
static int mixture_com_move(chain *c, uint64_t seed, uint32_t walker, uint64_t sweep_idx)
{
    int N = c->N;
    uint32_t r[4];
    rng4(seed, 0xfffffff1u, sweep_idx, walker, 6, r);
    double g1, g2;
    box_muller(r[0], r[1], &g1, &g2);
    double sd = 2.0 / sqrt((double)N);
    double dx = g1 * sd, dy = g2 * sd;
    double cx = 0.0, cy = 0.0;
    for (int i = 0; i < N; ++i) { cx += c->x[i]; cy += c->y[i]; }
    cx /= N; cy /= N;
    cplx p0[32], p1[32];
    for (int k = 0; k <= c->mx->kmix; ++k) p0[k] = p1[k] = 0;
    for (int i = 0; i < N; ++i) {
        cplx u0 = (c->x[i] + I * c->y[i]) / c->R, u1 = ((c->x[i] + dx) + I * (c->y[i] + dy)) / c->R;
        cplx q0 = u0, q1 = u1;
        for (int k = 1; k <= c->mx->kmix; ++k) { p0[k] += q0; p1[k] += q1; q0 *= u0; q1 *= u1; }
    }
    double zx = cx + dx, zy = cy + dy;
    double logA = -0.5 * N * ((zx * zx + zy * zy) - (cx * cx + cy * cy)) + (mix_log_w(c->mx, p1) - mix_log_w(c->mx, p0));
    int acc = logA >= 0.0 || log(u01(r[2])) < logA;
    double phi = 6.283185307179586 * u01(r[3]), cp = cos(phi), sp = sin(phi);
    for (int i = 0; i < N; ++i) {
        double x = c->x[i] + (acc ? dx : 0.0), y = c->y[i] + (acc ? dy : 0.0);
        c->x[i] = cp * x - sp * y;
        c->y[i] = sp * x + cp * y;
    }
    return acc;
}

/* ------------------------------------------------------------------------ */
/* Mixture chains on w |Psi_L|^2 (or_mixture above): sums over walkers x     */
/* samples of p_l* p_m / w and V p_l* p_m / w (S, H, block layout of         */
/* or_plain_run) and of 1 / w (wsum); w from the power sums of each kept     */
/* configuration.  The chains start from init_positions and update the      */
/* running p_k per move like the row-wise chains.                           */
/* ------------------------------------------------------------------------ */
int or_mixture_run(const or_system *sys, const or_basis *b, const or_mixture *mx, int walkers, int burnin,
                   int samples, int thin, double step, uint64_t seed, int threads, int com_moves, double *S,
                   double *H, double *wsum, double *acceptance, double *final_pos)
{
    int nb = or_block_size(b);
    memset(S, 0, sizeof(double) * 2 * nb);
    memset(H, 0, sizeof(double) * 2 * nb);
    if (mx->kmix < 1 || mx->kmix > 31 || mx->kmix > b->K) return -1;
    long acc_total = 0;
    double ws_total = 0.0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
#pragma omp parallel
    {
        double *Sl = calloc(2 * (size_t)nb, sizeof(double));
        double *Hl = calloc(2 * (size_t)nb, sizeof(double));
        double *buf = malloc(sizeof(double) * 2 * sys->N);
        double *pl = malloc(sizeof(double) * 2 * b->D);
        cplx *fk = malloc(sizeof(cplx) * (b->K + 1));
        long acc_l = 0;
        double ws_l = 0.0;
#pragma omp for schedule(dynamic, 1)
        for (int wk = 0; wk < walkers; ++wk) {
            chain c = {sys->N, sys->m, malloc(sizeof(double) * sys->N), malloc(sizeof(double) * sys->N), b, -1, sys->R, mx};
            cplx pk[64];
            init_positions(&c, seed, (uint32_t)wk);
            chain_power_sums(&c, pk);
            uint64_t sw = 0;
            for (int t = 0; t < burnin; ++t) sweep(&c, seed, (uint32_t)wk, sw++, step, pk);
            for (int s = 0; s < samples; ++s) {
                for (int t = 0; t < thin; ++t) acc_l += sweep(&c, seed, (uint32_t)wk, sw++, step, pk);
                if (com_moves) {
                    mixture_com_move(&c, seed, (uint32_t)wk, sw);
                    chain_power_sums(&c, pk);   /* the next sweep starts from the moved walker's p_k */
                }
                for (int i = 0; i < sys->N; ++i) { buf[2 * i] = c.x[i]; buf[2 * i + 1] = c.y[i]; }
                double V;
                or_features(sys, b, buf, pl, &V);
                power_sums(sys->N, sys->R, buf, b->K, fk);
                const double iw = exp(-mix_log_w(mx, fk));
                ws_l += iw;
                for (int sec = 0; sec < b->nsec; ++sec) {
                    int o = b->so[sec], d = b->so[sec + 1] - o;
                    size_t bo = 2 * (size_t)sector_block_offset(b, sec);
                    for (int a = 0; a < d; ++a) {
                        double ar = pl[2 * (o + a)], ai = -pl[2 * (o + a) + 1];
                        for (int q = 0; q < d; ++q) {
                            double br = pl[2 * (o + q)], bi = pl[2 * (o + q) + 1];
                            double xr = (ar * br - ai * bi) * iw, xi = (ar * bi + ai * br) * iw;
                            Sl[bo + 2 * (a * d + q)] += xr;
                            Sl[bo + 2 * (a * d + q) + 1] += xi;
                            Hl[bo + 2 * (a * d + q)] += V * xr;
                            Hl[bo + 2 * (a * d + q) + 1] += V * xi;
                        }
                    }
                }
            }
            if (final_pos)
                for (int i = 0; i < sys->N; ++i) {
                    final_pos[((size_t)wk * sys->N + i) * 2] = c.x[i];
                    final_pos[((size_t)wk * sys->N + i) * 2 + 1] = c.y[i];
                }
            free(c.x); free(c.y);
        }
#pragma omp critical
        {
            for (size_t q = 0; q < 2 * (size_t)nb; ++q) { S[q] += Sl[q]; H[q] += Hl[q]; }
            acc_total += acc_l;
            ws_total += ws_l;
        }
        free(Sl); free(Hl); free(buf); free(pl); free(fk);
    }
    if (wsum) *wsum = ws_total;
    if (acceptance) *acceptance = (double)acc_total / ((double)walkers * samples * thin * sys->N);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Reference row-wise estimator for one sector (em_spectrum.cu:16-137,       */
/* 318-390): for each row f, chains on |p_f Psi_L|^2 accumulate g/f, V g/f,   */
/* |g/f|^2; then H(f,g) = <V g/f>/sqrt<|g/f|^2>, M(f,g) = <g/f>/sqrt<|g/f|^2>,*/
/* then (X + X^H)/2.  Outputs d x d complex row-major (interleaved).          */
/* Optional Vc/Vs rows (cos/sin(dLexc theta) sums, em_dsf.cu:113-134) when     */
/* dl_exc > 0; optional final positions of every chain (final_pos).           */
/* ------------------------------------------------------------------------ */
int or_rowwise_sector(const or_system *sys, const or_basis *b, int sec, int walkers, int burnin,
                      int samples, double step, uint64_t seed, int threads, int dl_exc,
                      double *Hout, double *Mout, double *VCout, double *VSout, double *acceptance,
                      double *final_pos)
{
    int o = b->so[sec], d = b->so[sec + 1] - o;
    double *Hr = calloc(2 * (size_t)d * d, sizeof(double));
    double *Mr = calloc(2 * (size_t)d * d, sizeof(double));
    double *VCr = calloc(2 * (size_t)d * d, sizeof(double));
    double *VSr = calloc(2 * (size_t)d * d, sizeof(double));
    long acc_total = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    for (int f = 0; f < d; ++f) {
        double *fHg = calloc(2 * (size_t)d, sizeof(double));
        double *fg = calloc(2 * (size_t)d, sizeof(double));
        double *gg = calloc((size_t)d, sizeof(double));
        double *fVC = calloc(2 * (size_t)d, sizeof(double));
        double *fVS = calloc(2 * (size_t)d, sizeof(double));
        long acc_row = 0;
#pragma omp parallel
        {
            double *lH = calloc(2 * (size_t)d, sizeof(double));
            double *lg = calloc(2 * (size_t)d, sizeof(double));
            double *lgg = calloc((size_t)d, sizeof(double));
            double *lC = calloc(2 * (size_t)d, sizeof(double));
            double *lS = calloc(2 * (size_t)d, sizeof(double));
            double *buf = malloc(sizeof(double) * 2 * sys->N);
            double *pl = malloc(sizeof(double) * 2 * b->D);
            long la = 0;
#pragma omp for schedule(dynamic, 1)
            for (int wk = 0; wk < walkers; ++wk) {
                chain c = {sys->N, sys->m, malloc(sizeof(double) * sys->N), malloc(sizeof(double) * sys->N), b, o + f, sys->R, NULL};
                /* distinct stream per row: walker id offset by row */
                uint32_t wid = (uint32_t)(wk + (f + 1) * walkers);
                init_positions(&c, seed, wid);
                cplx pk[64];
                chain_power_sums(&c, pk);
                uint64_t sw = 0;
                for (int t = 0; t < burnin; ++t) sweep(&c, seed, wid, sw++, step, pk);
                double *wH = calloc(2 * (size_t)d, sizeof(double));
                double *wg = calloc(2 * (size_t)d, sizeof(double));
                double *wgg = calloc((size_t)d, sizeof(double));
                double *wC = calloc(2 * (size_t)d, sizeof(double));
                double *wS = calloc(2 * (size_t)d, sizeof(double));
                for (int s = 0; s < samples; ++s) {
                    la += sweep(&c, seed, wid, sw++, step, pk);
                    for (int i = 0; i < sys->N; ++i) { buf[2 * i] = c.x[i]; buf[2 * i + 1] = c.y[i]; }
                    double V;
                    or_features(sys, b, buf, pl, &V);
                    double vc = 0, vs = 0;
                    if (dl_exc > 0)
                        for (int i = 0; i < sys->N; ++i) {
                            double th = atan2(c.y[i], c.x[i]);
                            vc += cos(dl_exc * th);
                            vs += sin(dl_exc * th);
                        }
                    cplx pf = pl[2 * (o + f)] + I * pl[2 * (o + f) + 1];
                    for (int g = 0; g < d; ++g) {
                        cplx gf = (g == f) ? 1.0 : (pl[2 * (o + g)] + I * pl[2 * (o + g) + 1]) / pf;
                        wH[2 * g] += V * creal(gf); wH[2 * g + 1] += V * cimag(gf);
                        wg[2 * g] += creal(gf); wg[2 * g + 1] += cimag(gf);
                        wgg[g] += creal(gf) * creal(gf) + cimag(gf) * cimag(gf);
                        wC[2 * g] += vc * creal(gf); wC[2 * g + 1] += vc * cimag(gf);
                        wS[2 * g] += vs * creal(gf); wS[2 * g + 1] += vs * cimag(gf);
                    }
                }
                for (int g = 0; g < d; ++g) {
                    lH[2 * g] += wH[2 * g] / samples; lH[2 * g + 1] += wH[2 * g + 1] / samples;
                    lg[2 * g] += wg[2 * g] / samples; lg[2 * g + 1] += wg[2 * g + 1] / samples;
                    lgg[g] += wgg[g] / samples;
                    lC[2 * g] += wC[2 * g] / samples; lC[2 * g + 1] += wC[2 * g + 1] / samples;
                    lS[2 * g] += wS[2 * g] / samples; lS[2 * g + 1] += wS[2 * g + 1] / samples;
                }
                free(wH); free(wg); free(wgg); free(wC); free(wS);
                if (final_pos)   /* chain (row f, walker wk): [f walkers + wk][N][2] */
                    for (int i = 0; i < sys->N; ++i) {
                        final_pos[(((size_t)f * walkers + wk) * sys->N + i) * 2] = c.x[i];
                        final_pos[(((size_t)f * walkers + wk) * sys->N + i) * 2 + 1] = c.y[i];
                    }
                free(c.x); free(c.y);
            }
#pragma omp critical
            {
                for (int g = 0; g < d; ++g) {
                    fHg[2 * g] += lH[2 * g]; fHg[2 * g + 1] += lH[2 * g + 1];
                    fg[2 * g] += lg[2 * g]; fg[2 * g + 1] += lg[2 * g + 1];
                    gg[g] += lgg[g];
                    fVC[2 * g] += lC[2 * g]; fVC[2 * g + 1] += lC[2 * g + 1];
                    fVS[2 * g] += lS[2 * g]; fVS[2 * g + 1] += lS[2 * g + 1];
                }
                acc_row += la;
            }
            free(lH); free(lg); free(lgg); free(lC); free(lS); free(buf); free(pl);
        }
        for (int g = 0; g < d; ++g) {
            double nrm = sqrt(gg[g] / walkers);
            size_t ix = 2 * ((size_t)f * d + g);
            Hr[ix] = fHg[2 * g] / walkers / nrm; Hr[ix + 1] = fHg[2 * g + 1] / walkers / nrm;
            Mr[ix] = fg[2 * g] / walkers / nrm; Mr[ix + 1] = fg[2 * g + 1] / walkers / nrm;
            VCr[ix] = fVC[2 * g] / walkers / nrm; VCr[ix + 1] = fVC[2 * g + 1] / walkers / nrm;
            VSr[ix] = fVS[2 * g] / walkers / nrm; VSr[ix + 1] = fVS[2 * g + 1] / walkers / nrm;
        }
        acc_total += acc_row;
        free(fHg); free(fg); free(gg); free(fVC); free(fVS);
    }
    /* (X + X^H) / 2 */
    for (int a = 0; a < d; ++a)
        for (int c = 0; c < d; ++c) {
            size_t ac = 2 * ((size_t)a * d + c), ca = 2 * ((size_t)c * d + a);
            Hout[ac] = 0.5 * (Hr[ac] + Hr[ca]); Hout[ac + 1] = 0.5 * (Hr[ac + 1] - Hr[ca + 1]);
            Mout[ac] = 0.5 * (Mr[ac] + Mr[ca]); Mout[ac + 1] = 0.5 * (Mr[ac + 1] - Mr[ca + 1]);
            if (VCout) { VCout[ac] = 0.5 * (VCr[ac] + VCr[ca]); VCout[ac + 1] = 0.5 * (VCr[ac + 1] - VCr[ca + 1]); }
            if (VSout) { VSout[ac] = 0.5 * (VSr[ac] + VSr[ca]); VSout[ac + 1] = 0.5 * (VSr[ac + 1] - VSr[ca + 1]); }
        }
    if (acceptance) *acceptance = (double)acc_total / ((double)d * walkers * samples * sys->N);
    free(Hr); free(Mr); free(VCr); free(VSr);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* FP64 generalized Hermitian eigenproblem H v = E S v                        */
/* Cholesky S = L L^H, C = L^-1 H L^-H, cyclic complex Jacobi on C,           */
/* v = L^-H W (S-normalised).  Eigenvalues ascending.                         */
/* Returns 0, or -1 if S is not positive definite.                            */
/* ------------------------------------------------------------------------ */
int or_geneig(int d, const double *Sin, const double *Hin, double *evals, double *evecs)
{
    cplx *L = malloc(sizeof(cplx) * d * d);
    cplx *C = malloc(sizeof(cplx) * d * d);
    cplx *W = malloc(sizeof(cplx) * d * d);
    cplx *T = malloc(sizeof(cplx) * d * d);
    for (int i = 0; i < d * d; ++i) { L[i] = Sin[2 * i] + I * Sin[2 * i + 1]; C[i] = Hin[2 * i] + I * Hin[2 * i + 1]; }
    /* Cholesky, lower triangle, S = L L^H */
    for (int j = 0; j < d; ++j) {
        double s = creal(L[j * d + j]);
        for (int k = 0; k < j; ++k) s -= creal(L[j * d + k] * conj(L[j * d + k]));
        if (!(s > 0.0)) { free(L); free(C); free(W); free(T); return -1; }
        double ljj = sqrt(s);
        L[j * d + j] = ljj;
        for (int i = j + 1; i < d; ++i) {
            cplx t = L[i * d + j];
            for (int k = 0; k < j; ++k) t -= L[i * d + k] * conj(L[j * d + k]);
            L[i * d + j] = t / ljj;
        }
        for (int i = 0; i < j; ++i) L[i * d + j] = 0;
    }
    /* T = L^-1 H (forward substitution on columns) */
    for (int c = 0; c < d; ++c)
        for (int i = 0; i < d; ++i) {
            cplx t = C[i * d + c];
            for (int k = 0; k < i; ++k) t -= L[i * d + k] * T[k * d + c];
            T[i * d + c] = t / L[i * d + i];
        }
    /* C = T L^-H  <=>  C^H = L^-1 T^H: solve on rows */
    for (int r = 0; r < d; ++r)
        for (int i = 0; i < d; ++i) {
            /* (L^-1 T^H)[i][r] = (t_i) ; T^H[i][r] = conj(T[r][i]) */
            cplx t = conj(T[r * d + i]);
            for (int k = 0; k < i; ++k) t -= L[i * d + k] * W[k * d + r];
            W[i * d + r] = t / L[i * d + i];
        }
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) C[i * d + j] = conj(W[j * d + i]);
    /* symmetrise exactly */
    for (int i = 0; i < d; ++i) {
        C[i * d + i] = creal(C[i * d + i]);
        for (int j = i + 1; j < d; ++j) {
            cplx a = 0.5 * (C[i * d + j] + conj(C[j * d + i]));
            C[i * d + j] = a; C[j * d + i] = conj(a);
        }
    }
    for (int i = 0; i < d * d; ++i) W[i] = 0;
    for (int i = 0; i < d; ++i) W[i * d + i] = 1;
    /* cyclic Jacobi */
    for (int it = 0; it < 100; ++it) {
        double off = 0, tot = 0;
        for (int i = 0; i < d; ++i)
            for (int j = 0; j < d; ++j) {
                double a2 = creal(C[i * d + j] * conj(C[i * d + j]));
                tot += a2;
                if (i != j) off += a2;
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int p = 0; p < d - 1; ++p)
            for (int q = p + 1; q < d; ++q) {
                cplx apq = C[p * d + q];
                double aabs = cabs(apq);
                if (aabs < 1e-300) continue;
                double app = creal(C[p * d + p]), aqq = creal(C[q * d + q]);
                cplx ph = apq / aabs; /* e^{i phi} */
                double tau = (aqq - app) / (2.0 * aabs);
                double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                double cs = 1.0 / sqrt(1.0 + t * t), sn = t * cs;
                /* rotation G with columns p,q: G_pp = c, G_pq = s e^{i phi}, G_qp = -s e^{-i phi}, G_qq = c
                 * C <- G^H C G zeroes C_pq */
                for (int k = 0; k < d; ++k) { /* columns */
                    cplx ckp = C[k * d + p], ckq = C[k * d + q];
                    C[k * d + p] = cs * ckp - sn * conj(ph) * ckq;
                    C[k * d + q] = sn * ph * ckp + cs * ckq;
                }
                for (int k = 0; k < d; ++k) { /* rows */
                    cplx cpk = C[p * d + k], cqk = C[q * d + k];
                    C[p * d + k] = cs * cpk - sn * ph * cqk;
                    C[q * d + k] = sn * conj(ph) * cpk + cs * cqk;
                }
                C[p * d + q] = 0; C[q * d + p] = 0;
                for (int k = 0; k < d; ++k) {
                    cplx wkp = W[k * d + p], wkq = W[k * d + q];
                    W[k * d + p] = cs * wkp - sn * conj(ph) * wkq;
                    W[k * d + q] = sn * ph * wkp + cs * wkq;
                }
            }
    }
    /* sort ascending */
    int *idx = malloc(sizeof(int) * d);
    for (int i = 0; i < d; ++i) idx[i] = i;
    for (int i = 0; i < d; ++i)
        for (int j = i + 1; j < d; ++j)
            if (creal(C[idx[j] * d + idx[j]]) < creal(C[idx[i] * d + idx[i]])) { int t = idx[i]; idx[i] = idx[j]; idx[j] = t; }
    for (int i = 0; i < d; ++i) evals[i] = creal(C[idx[i] * d + idx[i]]);
    if (evecs) {
        /* v = L^-H w : back substitution with L^H (upper) */
        for (int c = 0; c < d; ++c) {
            int src = idx[c];
            for (int i = d - 1; i >= 0; --i) {
                cplx t = W[i * d + src];
                for (int k = i + 1; k < d; ++k) t -= conj(L[k * d + i]) * T[k * d + c];
                T[i * d + c] = t / L[i * d + i];
            }
        }
        for (int i = 0; i < d * d; ++i) { evecs[2 * i] = creal(T[i]); evecs[2 * i + 1] = cimag(T[i]); }
    }
    free(idx); free(L); free(C); free(W); free(T);
    return 0;
}

/* DSF weights (generalized_eigenproblem.cpp:54-95): drop row/col 0 of the
 * (1+d) x (1+d) matrices, solve the d x d block with eigenvectors, and return
 * E_j and |sum_a (VC - i VS)[0][1+a] v_j[a]|^2.  Returns 0 or -1 (not PD). */
int or_dsf_weights(int dp1, const double *H, const double *M, const double *VC, const double *VS,
                   double *energies, double *weights)
{
    int d = dp1 - 1;
    if (d <= 0) return -1;
    double *h = malloc(sizeof(double) * 2 * d * d), *m = malloc(sizeof(double) * 2 * d * d);
    double *v = malloc(sizeof(double) * 2 * d * d);
    for (int a = 0; a < d; ++a)
        for (int c = 0; c < d; ++c) {
            size_t s = 2 * ((size_t)(a + 1) * dp1 + (c + 1)), t = 2 * ((size_t)a * d + c);
            h[t] = H[s]; h[t + 1] = H[s + 1]; m[t] = M[s]; m[t + 1] = M[s + 1];
        }
    int rc = or_geneig(d, m, h, energies, v);
    if (rc == 0)
        for (int j = 0; j < d; ++j) {
            cplx me = 0;
            for (int a = 0; a < d; ++a) {
                cplx va = VC[2 * (1 + a)] + I * VC[2 * (1 + a) + 1];
                cplx vs = VS[2 * (1 + a)] + I * VS[2 * (1 + a) + 1];
                cplx ev = v[2 * ((size_t)a * d + j)] + I * v[2 * ((size_t)a * d + j) + 1];
                me += (va - I * vs) * ev;
            }
            weights[j] = creal(me * conj(me));
        }
    free(h); free(m); free(v);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Time evolution (modules/dynamics/dynamics.cpp:20-53, 73-125, 264-290):      */
/* dC/dt = -i M^-1 (H + F(t) U) C, C(0) = e_0, integrated with an adaptive     */
/* Cash-Karp 5(4) Runge-Kutta (controlled stepper: err_i = |xerr_i| /          */
/* (eps_abs + eps_rel (|x_i| + dt |dxdt_i|)), reject if max > 1 with           */
/* dt *= max(0.9 err^-1/3, 0.2), grow if max < 0.5 with dt *= 0.9 err^-1/5      */
/* capped at 5x), observed at t_k = k save_step, k < ceil(T / save_step):       */
/* rho_l(t) = C^H rho_l C and the norm C^H M C.                                */
/* ------------------------------------------------------------------------ */
typedef double (*or_time_fn)(void *ctx, double t);

static void cmatvec(int D, const cplx *A, const cplx *x, cplx *y)
{
    for (int i = 0; i < D; ++i) {
        cplx s = 0;
        for (int j = 0; j < D; ++j) s += A[(size_t)i * D + j] * x[j];
        y[i] = s;
    }
}

/* dC = -i (Heff + F Ueff) C */
static void rhs(int D, const cplx *Heff, const cplx *Ueff, double F, const cplx *C, cplx *out, cplx *tmp)
{
    cmatvec(D, Heff, C, out);
    cmatvec(D, Ueff, C, tmp);
    for (int i = 0; i < D; ++i) out[i] = -I * (out[i] + F * tmp[i]);
}

/* inverse of a complex matrix by Gauss-Jordan with partial pivoting (M^-1 of dynamics.cpp:25) */
static int cinverse(int D, const cplx *A, cplx *Ainv)
{
    cplx *a = malloc(sizeof(cplx) * D * D);
    memcpy(a, A, sizeof(cplx) * D * D);
    for (int i = 0; i < D * D; ++i) Ainv[i] = 0;
    for (int i = 0; i < D; ++i) Ainv[i * D + i] = 1;
    for (int c = 0; c < D; ++c) {
        int p = c;
        for (int r = c + 1; r < D; ++r) if (cabs(a[r * D + c]) > cabs(a[p * D + c])) p = r;
        if (cabs(a[p * D + c]) == 0.0) { free(a); return -1; }
        if (p != c)
            for (int k = 0; k < D; ++k) {
                cplx t = a[c * D + k]; a[c * D + k] = a[p * D + k]; a[p * D + k] = t;
                t = Ainv[c * D + k]; Ainv[c * D + k] = Ainv[p * D + k]; Ainv[p * D + k] = t;
            }
        cplx piv = a[c * D + c];
        for (int k = 0; k < D; ++k) { a[c * D + k] /= piv; Ainv[c * D + k] /= piv; }
        for (int r = 0; r < D; ++r) {
            if (r == c) continue;
            cplx f = a[r * D + c];
            if (f == 0) continue;
            for (int k = 0; k < D; ++k) { a[r * D + k] -= f * a[c * D + k]; Ainv[r * D + k] -= f * Ainv[c * D + k]; }
        }
    }
    free(a);
    return 0;
}

int or_evolve(int D, const double *Min, const double *Hin, const double *Uin, or_time_fn F, void *ctx,
              double end_time, double save_step, double eps_abs, double eps_rel, int nrho, const double *rhoin,
              double *out_rho, double *out_norm, long *n_steps)
{
    size_t DD = (size_t)D * D;
    cplx *M = malloc(sizeof(cplx) * DD), *Mi = malloc(sizeof(cplx) * DD);
    cplx *Heff = malloc(sizeof(cplx) * DD), *Ueff = malloc(sizeof(cplx) * DD), *A = malloc(sizeof(cplx) * DD);
    for (size_t q = 0; q < DD; ++q) M[q] = Min[2 * q] + I * Min[2 * q + 1];
    if (cinverse(D, M, Mi)) { free(M); free(Mi); free(Heff); free(Ueff); free(A); return -1; }
    for (size_t q = 0; q < DD; ++q) A[q] = Hin[2 * q] + I * Hin[2 * q + 1];
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
            cplx s = 0;
            for (int k = 0; k < D; ++k) s += Mi[(size_t)i * D + k] * A[(size_t)k * D + j];
            Heff[(size_t)i * D + j] = s;
        }
    for (size_t q = 0; q < DD; ++q) A[q] = Uin[2 * q] + I * Uin[2 * q + 1];
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
            cplx s = 0;
            for (int k = 0; k < D; ++k) s += Mi[(size_t)i * D + k] * A[(size_t)k * D + j];
            Ueff[(size_t)i * D + j] = s;
        }
    cplx *C = calloc(D, sizeof(cplx)), *k[6], *y = malloc(sizeof(cplx) * D), *tmp = malloc(sizeof(cplx) * D);
    cplx *Cn = malloc(sizeof(cplx) * D), *err = malloc(sizeof(cplx) * D);
    for (int s = 0; s < 6; ++s) k[s] = malloc(sizeof(cplx) * D);
    C[0] = 1;
    static const double c[6] = {0.0, 0.2, 0.3, 0.6, 1.0, 0.875};
    static const double a[6][5] = {{0, 0, 0, 0, 0},
                                   {0.2, 0, 0, 0, 0},
                                   {3.0 / 40, 9.0 / 40, 0, 0, 0},
                                   {0.3, -0.9, 1.2, 0, 0},
                                   {-11.0 / 54, 2.5, -70.0 / 27, 35.0 / 27, 0},
                                   {1631.0 / 55296, 175.0 / 512, 575.0 / 13824, 44275.0 / 110592, 253.0 / 4096}};
    static const double b5[6] = {37.0 / 378, 0, 250.0 / 621, 125.0 / 594, 0, 512.0 / 1771};
    static const double b4[6] = {2825.0 / 27648, 0, 18575.0 / 48384, 13525.0 / 55296, 277.0 / 14336, 0.25};
    long nt = (long)ceil(end_time / save_step);
    double t = 0.0, dt = save_step;
    long steps = 0;
    cplx *rho = malloc(sizeof(cplx) * DD);
    for (long kt = 0; kt < nt; ++kt) {
        double tk = kt * save_step;
        /* advance to tk */
        while (t < tk - 1e-14 * (1.0 + fabs(tk))) {
            double h = dt;
            int last = 0;
            if (t + h >= tk) { h = tk - t; last = 1; }
            for (;;) {
                rhs(D, Heff, Ueff, F(ctx, t), C, k[0], tmp);
                for (int s = 1; s < 6; ++s) {
                    for (int i = 0; i < D; ++i) {
                        cplx v = C[i];
                        for (int r = 0; r < s; ++r) v += h * a[s][r] * k[r][i];
                        y[i] = v;
                    }
                    rhs(D, Heff, Ueff, F(ctx, t + c[s] * h), y, k[s], tmp);
                }
                double emax = 0.0;
                for (int i = 0; i < D; ++i) {
                    cplx v = C[i], e = 0;
                    for (int s = 0; s < 6; ++s) { v += h * b5[s] * k[s][i]; e += h * (b5[s] - b4[s]) * k[s][i]; }
                    Cn[i] = v;
                    err[i] = e;
                    double r = cabs(e) / (eps_abs + eps_rel * (cabs(C[i]) + fabs(h) * cabs(k[0][i])));
                    if (r > emax) emax = r;
                }
                ++steps;
                if (emax > 1.0) {
                    double f = 0.9 * pow(emax, -1.0 / 3.0);
                    h *= (f > 0.2 ? f : 0.2);
                    last = 0;
                    continue;
                }
                memcpy(C, Cn, sizeof(cplx) * D);
                t += h;
                if (last) t = tk;
                if (emax < 0.5) {
                    double e = emax > pow(5.0, -5.0) ? emax : pow(5.0, -5.0);
                    h *= 0.9 * pow(e, -1.0 / 5.0);
                }
                if (!last || h > dt) dt = h;
                break;
            }
        }
        /* observe */
        double nrm = 0;
        for (int i = 0; i < D; ++i) {
            cplx s = 0;
            for (int j = 0; j < D; ++j) s += M[(size_t)i * D + j] * C[j];
            nrm += creal(conj(C[i]) * s);
        }
        if (out_norm) out_norm[kt] = nrm;
        for (int l = 0; l < nrho; ++l) {
            for (size_t q = 0; q < DD; ++q) rho[q] = rhoin[2 * (l * DD + q)] + I * rhoin[2 * (l * DD + q) + 1];
            cplx s = 0;
            for (int i = 0; i < D; ++i) {
                cplx r = 0;
                for (int j = 0; j < D; ++j) r += rho[(size_t)i * D + j] * C[j];
                s += conj(C[i]) * r;
            }
            out_rho[2 * (kt * nrho + l)] = creal(s);
            out_rho[2 * (kt * nrho + l) + 1] = cimag(s);
        }
    }
    if (n_steps) *n_steps = steps;
    free(M); free(Mi); free(Heff); free(Ueff); free(A); free(C); free(y); free(tmp); free(Cn); free(err); free(rho);
    for (int s = 0; s < 6; ++s) free(k[s]);
    return 0;
}
