/*
 * laughlinedge.h -- C ABI of the MI355X edge-state Monte-Carlo engine.
 *
 * Plain C types only (no HIP or torch types).  Errors are returned as codes; the
 * message of the last failure on the calling thread is le_last_error().  The engine
 * owns all device memory; callers own every host buffer they pass in.  One engine
 * per host thread; calls on one engine are serialised on its own HIP stream.
 *
 * Which reference interface each entry point replaces (Nardin94/LaughlinEdge):
 *   le_engine_create / le_engine_destroy
 *       the per-call setup of edgeMC::spectrum_compute: RNG states, partitions,
 *       capacity checks and result buffers (modules/edge_mc/em_spectrum.cu:210-252),
 *       with the sys_params constants (modules/sys_params.h:4-25) passed at run time
 *       and monteCarloParameters (modules/edge_montecarlo.h:45-88) as le_mc_params.
 *   le_basis_state, le_basis_describe
 *       iparts::integer_partitions host partitions (modules/integer_partitions.h:29-70,
 *       modules/partitions/integer_partitions.cu:10-226).
 *   le_tune_step, le_step_search
 *       edgeMC::optimal_step / mcmc_acceptanceBisection (modules/edge_mc/em_bisection.cu:7-153).
 *   le_burnin, le_sample, le_run
 *       the mcmc_hamiltonian launch and its host averaging
 *       (modules/edge_mc/em_spectrum.cu:139-208, 256-359), for all sectors at once.
 *   le_set_mixture, le_get_mixture
 *       no reference counterpart: the importance density of the mixture estimator.
 *   le_get_moments
 *       the D2H copy + block means of fHg/fg/gg (modules/edge_mc/em_spectrum.cu:297-335).
 *   le_geneig, le_engine_spectrum
 *       ges::generalizedEigenproblem::extract_spectrum
 *       (modules/eigensolver/generalized_eigenproblem.cpp:45-52), on the GPU in FP64.
 *   le_accumulate_positions, le_get_positions, le_set_positions
 *       no reference counterpart: fixed-sample accumulation and walker state I/O, so the
 *       same configurations can be fed to the FP64 CPU oracle.
 *   le_engine_create_occupation, le_get_occupation
 *       greenFunctionMC::occupation_numbers_compute (modules/green_function/
 *       gf_occupation_numbers.cu:309-562), all pairs from one two-sided chain.
 *   le_removal_weights
 *       the spectral function the reference plans but does not implement (README.md:14-15,
 *       modules/green_function.h:18,90).
 *   le_evolve
 *       timeEvolution::edgeDensityResponse_compute (modules/dynamics/dynamics.cpp:149-309).
 * Each declaration below repeats its reference interface; entry points with no reference
 * counterpart say so.
 */
#ifndef LAUGHLINEDGE_H
#define LAUGHLINEDGE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LE_OK 0
#define LE_ERR_ARG 1        /* invalid argument */
#define LE_ERR_CAPACITY 2   /* basis beyond the engine's capacity (reference: exit(1), em_spectrum.cu:235-252) */
#define LE_ERR_HIP 3        /* HIP runtime error */
#define LE_ERR_NOT_PD 4     /* metric not positive definite in the generalized eigensolve */
#define LE_ERR_BISECTION 5  /* step tuning failed to bracket (reference: exit(1), em_bisection.cu:146-149) */
#define LE_ERR_STATE 6      /* call out of order */
#define LE_ERR_NOCONV 7     /* eigensolver did not converge */

#define LE_EST_PLAIN 0      /* estimator: sample |Psi_L|^2 once, S and H for all rows (Gram on MFMA) */
#define LE_EST_ROWWISE 1    /* estimator: the reference's row-wise |p_f Psi_L|^2 chains */
#define LE_EST_MIXTURE 2    /* estimator: one population samples w |Psi_L|^2 (le_set_mixture),
                               S and H reweighted by 1 / w (Gram on MFMA) */

#define LE_PREC_FP32 0      /* FP32 sampler, FP32 MFMA tiles flushed to FP64 */
#define LE_PREC_FP64 1      /* FP64 sampler, FP64 MFMA accumulation (the plain FP64 form) */
#define LE_PREC_BF16X3 2    /* FP32 sampler; S/H tiles from BF16 hi/lo split products
                               (hi.hi + hi.lo + lo.hi) on MFMA, FP32 tiles flushed to FP64 */
#define LE_PREC_BF16X6 3    /* as LE_PREC_BF16X3 with a hi/mid/lo split: the six products
                               t_a.t_b with a + b < 3, dropping only terms of order 2^-24 */

/* sys_params.h, at run time.  V(r) = 0.5 w r^2 + eps r^4 + strength (r / RCl)^exponent,
 * RCl = sqrt(2 m N) (system_parameters.cpp:25-32 plus the w, eps form). */
typedef struct le_system_params {
    int particles_number;      /* N  (sys_params::particlesNumber) */
    int inverse_filling;       /* m  (sys_params::inverseFilling) */
    double conf_w;
    double conf_eps;
    double conf_strength;      /* sys_params::confinementStrength */
    int conf_exponent;         /* sys_params::confinementExponent */
} le_system_params;

/* edgeMC::monteCarloParameters, at run time (em_parameters.cu:9-31 defaults). */
typedef struct le_mc_params {
    int walkers;               /* threadsPerBlock * nBlocks in the reference */
    int burnin_sweeps;         /* burnInMoves */
    int samples;               /* MCSamples: kept samples per walker */
    int thin;                  /* sweeps per kept sample (reference: 1) */
    double step;               /* randomStep */
    int tune_sweeps;           /* targetAcceptanceMoves */
    double target_acceptance;  /* targetAcceptance */
    double acceptance_tol;     /* acceptance_tol */
    uint64_t seed;
    int precision;             /* LE_PREC_* */
    int lanes_per_walker;      /* 0: smallest power of two >= N/16 (N/32 for FP64); -1: timed autotune; else 2..64 */
    int samples_per_launch;    /* kept samples per sweep+feature+Gram round (0: auto) */
    int n_blocks;              /* equal time blocks for error bars (>= 1) */
    int trace_walkers;         /* walkers whose p_1..p_4, V series are recorded */
    int graphs;                /* 1 (default): capture each full sampling round (sweeps + features)
                                  in a hipGraph and replay it (rounds with user weights, the
                                  occupation chain and partial rounds run eagerly); 0: eager, with
                                  per-kernel times in le_run_stats */
    int symmetry_moves;        /* 1: after each kept sample's sweeps, redraw the centre of mass from
                                  its exact Gaussian and rotate by a uniform angle; 2: also redraw
                                  the scale of the coordinates relative to the centre of mass from
                                  its exact Gamma law (all exact moves of |Psi_L|^2, edge engine
                                  only); 0 (default): single-particle moves only, as the reference.
                                  With LE_EST_MIXTURE, 1 is a centre-of-mass random walk (variance
                                  4 / N per component) accepted with the |Psi_L|^2 and w ratios,
                                  then a uniform rotation (2 is refused) */
    int sweep_registers;       /* 1: keep each lane's particles in registers instead of LDS (edge
                                  engine: lanes_per_walker 8 or 16 with ceil(N / lanes) in [2, 25],
                                  13 or 25 for the mixture estimator and the learned proposal's
                                  register kernel; occupation engine: ceil((N - 1) / lanes)
                                  13 or 25);
                                  0 (default): LDS.  lanes_per_walker = -1 times both kinds and
                                  keeps the faster */
    int estimator;             /* LE_EST_PLAIN (default): one population of walkers samples
                                  |Psi_L|^2 and S, H are accumulated for all rows at once;
                                  LE_EST_ROWWISE: the reference's estimator (em_spectrum.cu:16-137,
                                  318-390): for every basis state f of every sector, `walkers`
                                  chains sample |p_f Psi_L|^2 and accumulate g/f, V g/f and
                                  |g/f|^2 over f's sector; le_get_moments then returns the
                                  normalised, symmetrised M(f, g) = <g/f>_f / sqrt<|g/f|^2>_f and
                                  H(f, g) = <V g/f>_f / sqrt<|g/f|^2>_f (unit diagonal) as S and H.
                                  Row-wise engines run the LDS sweep (no registers, symmetry moves,
                                  learned proposal, extra observables or fixed samples); all rows
                                  share one step; chain c = f walkers + w uses Philox walker id
                                  c + walkers;
                                  LE_EST_MIXTURE: one population of walkers samples w |Psi_L|^2
                                  with w = rho + sum_{k=1..L} c_k (|p_k|^2 / n_k)^(L/k), L the
                                  largest sector (le_set_mixture); S and H are the sums of
                                  conj(p_a) p_b / w and conj(p_a) V p_b / w over the sums of 1 / w.
                                  By weighted AM-GM every |p_lambda|^2 of every sector is at most
                                  a constant times w, so the reweighted integrands are bounded
                                  (the plain estimator's |p_1|^(2 dL) tails are not).  S and H
                                  only (no harmonics, user weights, dilation moves or learned
                                  proposal); register or LDS sweeps */
    int gram_xcd_spread;       /* XCDs that share one row split's S/H Gram tiles: 1, 2, 4 or 8
                                  (8: consecutive workgroups take consecutive tiles, spreading every
                                  split over all XCDs, whose L2s then each fetch most of every
                                  sample row; fewer XCDs per split fetch each row fewer times but
                                  load each L2 more); 0 (default): time 1, 2, 4 and 8 on the first
                                  full-size launch and keep the fastest (cached per sector layout) */
    double mixture_rho;        /* LE_EST_MIXTURE: weight c0 of the constant term of w (0 < rho < 1,
                                  default 0.5; see le_set_mixture) */
    int fused_step;            /* 1: le_sample runs one fused launch per kept sample (no reference
                                  counterpart; em_spectrum.cu:139-208 also fuses its sampler in one
                                  kernel): the sweeps, power sums and monomials of the sample and
                                  the S/H Gram tiles of the previous one as work items of one
                                  launch, so that VALU-bound sweeps and MFMA-bound tiles share the
                                  CUs.  Same chains and samples as the separate kernels.  Needs the
                                  plain estimator, FP32, particles in registers at 8 lanes per
                                  walker (ceil(N / 8) = 13 or 25), K <= 16, S and H only (no
                                  harmonics or user weights), no symmetry moves and the plain
                                  proposal; le_sample fails with LE_ERR_ARG otherwise.
                                  0 (default): separate kernels on two streams */
    int fused_rows;            /* rows per split of the fused Gram items (a multiple of 32; 0: 2048) */
    int gram_tile_cols;        /* Gram tile width: 64 (64x64 tiles, 512-thread workgroups) or 32 (64x32
                                  tiles, 256-thread workgroups; FP32 and FP64 accumulation only);
                                  0 (default): time both with each XCD placement on the first
                                  full-size launch and keep the fastest (cached per sector layout) */
} le_mc_params;

typedef struct le_engine le_engine;

/* Extra accumulated observables (all optional; pass NULL for S and H only).
 *  harmonics: angular harmonics L; for each L the weights sum_i cos(L theta_i) and
 *      sum_i sin(L theta_i) are accumulated between every sector pair (s1, s2) with
 *      |dL_s2 - dL_s1| = L (reference: em_dsf.cu:113-134 for the DSF, em_edgeResponse.tpp
 *      :133-163 and :456-470 for the density Fourier components and selection rule).
 *  n_user_weights: per-sample weights supplied by le_set_weight_callback, accumulated
 *      between every sector pair (reference: the excitation potential U,
 *      em_edgeResponse.tpp:129-164). */
typedef struct le_observables {
    int n_harmonics;           /* <= 8 */
    const int* harmonics;
    int n_user_weights;        /* <= 8 */
} le_observables;

/* Weight identifiers for le_get_block */
#define LE_W_ONE 0                      /* metric S */
#define LE_W_V 1                        /* confinement H */
#define LE_W_COS(L) (0x100 | (L))       /* sum_i cos(L theta_i) */
#define LE_W_SIN(L) (0x200 | (L))       /* sum_i sin(L theta_i) */
#define LE_W_USER(k) (0x10000 + (k))    /* k-th user weight */

/* User weight callback, called on the engine's stream after each kept-sample sweep:
 * positions: device, walkers x N complex (float when positions_are_double == 0, else
 * double); weights: device, n_user_weights rows of `walkers` values (float for an FP32
 * engine, double for FP64) with row stride ld; stream: the engine's hipStream_t.
 * Return 0 on success. */
typedef int (*le_weight_fn)(void* ctx, const void* positions, int positions_are_double, int walkers,
                            int N, void* weights, long ld, int weights_are_double, void* stream);

typedef struct le_engine_info {
    int D;                     /* basis size over all sectors */
    int K;                     /* largest power sum degree */
    int n_sectors;
    int longest_partition;     /* longest compressed partition */
    long block_elems;          /* sum_s d_s^2 */
    int lanes_per_walker;      /* mapping in use */
    int samples_per_launch;
    int n_tiles;               /* 64x64 Gram tiles per weight */
    long x_ld;                 /* real columns per sample row */
    int sweep_in_registers;    /* 1 when the plain sweep keeps particles in registers */
    double gram_mfma_flop_per_row; /* FP32-equivalent MFMA FLOPs the Gram launches issue per
                                      sample row: 2 x 16 x 16 per computed 16x16 block (those
                                      holding requested entries, an odd one padded) and weight */
    int gram_xcd_spread;       /* XCDs sharing one row split's Gram tiles in use (the autotuned
                                  choice once the first full-size round has run) */
    long chains;               /* Markov chains held: walkers, or D x walkers for the row-wise
                                  estimator (the leading dimension of le_get/set_positions) */
    int gram_tile_cols;        /* Gram tile width in use (64 or 32; the autotuned choice once the
                                  first full-size round has run); n_tiles and
                                  gram_mfma_flop_per_row describe that width's tile list */
} le_engine_info;

typedef struct le_run_stats {
    double step;               /* proposal width in use */
    double acceptance;         /* accepted / proposed over the sampling sweeps */
    long long sweeps;          /* sweeps per walker so far (burn-in included) */
    long long samples;         /* kept samples accumulated (all walkers) */
    double ms_sweep;           /* device time, sampling sweeps */
    double ms_features;        /* device time, features */
    double ms_gram;            /* device time, Gram accumulation */
    double ms_sample_total;    /* wall time of le_sample calls */
    long long gram_launches;
    long long sweep_launches;
    double ms_sampler_graph;   /* device time of replayed sampling graphs (sweeps + features) */
    double ms_gram_launch_min; /* shortest full-size Gram launch (samples_per_launch x walkers rows)
                                  of le_sample so far, 0 before the first; the last launch of each
                                  call runs with no sampling beside it, so this is the kernel alone */
    double ms_step;            /* device time of the fused launches (fused_step = 1), the closing
                                  Gram launch of each le_sample call included */
    long long step_launches;   /* fused launches so far */
} le_run_stats;

void le_mc_params_default(le_mc_params* p);
const char* le_last_error(void);

int le_engine_create(const le_system_params* sys, const le_mc_params* mc, const int* sectors,
                     int n_sectors, const le_observables* obs, le_engine** out);
int le_set_weight_callback(le_engine* e, le_weight_fn fn, void* ctx);
int le_engine_destroy(le_engine* e);
int le_engine_get_info(const le_engine* e, le_engine_info* info);
/* sector dims (n_sectors) and basis offsets (n_sectors + 1) */
int le_sector_dims(const le_engine* e, int* dims, int* offsets);
/* state -> dL and compressed (part, multiplicity) pairs; pairs holds up to 2*8 ints */
int le_basis_state(const le_engine* e, int state, int* dL, int* pairs, int* n_pairs);

/* Basis without an engine (no GPU call): D states, largest degree K; per state (up to
 * `capacity`): dL, up to 8 (part, multiplicity) pairs at pairs[16 t ..], pair count. */
int le_basis_describe(int N, const int* sectors, int n_sectors, int capacity, int* D, int* K,
                      int* dls, int* pairs, int* n_pairs);

int le_init_walkers(le_engine* e);
/* Tune the proposal width to target_acceptance +- acceptance_tol: le_step_search over the
 * acceptance of tune_sweeps sweeps of freshly initialised walkers per trial width
 * (edgeMC::optimal_step, modules/edge_mc/em_bisection.cu:49-153).  Leaves fresh walkers. */
int le_tune_step(le_engine* e, double* step_out);

/* The step search of edgeMC::optimal_step (em_bisection.cu:80-150) over a caller's acceptance
 * measurement (no GPU call of its own): measure(ctx, step, &acceptance) returns LE_OK or an
 * error code that the search returns unchanged.  From step0: keep it when |a - target| < tol;
 * else double (a > target) or halve the step until the target is bracketed (at most 41
 * doublings or halvings, else LE_ERR_BISECTION); then interpolate linearly between the bracket ends,
 * replacing the end on the guess's side (LE_ERR_BISECTION when a guess falls outside the
 * bracket, where the reference exits, or after 60 guesses).  *n_measured (optional) counts
 * the measurements. */
typedef int (*le_accept_fn)(void* ctx, double step, double* acceptance);
int le_step_search(le_accept_fn measure, void* ctx, double step0, double target, double tol, double* step_out,
                   int* n_measured);
/* Burn-in.  A mixture engine (LE_EST_MIXTURE) whose mixture is not set yet runs the first
 * half of these sweeps on |Psi_L|^2, then sets it from the walkers (n_k = mean |p_k|^2 over
 * the walkers, le_set_mixture with mixture_rho) and runs the second half on w |Psi_L|^2. */
int le_burnin(le_engine* e, int sweeps);
/* Mixture density of a LE_EST_MIXTURE engine (no reference counterpart):
 *   w = rho + sum_{k=1..L} exp(a_k ln|p_k|^2 + b_k),  a_k = L / k,
 *   b_k = ln((1 - rho) / L) - lnGamma(a_k + 1) - a_k ln n_k,
 * with L the largest sector dL and p_k the power sums of z / sqrt(2 m N); each term has mean
 * (1 - rho) / L when |p_k|^2 / n_k is a unit exponential.  nk[0..L-1] > 0, 0 < rho < 1.
 * Sets the density for every later sweep (le_sample calls it from the walkers when it was not
 * set); the walkers are kept. */
int le_set_mixture(le_engine* e, const double* nk, int n_k, double rho);
/* The density in use: L (0 when not set), c0 = rho, a[0..L-1], b[0..L-1] (buffers of >= L). */
int le_get_mixture(const le_engine* e, int* L, double* c0, double* a, double* b);
/* sample `samples` kept samples per walker into the current block; closes the block */
int le_sample(le_engine* e, int samples);
/* init + tune (if tune_sweeps > 0) + burn-in + n_blocks x (samples / n_blocks) */
int le_run(le_engine* e);
int le_get_stats(const le_engine* e, le_run_stats* st);
int le_n_blocks_done(const le_engine* e);
/* Mean over closed blocks (block = -1) or one block's means.  S, H: sum_s d_s^2 complex
 * values, interleaved re/im, per sector d x d row-major, S_ab = <conj(p_a) p_b>, power
 * sums of z / sqrt(2 m N); H carries sum_i V(r_i). */
int le_get_moments(const le_engine* e, int block, double* S, double* H);
/* Mean (block = -1) or one block of a weight's complex block between sectors s1 and s2:
 * d_s1 x d_s2 complex, interleaved, out_ab = <conj(p_a) w p_b>.  s1 > s2 returns the
 * conjugate transpose of (s2, s1).  Blocks that were not accumulated give LE_ERR_ARG. */
int le_get_block(const le_engine* e, int block, int weight, int s1, int s2, double* out);
/* acceptance trace: per kept sample, per traced walker: Re/Im p_1..p_4, V (float) */
int le_get_trace(const le_engine* e, float* out, long capacity, long* n_rows);

int le_get_positions(le_engine* e, double* out);        /* walkers x N x 2 */
int le_set_positions(le_engine* e, const double* in);
/* fixed samples: n configurations (n x N x 2, FP64) accumulated into a fresh block */
int le_accumulate_positions(le_engine* e, const double* pos, long n);
int le_reset_moments(le_engine* e);
/* Sweeps per kept sample from the next le_sample on (>= 1).  No reference counterpart: the
 * reference keeps one sample per sweep (em_spectrum.cu:49-115); thinning trades sample
 * rate for less correlated samples. */
int le_set_thin(le_engine* e, int thin);

/* FP64 generalized eigensolve on the GPU: d x d complex (interleaved) S, H.
 * evals ascending; evecs (optional, d x d complex, column j = eigenvector j, S-normalised). */
int le_geneig(int d, const double* S, const double* H, double* evals, double* evecs);
/* le_geneig without device allocations.  No reference counterpart: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  S_dev, H_dev (d x d complex, interleaved), evals_dev (d) and the optional
 * evecs_dev (d x d complex) and status_dev (one int) are DEVICE pointers; the work is enqueued on
 * `stream` (a hipStream_t; NULL = the default stream) and the call returns without waiting.
 * When the stream reaches it: evals ascending, evecs as le_geneig's, *status_dev = 0, or 1 when
 * the metric is not positive definite (evals and evecs then undefined).  The workspace is a
 * fixed device array of the library (one per device, 8.4 MB), so d <= LE_GENEIG_NOALLOC_MAX_DIM
 * (else LE_ERR_ARG) and calls that overlap on one device (other streams or threads) race: the
 * array keeps the last call's contents.  The call records no memset, so it may be captured. */
#define LE_GENEIG_NOALLOC_MAX_DIM 256
int le_geneig_noalloc(int d, const double* S_dev, const double* H_dev, double* evals_dev, double* evecs_dev,
                      int* status_dev, void* stream);
/* per-sector eigenvalues of the engine's moments (block -1 = mean); evals has d_s entries */
int le_engine_spectrum(le_engine* e, int sector, int block, double* evals);

/* Particle-removal spectral weights (the spectral function, which the reference plans but does
 * not implement: README.md:14-15 "coming soon", the commented-out spectral_function_sampler
 * in modules/green_function.h:18,90).  The engine holds the edge space of N particles; the
 * droplet of N + 1 particles loses orbital l = m N - dL, dL = sectors[sector] <= N, and
 *   c_l Psi_L^{N+1}  is proportional to  e^(m)_dL(z_1..z_N) Psi_L^N,
 * e^(m)_k being the elementary symmetric polynomial of degree k of the multiset {z_j, each m
 * times}.  On return (block -1 = mean): energies = the sector's ascending generalized
 * eigenvalues E_j, weights[j] = |<j| c_l Psi>|^2 / |c_l Psi|^2 (they sum to 1), *orbital = l.
 * Times the occupation n_l of the (N+1)-particle Laughlin state (le_get_occupation) they give
 * the removal spectral function A(l, w) = n_l sum_j weights[j] delta(w - (E_0^{N+1} - E_j)). */
int le_removal_weights(le_engine* e, int sector, int block, int* orbital, double* energies, double* weights);

/* Time evolution (timeEvolution::edgeDensityResponse_compute, modules/dynamics/dynamics.cpp
 * :149-309) of C(0) = e_0 under M dC/dt = -i (H + F(t) U_u) C for n_u perturbations at once,
 * FP64 on the GPU.  M and H are D x D complex (only the diagonal blocks of sizes
 * block_dims[] are used: angular momentum is conserved by them), U and rho are n_u and
 * n_rho D x D complex matrices (the Hermitian part of each U_u is used; the reference's U is
 * written self-adjoint).  Observed at t_k = k save_step, k < ceil(end_time /
 * save_step): out_rho[u][k][l] = C^H rho_l C (complex), out_norm[u][k] = C^H M C.
 * Adaptive Cash-Karp 5(4) with the given tolerances (the reference: 1e-10, 1e-8). */
typedef double (*le_time_fn)(void* ctx, double t);
int le_evolve(int D, int n_blocks, const int* block_dims, const double* M, const double* H, int n_u, const double* U,
              int n_rho, const double* rho, le_time_fn F, void* ctx, double end_time, double save_step,
              double eps_abs, double eps_rel, double* out_rho, double* out_norm, long* n_steps);

/* Occupation numbers <L| a^dagger_l a_l |R> for all pairs L, R of sector dL and l < lcut
 * (greenFunctionMC::occupation_numbers_compute, gf_occupation_numbers.cu:309-562): one
 * two-sided chain on |Psi(z_0..z_{N-1})| |Psi(z_1..z_N)| (positions hold N + 1 particles)
 * serves every pair.  le_run / le_sample / le_accumulate_positions work as for the edge
 * engine; le_get_occupation returns the raw means O[a][b][l] (d x d x lcut complex). */
int le_engine_create_occupation(const le_system_params* sys, const le_mc_params* mc, int dL, int lcut,
                                le_engine** out);
int le_get_occupation(const le_engine* e, int block, double* out);

/* Learned proposal (no reference counterpart): a 3 -> 64 -> 64 -> 2 tanh MLP maps the
 * features (|z|/R, the distance to the nearest particle and the mean distance to the 8 nearest,
 * both capped at 8) of the particle being moved to radial/tangential Gaussian widths; moves are
 * accepted with the exact Metropolis-Hastings ratio (widths re-evaluated at the proposed
 * position), so |Psi_L|^2 stays the target.  Training maximises the expected squared jump
 * distance on configurations of the plain chain with Adam; the trained network is
 * tabulated on a table_size^3 grid that the sampler interpolates.  FP32 engines only. */
#define LE_PROPOSAL_PLAIN 0
#define LE_PROPOSAL_LEARNED 1
typedef struct le_train_params {
    int steps;                 /* Adam steps (2000) */
    double lr;                 /* 1e-3 */
    int batch;                 /* training triples per step (4096) */
    int n_configs;             /* configurations drawn from the plain chain (1e5) */
    int snapshot_sweeps;       /* plain sweeps between snapshots of the walkers */
    uint64_t seed;             /* weights and training noise (99) */
    double sigma0;             /* width scale; <= 0: the engine's current step */
    int table_size;            /* grid points per feature (12) */
} le_train_params;
typedef struct le_train_report {
    double loss_first, loss_last;      /* -<min(1, alpha) |dz|^2> */
    double accept_first, accept_last;  /* mean min(1, alpha) */
    double ms;                         /* wall time */
} le_train_report;
void le_train_params_default(le_train_params* p);
int le_train_proposal(le_engine* e, const le_train_params* tp, le_train_report* rep);
int le_set_proposal(le_engine* e, int kind);
int le_get_proposal_params(const le_engine* e, float* out);          /* 4546 floats */
int le_set_proposal_params(le_engine* e, const float* in, double sigma0, int table_size);
/* loss and parameter gradient on given triples (configs: n_cfg x N x 2 float) */
int le_proposal_loss_grad(le_engine* e, const float* configs, int n_cfg, const int* cfg, const int* part,
                          const float* noise, int B, double sigma0, double* loss, float* grad);

/* eigenvalues of every sector in one launch (block -1 = mean): evals has D entries in basis
 * order; status (optional, n_sectors): 0 ok, 1 metric not positive definite */
int le_engine_spectra(le_engine* e, int block, double* evals, int* status);

#ifdef __cplusplus
}
#endif

#endif
