// Host-side launch interface of the edge-state Monte-Carlo kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace le {

// One potential: V(r) = 0.5 w r^2 + eps r^4 + strength (r / R)^expo
struct Potential {
    double w, eps, strength;
    int expo;
    double R;
};

// Mixture importance density w(z) |Psi_L|^2 with w = c0 + sum_{k=1..kmix} exp(a_k ln|p_k|^2 + b_k),
// p_k = sum_i (z_i / R)^k.  kmix = 0: no mixture.
constexpr int kMixMax = 24;
struct MixParams {
    int kmix;
    double c0;
    double a[kMixMax];
    double b[kMixMax];
};

struct SweepArgs {
    void* pos;               // [walkers][N] float2 or double2
    unsigned* accepted;      // [walkers] accepted-move counters (added to)
    unsigned long long seed;
    unsigned long long sweep0;   // global sweep index of the first sweep (Philox counter)
    int N, m, walkers, nsweeps;
    double step;
    const unsigned long long* sweep_base;   // optional device value added to sweep0 (graph replays)
    int dilate;              // k_symmetry: also redraw the scale of the relative coordinates
    // row-wise estimator (launch_sweep_row): chain c samples |p_f Psi_L|^2 for the row state
    // f = c / row_walkers; its Philox walker id is c + walker_id0
    int walker_id0;
    int row_walkers;
    const uint32_t* row_fac;         // [D][8] packed (k | r << 8) factors of every state
    double2* row_pk;                 // [chains][8] running p_k of the row state's factors
    double R;                        // power sums of z / R
    // mixture density (launch_sweep / launch_sweep_reg with mix.kmix > 0): the chains sample
    // w |Psi_L|^2, starting from the p_k in mix_pk [walkers][kMixMax] (launch_mix_init)
    MixParams mix;
    double2* mix_pk;
};

// Row-wise accumulation (the reference's estimator): per row f (a basis state of sector s, d
// states from basis offset o), over the row_walkers chains of f, sums of g/f = p_g / p_f (1 at
// g = f), V g/f and |g/f|^2 for g in the sector, added to acc at [q][bo_s + a d + g'],
// q = Re fg, Im fg, Re fHg, Im fHg, gg (a = f - o, g' = g - o).
struct RowAccumArgs {
    const void* X;           // [chains][ldx] real-embedded monomials of the current sample
    long ldx;
    const void* V;           // [chains] confinement of the current sample
    const int* col_re;       // [D]
    const int* col_dim;      // [D]
    const int* row_sector;   // [D] sector of each state
    const int* sec_off;      // [nsec + 1] basis offsets
    const long* sec_bo;      // [nsec] block offsets sum_{q<s} d_q^2
    long block_elems;        // sum_s d_s^2
    int D, row_walkers;
    double* acc;             // [5][block_elems]
};

constexpr int kMaxHarmonics = 8;

struct FeatureArgs {
    const void* pos;         // [walkers][N] float2 or double2 (or double2 for fixed samples)
    int pos_is_double;       // fixed-sample path feeds FP64 positions to an FP32 engine
    void* X;                 // [rows][ldx] float or double: real-embedded monomials
    long ldx;
    void* W;                 // [nweights][wld] float or double: row weights
    long wld;                // stride between weight rows (= rows of X)
    long row0;               // first output row
    int N, walkers, D, K;
    const uint32_t* fac;     // [D][8] packed (k | r << 8)
    const int* col_re;       // [D] column of Re p_t (Im at col_re + col_dim)
    const int* col_dim;      // [D] d of the state's sector
    const unsigned long long* prod;  // [D] packed product plan: count (4 bits) + 8 x 7-bit power-table indices
    const int* pw_k;         // [npw] power table entry e = p_{pw_k[e]}^{pw_r[e]}
    const int* pw_r;
    int npw;                 // <= 128
    double invR;
    Potential pot;
    int nh;                  // angular harmonics L_h: weights sum_i cos(L theta_i), sum_i sin(L theta_i)
    int harm[kMaxHarmonics];
    float* trace;            // optional [trace_rows][9]: Re/Im p1..p4, V
    int trace_walkers;
    long trace_row0;
    const long long* trace_base;   // optional device value added to trace_row0 (graph replays)
    double* ps;              // [walkers][ps_ld] scratch: Re/Im p_1..p_K per sample
    int ps_ld;
    // mixture estimator (mix.kmix > 0): row weights V / w (weight row 0) and 1 / w (weight row
    // mix_row), and 1 / w added to mix_acc[row]
    MixParams mix;
    long mix_row;
    double* mix_acc;
    double2* mix_pk_out;     // optional [walkers][kMixMax]: p_k (k <= mix.kmix) of each sample, the
                             // mixture chains' starting values for their next sweep launch
};

// Two-sided chain features (occupation numbers): positions [walkers][N+1]; bulk power
// sums over z_1..z_{N-1}; left monomials (sector copy 0) use + z_0^k, right (copy 1)
// + z_N^k; weights w_l = e^{i phi} phi_l(z_0) conj(phi_l(z_N)), l < lcut, as rows
// (Re w_l, Im w_l) = (2 l, 2 l + 1) of W.
struct Features2Args {
    const void* pos;
    int pos_is_double;
    void* X;
    long ldx;
    void* W;
    long wld;
    long row0;
    int N, m, walkers, D, K, lcut;
    const uint32_t* fac;     // [D][8]
    const int* col_re;
    const int* col_dim;
    const int* side;         // [D] 0: left power sums, 1: right
    const double* ln2pifact; // [lcut] ln(2 pi l!)
    double invR;
    double* ps;              // [walkers][ps_ld] scratch: bulk power sums (2 K) and the unit phase (2)
    int ps_ld;               // >= 2 K + 2
};

// A 64x64 upper Gram tile (or a 64x32 one: column blocks 0 and 1 only) and the 16x16 blocks of
// it that hold requested entries, two per wave and both in one row block: byte a of wblk[w] =
// (row block << 2) | column block, bit 4 set when the block only pads an odd count and is not
// written back.  Waves w >= nwave take no blocks.
struct Tile {
    int i0, j0;              // column origins (multiples of 64)
    int nwave;
    uint32_t wblk[8];
};

// One Gram job: up to two row weights (nullptr = the constant 1) sharing a tile list.
// Weights and accumulators of one job in a launch that covers several jobs with one tile list.
struct GramJob {
    const void* w0;
    const void* w1;
    double* C0;
    double* C1;
};

struct GramArgs {
    const void* X;           // [rows][ldx]
    long ldx;
    const void* w0;          // [rows] or nullptr
    const void* w1;          // [rows] or nullptr (unused when nw == 1)
    const GramJob* jobs;     // nullptr: w0, w1, C0, C1 here; else job blockIdx.y of njobs
    int njobs;
    int nw;
    long rows;
    long rows_per_split;
    int splits;              // ceil(rows / rows_per_split)
    const Tile* tiles;
    int ntiles;
    double* C0;              // [ldc][ldc] FP64 accumulators (upper tiles)
    double* C1;
    long ldc;
    int xcd_spread;          // 1, 2 or 4: the tiles of one row split run on that many XCDs;
                             // otherwise consecutive workgroups take consecutive tiles (all 8)
    int tile_cols;           // 64 (or 0): 64x64 tiles; 32: 64x32 tiles (FP32 and FP64 only)
};

// Workgroups of a Gram launch: gram_grid(ntiles, splits, xcd_spread)
long gram_grid(int ntiles, int splits, int xcd_spread);

// One fused sampling step (launch_step): the sweeps, power sums and monomials of every walker
// (sw, f: the current sample's rows) and the S/H Gram tiles of the previous sample's rows (g,
// two weights, g.jobs unused; g.rows = 0 for none), as work items of one launch.  Queue q
// (q < 8) serves the workgroups b with b mod 8 = q: sweep units u = q + 8 j (8 wavefronts,
// 8 x 64 / G walkers) and Gram items (tile, split) with split = q + 8 j'.  queue[q] counts the
// items taken, queue[8 + q] the workgroups that finished; both are zero at launch, and the last
// workgroup of a queue to finish clears them.
constexpr int kStepQueueWords = 16;
struct StepArgs {
    SweepArgs sw;
    FeatureArgs f;
    GramArgs g;
    int units;
    unsigned* queue;         // [kStepQueueWords]
};

// Many weights on one small block: C_q[r0 + i][c0 + j] += sum_s w_q(s) X[s][r0 + i] X[s][c0 + j]
// for i < nr, j < nc (both <= 64), q < nweights, with weight q's samples at W + q wld and its
// accumulator jobs[q / 2].C0 (q even) or .C1 (q odd).
struct WeightGramArgs {
    const void* X;           // [rows][ldx]
    long ldx;
    const void* W;
    long wld;
    int nweights;
    const GramJob* jobs;
    int r0, nr, c0, nc;
    long rows;
    long rows_per_split;     // a multiple of 32
    int splits;
    long ldc;
};

// Complex block (s1, s2) of one weight's accumulator:
// out_ab = (C[p_a][p_b] + C[q_a][q_b]) + i (C[p_a][q_b] - C[q_a][p_b]), scaled.
struct ExtractItem {
    const double* C;
    int c1, d1;              // first Re column and size of the row sector
    int c2, d2;              // same for the column sector
    long out_off;            // complex-element offset of the block in the output
};

struct ExtractArgs {
    const ExtractItem* items;
    int nitems;
    long total;              // complex elements over all items
    long ldc;
    double scale;
    double* out;             // interleaved complex
};

struct GenEigProblem {
    int d;
    long in_off;             // complex-element offset of S, H, evecs blocks
    long eval_off;
    long work_off;           // complex-element offset into the work buffer (8 d^2 + 4 d)
};

// precision selector for templated paths
enum Prec { kFP32 = 0, kFP64 = 1, kBF16X3 = 2, kBF16X6 = 3 };   // kBF16X*: Gram only (FP32 data)

hipError_t launch_init_positions(void* pos, int prec, int N, int walkers, double R,
                                 unsigned long long seed, hipStream_t s, int walker_id0 = 0);
// row-wise estimator: Metropolis on |p_f Psi_L|^2 (LDS sweep with the importance ratio), the
// chains' starting p_k from their positions, and the per-sample row accumulation
hipError_t launch_sweep_row(const SweepArgs& a, int prec, int lanes_per_walker, hipStream_t s);
hipError_t launch_row_init(const SweepArgs& a, int prec, hipStream_t s);
hipError_t launch_row_accum(const RowAccumArgs& a, int prec, hipStream_t s);
hipError_t launch_sweep(const SweepArgs& a, int prec, int lanes_per_walker, hipStream_t s);
// mixture chains: a.mix_pk[w][k - 1] = p_k (k <= a.mix.kmix) of walker w's positions, in FP64
hipError_t launch_mix_init(const SweepArgs& a, int prec, hipStream_t s);
// exact centre-of-mass heat bath + global rotation of every walker (a.sweep0 + *a.sweep_base
// selects the random numbers)
hipError_t launch_symmetry(const SweepArgs& a, int prec, hipStream_t s);
// mixture chains (a.mix): centre-of-mass random walk accepted with the |Psi_L|^2 and w ratios,
// then a uniform rotation
hipError_t launch_symmetry_mix(const SweepArgs& a, int prec, hipStream_t s);
// two-sided chain: a.N is the number of particles per state (positions hold N + 1)
hipError_t launch_sweep2(const SweepArgs& a, int prec, int lanes_per_walker, hipStream_t s);
// two-sided chain with the N - 1 bulk particles in registers: 8 or 16 lanes, ceil((N - 1) / lanes)
// = 13 or 25 slots; bit-identical to launch_sweep2 with the same lanes
bool sweep2_reg_supported(int lanes_per_walker, int N);
hipError_t launch_sweep2_reg(const SweepArgs& a, int prec, int lanes_per_walker, hipStream_t s);
hipError_t launch_features2(const Features2Args& a, int prec, hipStream_t s);
hipError_t launch_features(const FeatureArgs& a, int prec, hipStream_t s);
hipError_t launch_gram(const GramArgs& a, int prec, int splits, hipStream_t s);
// fused steps: FP32 plain chains with particles in registers at 8 lanes per walker
// (ceil(N / 8) = 13 or 25), power sums up to K <= 16, no harmonics, trace allowed
bool step_supported(int lanes_per_walker, int N, int K);
hipError_t launch_step(const StepArgs& a, int lanes_per_walker, hipStream_t s);
hipError_t launch_gram_weights(const WeightGramArgs& a, int prec, hipStream_t s);
hipError_t launch_extract(const ExtractArgs& a, hipStream_t s);
hipError_t launch_geneig(const double* S, const double* H, double* work, const GenEigProblem* probs,
                         int nprob, int dmax, double* evals, double* evecs, int* status, hipStream_t s);
hipError_t launch_zero(double* p, long n, hipStream_t s);
// bases[0] = sweep base, bases[1] = trace row base (read by replayed sample graphs)
hipError_t launch_set_bases(unsigned long long* bases, unsigned long long sweep_base, long long trace_base,
                            hipStream_t s);

size_t sweep_lds_bytes(int prec, int lanes_per_walker, int N);
// register-resident sweeps (k_sweep_reg): the (G, ceil(N / G)) pairs that are instantiated
// register sweeps: G in {8, 16} with ceil(N / G) in {13, 25}, or (not for the mixture chains)
// any ceil(N / G) in [2, 25]
bool sweep_reg_supported(int lanes_per_walker, int N, bool mixture = false);
hipError_t launch_sweep_reg(const SweepArgs& a, int prec, int lanes_per_walker, hipStream_t s);
size_t sweep2_lds_bytes(int prec, int lanes_per_walker, int N);
bool sweep_supported(int lanes_per_walker);

}  // namespace le
