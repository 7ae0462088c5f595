// C-ABI engine: owns device memory, drives the sampler / features / Gram / eigensolver
// kernels on one HIP stream, and keeps per-block FP64 moments for error bars.
//
// Reference drivers this replaces (Nardin94/LaughlinEdge):
//   edgeMC::spectrum_compute host flow .... modules/edge_mc/em_spectrum.cu:210-422
//   edgeMC::optimal_step (step bisection) . modules/edge_mc/em_bisection.cu:49-153
//   host block averaging .................. modules/edge_mc/em_spectrum.cu:318-335
// The reference runs one chain per basis row f on |p_f Psi_L|^2; this engine samples
// |Psi_L|^2 once for all rows and sectors and accumulates S = <conj(p_a) p_b> and
// H = <conj(p_a) V p_b> as weighted Gram matrices of the sample matrix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/laughlinedge.h"
#include "le_basis.hpp"
#include "le_evolve.hpp"
#include "le_kernels.hpp"
#include "le_learned.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                      \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess)                                                             \
            return fail(LE_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));  \
    } while (0)

template <typename T>
hipError_t dalloc(T** p, size_t count)
{
    *p = nullptr;
    if (count == 0) return hipSuccess;
    return hipMalloc((void**)p, count * sizeof(T));
}

// Gram tiles of 64 rows x TC columns (TC = 64, 32) over the sector pairs `blocks`, sector s in
// sample-row columns [colb[s], colb[s] + 2 d_s).  Per tile, the 16x16 blocks that hold an entry
// the extraction reads (one with row <= column inside a requested block or its transpose), two
// to a wave within a row block so that they share the row operand; an odd block of a row block
// is computed again and not written.  The Gram kernels compute only these (N = 200, dL 1..16:
// 0.76 of the products of whole tiles).  flop_row gains the MFMA FLOP per sample row of nw
// weights.
std::vector<le::Tile> gram_tiles(const le::Basis& basis, const std::vector<int>& colb,
                                 const std::vector<std::pair<int, int>>& blocks, int TC, int nw, double& flop_row)
{
    auto meets = [](int r0, int r1, int q0, int q1, int rb0, int rb1, int qb0, int qb1) {
        const int ra = std::max(r0, rb0), rb = std::min(r1, rb1);
        const int qa = std::max(q0, qb0), qb = std::min(q1, qb1);
        return ra <= rb && qa <= qb && ra <= qb;
    };
    std::set<std::pair<int, int>> tset;
    for (auto& pr : blocks) {
        const int c1 = colb[pr.first], e1 = c1 + 2 * basis.dim(pr.first) - 1;
        const int c2 = colb[pr.second], e2 = c2 + 2 * basis.dim(pr.second) - 1;
        for (int a = c1 / 64; a <= e1 / 64; ++a)
            for (int b = c2 / TC; b <= e2 / TC; ++b)
                if (64 * a <= TC * b + TC - 1) tset.insert({a, b});
    }
    std::vector<le::Tile> th;
    for (auto& pr : tset) {
        le::Tile t{};
        t.i0 = pr.first * 64;
        t.j0 = pr.second * TC;
        std::vector<uint32_t> act;
        for (int rb = 0; rb < 4; ++rb) {
            const size_t first = act.size();
            for (int cb = 0; cb < TC / 16; ++cb) {
                const int r0 = t.i0 + 16 * rb, q0 = t.j0 + 16 * cb;
                for (auto& bp : blocks) {
                    const int c1 = colb[bp.first], e1 = c1 + 2 * basis.dim(bp.first) - 1;
                    const int c2 = colb[bp.second], e2 = c2 + 2 * basis.dim(bp.second) - 1;
                    if (meets(r0, r0 + 15, q0, q0 + 15, c1, e1, c2, e2) ||
                        meets(r0, r0 + 15, q0, q0 + 15, c2, e2, c1, e1)) {
                        act.push_back((uint32_t)(rb << 2 | cb));
                        break;
                    }
                }
            }
            if ((act.size() - first) & 1) act.push_back(act.back() | 16u);   // computed again, not written
        }
        if (act.empty()) continue;
        t.nwave = (int)act.size() / 2;
        flop_row += 2.0 * 16 * 16 * (double)act.size() * nw;
        for (int w = 0; w < t.nwave; ++w) t.wblk[w] = act[2 * w] | act[2 * w + 1] << 8;
        th.push_back(t);
    }
    return th;
}

// Default walker mapping: the smallest power of two >= N / 16 lanes per walker (about 13-25
// pair slots per lane), in [2, 64].  Deterministic, so a seed fixes the chain.
// Lanes per walker without timing: the smallest power of two with at most 16 particle
// slots per lane (FP32), 32 for FP64, whose sweeps are the slower ones and prefer longer
// per-lane loops (N = 200 on MI355X: 0.52 ms FP32 at 16 lanes, 1.54 ms FP64 at 8 lanes).
constexpr int kPsLd = 50;        // Re/Im p_1..p_24 per sample (+ the two-sided chain's phase)
constexpr int kMixChunk = 8;          // mixture chains: sweeps per launch (FP32 running p_k drift)

// Sweep mapping: lanes per walker in the low byte; kRegSweep marks the register-resident
// kernel (k_sweep_reg) instead of the LDS one.
constexpr int kRegSweep = 0x100;
inline int map_lanes(int G) { return G & 0xff; }
inline bool map_reg(int G) { return (G & kRegSweep) != 0; }

int default_lanes(int N, int prec)
{
    const int per_lane = prec == LE_PREC_FP64 ? 32 : 16;
    int g = 2;
    while (g < 64 && g * per_lane < N) g <<= 1;
    return g;
}

std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int>, int> g_tuned_lanes;   // (N, walkers, prec, two-sided) -> G
// Gram tile placement chosen per (rows per launch, Gram kind, sector layout)
std::map<std::tuple<long, int, std::vector<int>, int>, std::pair<int, int>> g_tuned_spread;   // -> (spread, tile columns)

}  // namespace

namespace le {
void set_error(const std::string& msg) { g_err = msg; }
}  // namespace le

// Block records.  Records of up to 64 MB live in pinned host memory: each closed block is
// copied from the device straight into its slot (no host-side copy); slots come from chunks of
// ~64 MB allocated once and reused after le_reset_moments.  A copy into freshly allocated
// pageable memory cost ~1.5 ms per 4 MB record in page faults, inside every sampling call.
// Larger records (many weights on large bases) go through the pinned staging buffer into
// pageable vectors, as pinned chunks of their size would tie up gigabytes.
struct RecordStore {
    std::vector<double*> chunks;
    std::vector<std::vector<double>> big;   // records too large for pinned slots
    std::vector<const double*> recs;        // record b
    size_t len = 0;                         // doubles per record
    size_t per_chunk = 1;
    size_t used = 0;                        // pinned slots handed out
    bool pinned = true;
    static constexpr size_t kChunkBytes = (size_t)64 << 20;
    hipError_t init(size_t doubles)
    {
        len = std::max<size_t>(2, doubles);
        pinned = len * sizeof(double) <= kChunkBytes;
        per_chunk = std::max<size_t>(1, kChunkBytes / (len * sizeof(double)));
        return pinned ? grow() : hipSuccess;
    }
    hipError_t grow()
    {
        double* c = nullptr;
        hipError_t h = hipHostMalloc((void**)&c, sizeof(double) * len * per_chunk, hipHostMallocDefault);
        if (h == hipSuccess) chunks.push_back(c);
        return h;
    }
    // where the next record is copied from the device: its pinned slot, or the staging buffer
    hipError_t next(double* staging, double** out)
    {
        if (!pinned) {
            *out = staging;
            return hipSuccess;
        }
        if (used == chunks.size() * per_chunk) {
            hipError_t h = grow();
            if (h != hipSuccess) return h;
        }
        *out = chunks[used / per_chunk] + (used % per_chunk) * len;
        return hipSuccess;
    }
    // after the copy has completed
    void push(const double* slot)
    {
        if (pinned) {
            recs.push_back(slot);
            ++used;
        } else {
            big.emplace_back(slot, slot + len);
            recs.push_back(big.back().data());
        }
    }
    size_t size() const { return recs.size(); }
    const double* operator[](size_t b) const { return recs[b]; }
    void clear()
    {
        recs.clear();
        big.clear();
        used = 0;
    }
    ~RecordStore()
    {
        for (double* c : chunks) (void)hipHostFree(c);
    }
};

struct le_engine {
    le_system_params sys{};
    le_mc_params mc{};
    le::Basis basis;
    int prec = LE_PREC_FP32;        // data precision of positions, samples and tiles
    int gram_prec = le::kFP32;      // accumulation kernel (kBF16X3, kBF16X6: split products)
    size_t rsz = 4;
    int G = 8;
    int spl = 1;                 // kept samples per launch round
    double R = 1.0;
    double step = 0.5;
    hipStream_t stream = nullptr;

    // device buffers
    void* pos = nullptr;         // walkers x N x (float2|double2)
    unsigned* acc = nullptr;
    double* ps_d = nullptr;      // [rows][kPsLd] per-sample power sums (features scratch)
    void* X = nullptr;           // rows x ldx: Xb[0] (fixed-sample and two-sided paths)
    void* Xb[2] = {nullptr, nullptr};     // double buffer: sampler fills one, Gram reads the other
    void* Wbb[2] = {nullptr, nullptr};
    hipStream_t gstream = nullptr;        // Gram accumulation stream
    hipEvent_t filled[2] = {nullptr, nullptr}, drained[2] = {nullptr, nullptr};
    std::vector<hipEvent_t> evpool;
    size_t evused = 0;
    hipEvent_t next_event();
    uint32_t* fac = nullptr;
    int* col_re = nullptr;
    int* col_dim = nullptr;
    unsigned long long* prod_d = nullptr;   // per-state product plan over the power table
    int* pwk_d = nullptr;
    int* pwr_d = nullptr;
    int npw = 0;
    long ldx = 0;
    // first column of sector s in a sample row; sector s takes [colb[s], colb[s] + 2 d_s)
    std::vector<int> colb;
    // weights: row 0 = V, rows 1..2 nh = cos/sin harmonics, then user weights
    void* Wb = nullptr;
    long wrows = 0;
    int nh = 0;
    int harm[le::kMaxHarmonics] = {0};
    int n_user = 0;
    bool two_sided = false;          // occupation-number chain (N + 1 coordinates)
    int lcut = 0;
    double* lnfact_d = nullptr;
    int* side_d = nullptr;
    int npos() const { return sys.particles_number + (two_sided ? 1 : 0); }
    // learned proposal
    int proposal = LE_PROPOSAL_PLAIN;
    float* mlp = nullptr;            // kMlpParams
    float* table = nullptr;          // T^3 x 2
    int table_T = 0;
    float table_lo[3] = {0.f, 0.f, 0.f}, table_hi[3] = {1.5f, 4.f, 8.f};
    double sigma0 = 0.0;
    le_weight_fn wfn = nullptr;
    void* wctx = nullptr;
    struct Job {
        int wid[2];                  // weight identifiers (LE_W_*)
        int nw;
        std::vector<std::pair<int, int>> blocks;   // sector pairs s1 <= s2
        le::Tile* tiles = nullptr;
        int ntiles = 0;
        le::Tile* tiles32 = nullptr;   // the same entries on 64x32 tiles
        int ntiles32 = 0;
        double* C[2] = {nullptr, nullptr};
    };
    std::vector<Job> jobs;
    // consecutive jobs with the same sector blocks and weight count: one Gram launch each, over
    // a device table of their weights and accumulators (one table per sample buffer)
    struct Batch {
        int first, count;
    };
    std::vector<Batch> batches;
    le::GramJob* jobtab[2] = {nullptr, nullptr};
    double* cpool = nullptr;         // every job's FP64 accumulators, one allocation: one clear
    size_t cpool_elems = 0;
    bool weights_gram = false;       // occupation weights on a block of <= 64 x 64 real columns:
                                     // one k_gram_weights launch for all of them
    int ntiles = 0;                  // tiles of the S/H job
    // row-wise estimator: mc.walkers = D x row_W chains, chain c on row state c / row_W
    bool rowwise = false;
    int row_W = 0;
    double2* row_pk = nullptr;       // [chains][8] running p_k of each chain's row-state factors
    double* row_acc = nullptr;       // [5][sum_s d_s^2] sums of Re, Im g/f, Re, Im V g/f, |g/f|^2
    int* row_sector = nullptr;       // [D]
    int* sec_off_d = nullptr;        // [nsec + 1]
    long* sec_bo_d = nullptr;        // [nsec]
    int sample_rowwise(int samples);
    hipError_t init_positions();
    // mixture estimator: walkers sample w |Psi_L|^2 (le::MixParams; kmix = 0 until set)
    bool mixture = false;
    le::MixParams mixp{};
    double2* mix_pk = nullptr;       // [walkers][kMixMax] p_k at the start of each sweep launch
    double* mix_acc = nullptr;       // [wrows] per-row sums of 1 / w in the open block
    double* mix_acc_h = nullptr;     // pinned staging of mix_acc
    long mix_row = 0;                // weight row of 1 / w (row 0 holds V / w)
    int mix_L = 0;                   // largest sector dL: the terms k = 1..L
    int set_mixture(const double* nk, int n, double rho);
    int calibrate_mixture();
    // mixture weights in the features; walkers: the rows are the walkers' current samples, whose
    // p_k start the next sweep launch
    void mix_features(le::FeatureArgs& f, bool walkers) const;
    int gram_spread = 8;             // XCDs sharing one row split's Gram tiles (1, 2, 4 or 8)
    int gram_tc = 64;                // Gram tile columns in use (64, or 32 for the FP32 / FP64 tiles)
    double gram_flop_row32 = 0;      // MFMA FLOPs per sample row issued with 32-column tiles
    bool need_spread_tune = false;   // XCD placement or tile width to time on the first full-size round
    long long block_grams = 0;       // Gram launches into the open block (the tuning needs none yet)
    int tune_gram_spread(int buf);
    double gram_flop_row = 0;        // MFMA FLOPs issued per sample row over all Gram jobs
    struct Item {
        int wid, s1, s2;
        long off;                    // complex offset in a block record
    };
    std::vector<Item> items;
    long rec_elems = 0;              // complex elements per block record
    le::ExtractItem* items_d = nullptr;
    double* rec_d = nullptr;
    double* rec_h = nullptr;         // pinned staging of one block record too large for the pinned slots
    unsigned* acc_h = nullptr;       // pinned staging of the per-walker accepted-move counts
    float* trace_d = nullptr;
    long trace_cap_rows = 0;
    long trace_rows = 0;
    double* eig_work = nullptr;
    size_t eig_work_elems = 0;

    // host state: one record (all items) per closed block
    RecordStore rec;
    std::vector<long long> blk_samples;
    std::vector<double> blk_norm;    // per block: the samples, or the sum of 1 / w (mixture)
    bool walkers_ready = false;
    unsigned long long sweep_counter = 0;
    long long sweeps_done = 0;
    le_run_stats st{};
    unsigned long long acc_base = 0;   // accepted moves before the sampling sweeps
    long long acc_moves = 0;           // proposed moves during sampling sweeps
    unsigned long long acc_sampling = 0;
    std::vector<hipEvent_t> ev;
    // sampler rounds captured as graphs (mc.graphs): one per (buffer, traced) and the
    // parameters baked into their kernel arguments
    struct SampleGraph {
        hipGraphExec_t exec;
        int buf;
        bool traced;
        int thin, proposal, G, n, table_T;
        double step;
        const float* table;
    };
    std::vector<SampleGraph> graphs;
    unsigned long long* bases_d = nullptr;   // [0] sweep base, [1] trace row base
    // fused steps (mc.fused_step): the work queues of launch_step, zero between launches
    unsigned* step_q = nullptr;
    int fused_check() const;
    int sample_fused(int samples, le::FeatureArgs f, int tw);

    ~le_engine();
    int setup(const le_observables* obs);
    const void* weight_ptr(int wid, int buf = 0) const;
    long item_offset(int wid, int s1, int s2) const;
    int autotune_lanes();
    // mix_fresh: mix_pk already holds the p_k of `positions` (written by the last features launch)
    int run_sweeps(int nsweeps, unsigned long long sweep0, void* positions,
                   const unsigned long long* sweep_base = nullptr, bool mix_fresh = false);
    int sample_graph(le::FeatureArgs f, int buf, bool traced, int tw, hipGraphExec_t* out);
    int capture_round(le::FeatureArgs f, int buf, bool traced, int tw, hipGraphExec_t* out);
    int symmetry(unsigned long long sweep0, const unsigned long long* sweep_base);
    void drop_graphs();
    int close_block(long long samples);
    int gram_rows(long rows, int buf = 0, hipStream_t s = nullptr);
    double measure_acceptance(double trial_step, int attempt, int* rc);
    hipError_t features2(const void* positions, int pos_is_double, int nsamp, long row0);
};

hipError_t le_engine::features2(const void* positions, int pos_is_double, int nsamp, long row0)
{
    le::Features2Args f{};
    f.pos = positions;
    f.pos_is_double = pos_is_double;
    f.X = X;
    f.ldx = ldx;
    f.W = (char*)Wb + (size_t)(1 + 2 * nh) * wrows * rsz;   // user rows: (Re w_l, Im w_l)
    f.wld = wrows;
    f.row0 = row0;
    f.N = sys.particles_number;
    f.m = sys.inverse_filling;
    f.walkers = nsamp;
    f.D = basis.D;
    f.K = basis.K;
    f.lcut = lcut;
    f.fac = fac;
    f.col_re = col_re;
    f.col_dim = col_dim;
    f.side = side_d;
    f.ln2pifact = lnfact_d;
    f.invR = 1.0 / R;
    f.ps = ps_d;
    f.ps_ld = kPsLd;
    return le::launch_features2(f, prec, stream);
}

void le_engine::drop_graphs()
{
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
}

le_engine::~le_engine()
{
    drop_graphs();
    (void)hipFree(bases_d);
    (void)hipFree(pos);
    (void)hipFree(acc);
    (void)hipFree(ps_d);
    (void)hipFree(Xb[0]);
    (void)hipFree(Xb[1]);
    (void)hipFree(fac);
    (void)hipFree(col_re);
    (void)hipFree(col_dim);
    (void)hipFree(prod_d);
    (void)hipFree(pwk_d);
    (void)hipFree(pwr_d);
    (void)hipFree(Wbb[0]);
    (void)hipFree(Wbb[1]);
    for (auto q : evpool) (void)hipEventDestroy(q);
    for (int b = 0; b < 2; ++b) {
        if (filled[b]) (void)hipEventDestroy(filled[b]);
        if (drained[b]) (void)hipEventDestroy(drained[b]);
    }
    if (gstream) (void)hipStreamDestroy(gstream);
    for (auto& j : jobs) {
        (void)hipFree(j.tiles);
        (void)hipFree(j.tiles32);
    }
    (void)hipFree(jobtab[0]);
    (void)hipFree(jobtab[1]);
    (void)hipFree(cpool);
    (void)hipFree(items_d);
    (void)hipFree(rec_d);
    (void)hipFree(row_pk);
    (void)hipFree(row_acc);
    (void)hipFree(row_sector);
    (void)hipFree(sec_off_d);
    (void)hipFree(sec_bo_d);
    if (rec_h) (void)hipHostFree(rec_h);
    if (acc_h) (void)hipHostFree(acc_h);
    (void)hipFree(lnfact_d);
    (void)hipFree(side_d);
    (void)hipFree(mlp);
    (void)hipFree(table);
    (void)hipFree(trace_d);
    (void)hipFree(eig_work);
    (void)hipFree(mix_pk);
    (void)hipFree(mix_acc);
    (void)hipFree(step_q);
    if (mix_acc_h) (void)hipHostFree(mix_acc_h);
    for (auto e : ev) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
}

extern "C" void le_mc_params_default(le_mc_params* p)
{
    // em_parameters.cu:9-31 defaults, walkers = 32 x 128
    std::memset(p, 0, sizeof(*p));
    p->walkers = 4096;
    p->burnin_sweeps = 1000;
    p->samples = 50000;
    p->thin = 1;
    p->step = 0.8;
    p->tune_sweeps = 100;
    p->target_acceptance = 0.5;
    p->acceptance_tol = 0.01;
    p->seed = 1234;
    p->precision = LE_PREC_FP32;
    p->lanes_per_walker = 0;
    p->samples_per_launch = 0;
    p->n_blocks = 1;
    p->trace_walkers = 0;
    p->graphs = 1;
    p->symmetry_moves = 0;
    p->sweep_registers = 0;
    p->estimator = LE_EST_PLAIN;
    p->gram_xcd_spread = 0;
    p->mixture_rho = 0.5;
    p->fused_step = 0;
    p->fused_rows = 0;
    p->gram_tile_cols = 0;
}

extern "C" const char* le_last_error(void) { return g_err.c_str(); }

const void* le_engine::weight_ptr(int wid, int buf) const
{
    const size_t rsz_ = rsz;
    const void* base = Wbb[buf];
    auto row = [&](long r) { return (const void*)((const char*)base + (size_t)r * wrows * rsz_); };
    if (wid == LE_W_ONE) return mixture ? row(mix_row) : nullptr;
    if (wid == LE_W_V) return row(0);
    if (wid >= LE_W_USER(0)) {
        const int k = wid - LE_W_USER(0);
        return k < n_user ? row(1 + 2 * nh + k) : nullptr;
    }
    const int kind = wid >> 8, par = wid & 0xff;
    if (kind == 1 || kind == 2) {
        for (int h = 0; h < nh; ++h)
            if (harm[h] == par) return row(1 + 2 * h + (kind == 2 ? 1 : 0));
    }
    return nullptr;
}

long le_engine::item_offset(int wid, int s1, int s2) const
{
    for (const auto& it : items)
        if (it.wid == wid && it.s1 == s1 && it.s2 == s2) return it.off;
    return -1;
}

int le_engine::setup(const le_observables* obs)
{
    const int N = sys.particles_number;
    const int D = basis.D;
    const int ns = basis.nsec();
    R = std::sqrt(2.0 * sys.inverse_filling * N);
    step = mc.step;
    rsz = (prec == LE_PREC_FP64) ? sizeof(double) : sizeof(float);
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));

    if (obs) {
        nh = obs->n_harmonics;
        for (int h = 0; h < nh; ++h) harm[h] = obs->harmonics[h];
        n_user = obs->n_user_weights;
    }

    // Gram jobs: {1, V} on diagonal sector blocks; {cos L, sin L} on pairs |dL2 - dL1| = L;
    // user weights (two per job) on every pair
    auto add_job = [&](int w0, int w1, int nw, const std::vector<std::pair<int, int>>& blocks) {
        Job j;
        j.wid[0] = w0;
        j.wid[1] = w1;
        j.nw = nw;
        j.blocks = blocks;
        jobs.push_back(j);
    };
    std::vector<std::pair<int, int>> diag, all;
    for (int s = 0; s < ns; ++s)
        if (basis.dim(s) > 0) diag.push_back({s, s});
    for (int a = 0; a < ns; ++a)
        for (int b = a; b < ns; ++b)
            if (basis.dim(a) > 0 && basis.dim(b) > 0) all.push_back({a, b});
    if (two_sided) {
        // occupation weights (Re w_l, Im w_l) between the left (0) and right (1) copies
        std::vector<std::pair<int, int>> lr = {{0, 1}};
        for (int k = 0; k < n_user; k += 2) add_job(LE_W_USER(k), LE_W_USER(k + 1), 2, lr);
    } else {
        add_job(LE_W_ONE, LE_W_V, 2, diag);
    }
    for (int h = 0; h < nh && !two_sided; ++h) {
        std::vector<std::pair<int, int>> bl;
        for (auto& pr : all)
            if (std::abs(basis.sectors[pr.second] - basis.sectors[pr.first]) == harm[h]) bl.push_back(pr);
        if (!bl.empty()) add_job(LE_W_COS(harm[h]), LE_W_SIN(harm[h]), 2, bl);
    }
    for (int k = 0; k < n_user && !two_sided; k += 2) {
        if (k + 1 < n_user) add_job(LE_W_USER(k), LE_W_USER(k + 1), 2, all);
        else add_job(LE_W_USER(k), 0, 1, all);
    }
    // Sample-row layout: sector s takes columns [colb[s], colb[s] + 2 d_s), Re then Im.  Three
    // candidates: sectors packed back to back; every sector starting on a multiple of 16
    // columns, so that no 16x16 Gram block straddles two sectors; only sectors of 16 columns
    // or more aligned.  The one with the least Gram work is kept, counting the computed
    // 16x16 blocks of all jobs' 64-column tiles plus 1.5 per tile for its staging
    // (N = 200, dL 1..16: aligned, 1180 blocks in 97 tiles against 1258 in 103 packed; many
    // small sectors stay packed).  Padding columns stay zero.
    {
        double best = -1.0;
        for (int mode = 0; mode < 3; ++mode) {
            std::vector<int> cb(ns + 1, 0);
            for (int s = 0; s < ns; ++s) {
                const int w = 2 * basis.dim(s);
                int start = cb[s];
                if (mode == 1 || (mode == 2 && w >= 16)) start = (start + 15) / 16 * 16;
                cb[s] = start;
                cb[s + 1] = start + w;
            }
            double cost = 0.0;
            for (auto& j : jobs) {
                double fl = 0.0;
                const std::vector<le::Tile> th = gram_tiles(basis, cb, j.blocks, 64, 1, fl);
                cost += fl / (2.0 * 16 * 16) + 1.5 * (double)th.size();
            }
            if (best < 0.0 || cost < best) {
                best = cost;
                colb = cb;
            }
        }
    }
    ldx = std::max<long>(64, ((long)colb[ns] + 63) / 64 * 64);
    std::vector<int> cre(D), cdim(D);
    for (int s = 0; s < ns; ++s) {
        const int o = basis.sector_offset[s], d = basis.dim(s);
        for (int t = 0; t < d; ++t) {
            cre[o + t] = colb[s] + t;
            cdim[o + t] = d;
        }
    }
    for (auto& j : jobs) {
        std::vector<le::Tile> th = gram_tiles(basis, colb, j.blocks, 64, j.nw, gram_flop_row);
        std::vector<le::Tile> th32 = gram_tiles(basis, colb, j.blocks, 32, j.nw, gram_flop_row32);
        j.ntiles = (int)th.size();
        HIPCHK(dalloc(&j.tiles, (size_t)std::max(1, j.ntiles)));
        if (j.ntiles) HIPCHK(hipMemcpy(j.tiles, th.data(), j.ntiles * sizeof(le::Tile), hipMemcpyHostToDevice));
        j.ntiles32 = (int)th32.size();
        HIPCHK(dalloc(&j.tiles32, (size_t)std::max(1, j.ntiles32)));
        if (j.ntiles32)
            HIPCHK(hipMemcpy(j.tiles32, th32.data(), j.ntiles32 * sizeof(le::Tile), hipMemcpyHostToDevice));
        cpool_elems += (size_t)j.nw * ldx * ldx;
        for (int q = 0; q < j.nw; ++q)
            for (auto& pr : j.blocks) {
                items.push_back({j.wid[q], pr.first, pr.second, rec_elems});
                rec_elems += (long)basis.dim(pr.first) * basis.dim(pr.second);
            }
    }
    HIPCHK(dalloc(&cpool, std::max<size_t>(1, cpool_elems)));
    {
        size_t o = 0;
        for (auto& j : jobs)
            for (int q = 0; q < j.nw; ++q, o += (size_t)ldx * ldx) j.C[q] = cpool + o;
    }
    HIPCHK(le::launch_zero(cpool, (long)cpool_elems, stream));
    ntiles = jobs[0].ntiles;
    std::vector<le::ExtractItem> xi;
    for (auto& it : items) {
        const double* C = nullptr;
        for (auto& j : jobs)
            for (int q = 0; q < j.nw; ++q)
                if (j.wid[q] == it.wid) C = j.C[q];
        xi.push_back({C, colb[it.s1], basis.dim(it.s1), colb[it.s2],
                      basis.dim(it.s2), it.off});
    }
    HIPCHK(dalloc(&items_d, std::max<size_t>(1, xi.size())));
    if (!xi.empty()) HIPCHK(hipMemcpy(items_d, xi.data(), xi.size() * sizeof(le::ExtractItem), hipMemcpyHostToDevice));
    HIPCHK(dalloc(&rec_d, (size_t)std::max<long>(2, 2 * rec_elems)));
    // pinned host staging: a pageable copy of the block record went through the runtime's
    // bounce buffer at ~2 GB/s (2 ms for the 4 MB record of N = 200, dL 1..16)
    HIPCHK(hipHostMalloc((void**)&rec_h, sizeof(double) * (size_t)std::max<long>(2, 2 * rec_elems),
                         hipHostMallocDefault));
    if (!rowwise) HIPCHK(rec.init((size_t)2 * rec_elems));
    HIPCHK(hipHostMalloc((void**)&acc_h, sizeof(unsigned) * (size_t)std::max(1, mc.walkers), hipHostMallocDefault));

    // kept samples per launch round: ~128k rows, X under 1 GiB
    const long row_bytes = ldx * (long)rsz;
    long max_rows = std::max<long>(mc.walkers, (1L << 30) / row_bytes);
    if (mc.samples_per_launch > 0) {
        spl = mc.samples_per_launch;
    } else {
        long want = (131072 + mc.walkers - 1) / mc.walkers;
        spl = (int)std::max<long>(1, std::min<long>(want, max_rows / mc.walkers));
    }
    const long rows = (long)spl * mc.walkers;
    wrows = rows;

    HIPCHK(hipMalloc(&pos, (size_t)mc.walkers * npos() * 2 * rsz));
    if (two_sided) {
        std::vector<double> lf(lcut);
        for (int l = 0; l < lcut; ++l) lf[l] = std::lgamma((double)l + 1.0) + std::log(2.0 * M_PI);
        std::vector<int> side(D);
        for (int t = 0; t < D; ++t) side[t] = t >= basis.sector_offset[1] ? 1 : 0;
        HIPCHK(dalloc(&lnfact_d, (size_t)std::max(1, lcut)));
        HIPCHK(hipMemcpy(lnfact_d, lf.data(), sizeof(double) * lcut, hipMemcpyHostToDevice));
        HIPCHK(dalloc(&side_d, (size_t)std::max(1, D)));
        HIPCHK(hipMemcpy(side_d, side.data(), sizeof(int) * D, hipMemcpyHostToDevice));
    }
    HIPCHK(dalloc(&acc, (size_t)mc.walkers));
    HIPCHK(dalloc(&ps_d, (size_t)rows * kPsLd));
    if (mixture) {
        mix_row = 1 + 2 * nh + n_user;
        HIPCHK(dalloc(&mix_pk, (size_t)mc.walkers * le::kMixMax));
        HIPCHK(dalloc(&mix_acc, (size_t)rows));
        HIPCHK(le::launch_zero(mix_acc, rows, stream));
        HIPCHK(hipHostMalloc((void**)&mix_acc_h, sizeof(double) * (size_t)rows, hipHostMallocDefault));
    }
    for (int b = 0; b < 2; ++b) {
        HIPCHK(hipMalloc(&Xb[b], (size_t)rows * ldx * rsz));
        HIPCHK(hipMemsetAsync(Xb[b], 0, (size_t)rows * ldx * rsz, stream));
        HIPCHK(hipMalloc(&Wbb[b], (size_t)(1 + 2 * nh + n_user + (mixture ? 1 : 0)) * rows * rsz));
        HIPCHK(hipMemsetAsync(Wbb[b], 0, (size_t)(1 + 2 * nh + n_user + (mixture ? 1 : 0)) * rows * rsz, stream));
        HIPCHK(hipEventCreateWithFlags(&filled[b], hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&drained[b], hipEventDisableTiming));
    }
    X = Xb[0];
    Wb = Wbb[0];
    for (size_t a = 0; a < jobs.size();) {
        size_t b = a + 1;
        while (b < jobs.size() && b - a < 4096 && jobs[b].nw == jobs[a].nw && jobs[b].blocks == jobs[a].blocks) ++b;
        batches.push_back({(int)a, (int)(b - a)});
        a = b;
    }
    weights_gram = two_sided && (gram_prec == le::kFP32 || gram_prec == le::kFP64) && 2 * basis.dim(0) <= 64 && n_user > 0;
    if (weights_gram) {
        const int P = 4 * basis.dim(0) * basis.dim(1), nwt = (n_user + 63) / 64;
        gram_flop_row = 0;
        for (int wt = 0; wt < nwt; ++wt)
            for (int pt = 0; pt < (P + 63) / 64; ++pt)
                for (int w = 0; w < 8; ++w)
                    if (64 * wt + 32 * (w / 4) < n_user && 64 * pt + 16 * (w % 4) < P) gram_flop_row += 2.0 * 2 * 16 * 16;
    }
    for (int b = 0; b < 2; ++b) {
        std::vector<le::GramJob> tab;
        for (auto& j : jobs)
            tab.push_back({weight_ptr(j.wid[0], b), j.nw > 1 ? weight_ptr(j.wid[1], b) : nullptr, j.C[0], j.C[1]});
        HIPCHK(dalloc(&jobtab[b], std::max<size_t>(1, tab.size())));
        if (!tab.empty())
            HIPCHK(hipMemcpy(jobtab[b], tab.data(), tab.size() * sizeof(le::GramJob), hipMemcpyHostToDevice));
    }
    HIPCHK(hipStreamCreateWithFlags(&gstream, hipStreamNonBlocking));
    auto ft = basis.factor_table();
    HIPCHK(dalloc(&fac, std::max<size_t>(1, ft.size())));
    if (!ft.empty()) HIPCHK(hipMemcpy(fac, ft.data(), ft.size() * 4, hipMemcpyHostToDevice));
    {
        // power table p_k^r for k r <= K, and each state's plan of table indices
        std::vector<int> pk, pr;
        std::vector<int> off(basis.K + 2, 0);
        for (int k = 1; k <= basis.K; ++k) {
            off[k] = (int)pk.size();
            for (int r = 1; k * r <= basis.K; ++r) {
                pk.push_back(k);
                pr.push_back(r);
            }
        }
        npw = (int)pk.size();
        if (npw > 128) return fail(LE_ERR_CAPACITY, "power table larger than 128 entries");
        std::vector<unsigned long long> plan(std::max(1, D), 0ull);
        for (int t = 0; t < D; ++t) {
            int cnt = 0;
            unsigned long long w = 0;
            for (const auto& pp : basis.states[t].pairs) {
                if (pp.number == 0) continue;
                const unsigned long long idx = (unsigned long long)(off[pp.number] + pp.repetitions - 1);
                w |= idx << (4 + 7 * cnt);
                ++cnt;
            }
            plan[t] = w | (unsigned long long)cnt;
        }
        HIPCHK(dalloc(&prod_d, plan.size()));
        HIPCHK(hipMemcpy(prod_d, plan.data(), plan.size() * sizeof(unsigned long long), hipMemcpyHostToDevice));
        HIPCHK(dalloc(&pwk_d, (size_t)std::max(1, npw)));
        HIPCHK(dalloc(&pwr_d, (size_t)std::max(1, npw)));
        if (npw) {
            HIPCHK(hipMemcpy(pwk_d, pk.data(), npw * sizeof(int), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(pwr_d, pr.data(), npw * sizeof(int), hipMemcpyHostToDevice));
        }
    }
    HIPCHK(dalloc(&col_re, (size_t)std::max(1, D)));
    HIPCHK(dalloc(&col_dim, (size_t)std::max(1, D)));
    if (D) {
        HIPCHK(hipMemcpy(col_re, cre.data(), D * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(col_dim, cdim.data(), D * 4, hipMemcpyHostToDevice));
    }
    if (mc.trace_walkers > 0) {
        const int tw = std::min(mc.trace_walkers, mc.walkers);
        trace_cap_rows = (long)std::max(1, mc.samples) * tw;
        HIPCHK(dalloc(&trace_d, (size_t)trace_cap_rows * 9));
    }
    if (mc.fused_step) {
        HIPCHK(dalloc(&step_q, (size_t)le::kStepQueueWords));
        HIPCHK(hipMemsetAsync(step_q, 0, sizeof(unsigned) * le::kStepQueueWords, stream));
    }
    if (rowwise) {
        const long be = basis.block_elems();
        std::vector<int> rs(std::max(1, D)), so(basis.sector_offset.begin(), basis.sector_offset.end());
        std::vector<long> bo(std::max(1, ns));
        for (int s = 0; s < ns; ++s) {
            bo[s] = basis.block_offset(s);
            for (int t = basis.sector_offset[s]; t < basis.sector_offset[s + 1]; ++t) rs[t] = s;
        }
        HIPCHK(dalloc(&row_pk, (size_t)mc.walkers * le::kMaxFactors));
        HIPCHK(dalloc(&row_acc, (size_t)std::max<long>(1, 5 * be)));
        HIPCHK(le::launch_zero(row_acc, 5 * be, stream));
        HIPCHK(dalloc(&row_sector, rs.size()));
        HIPCHK(hipMemcpy(row_sector, rs.data(), rs.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(dalloc(&sec_off_d, so.size()));
        HIPCHK(hipMemcpy(sec_off_d, so.data(), so.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(dalloc(&sec_bo_d, bo.size()));
        HIPCHK(hipMemcpy(sec_bo_d, bo.data(), bo.size() * sizeof(long), hipMemcpyHostToDevice));
        if (rec_h) (void)hipHostFree(rec_h);
        HIPCHK(hipHostMalloc((void**)&rec_h, sizeof(double) * (size_t)std::max<long>(2, 5 * be), hipHostMallocDefault));
        HIPCHK(rec.init((size_t)5 * be));
    }
    ev.resize(2 * spl + 4);
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    HIPCHK(hipStreamSynchronize(stream));
    st.step = step;
    return LE_OK;
}

extern "C" int le_engine_create(const le_system_params* sys, const le_mc_params* mc, const int* sectors,
                                int n_sectors, const le_observables* obs, le_engine** out)
{
    if (obs) {
        if (obs->n_harmonics < 0 || obs->n_harmonics > le::kMaxHarmonics || (obs->n_harmonics && !obs->harmonics))
            return fail(LE_ERR_ARG, "n_harmonics must be in [0, 8] with a harmonics array");
        for (int h = 0; h < obs->n_harmonics; ++h)
            if (obs->harmonics[h] < 0 || obs->harmonics[h] > 255) return fail(LE_ERR_ARG, "harmonic L must be in [0, 255]");
        if (obs->n_user_weights < 0 || obs->n_user_weights > 8) return fail(LE_ERR_ARG, "n_user_weights must be in [0, 8]");
    }
    if (!sys || !mc || !out || (n_sectors > 0 && !sectors) || n_sectors <= 0)
        return fail(LE_ERR_ARG, "le_engine_create: null argument or no sectors");
    *out = nullptr;
    if (sys->particles_number < 2 || sys->inverse_filling < 1)
        return fail(LE_ERR_ARG, "need particles_number >= 2 and inverse_filling >= 1");
    if (mc->walkers < 1 || mc->thin < 1 || mc->n_blocks < 1 || mc->samples < 0 || mc->burnin_sweeps < 0)
        return fail(LE_ERR_ARG, "walkers, thin and n_blocks must be >= 1; samples, burn-in >= 0");
    if (mc->precision < LE_PREC_FP32 || mc->precision > LE_PREC_BF16X6)
        return fail(LE_ERR_ARG, "unknown precision");
    if (mc->lanes_per_walker > 0 && !le::sweep_supported(mc->lanes_per_walker))
        return fail(LE_ERR_ARG, "lanes_per_walker must be -1, 0, 2, 4, 8, 16, 32 or 64");
    if (mc->graphs != 0 && mc->graphs != 1) return fail(LE_ERR_ARG, "graphs must be 0 or 1");
    if (mc->symmetry_moves < 0 || mc->symmetry_moves > 2) return fail(LE_ERR_ARG, "symmetry_moves must be 0, 1 or 2");
    if (mc->sweep_registers != 0 && mc->sweep_registers != 1) return fail(LE_ERR_ARG, "sweep_registers must be 0 or 1");
    if (mc->gram_xcd_spread != 0 && mc->gram_xcd_spread != 1 && mc->gram_xcd_spread != 2 && mc->gram_xcd_spread != 4 &&
        mc->gram_xcd_spread != 8)
        return fail(LE_ERR_ARG, "gram_xcd_spread must be 0, 1, 2, 4 or 8");
    if (mc->gram_tile_cols != 0 && mc->gram_tile_cols != 32 && mc->gram_tile_cols != 64)
        return fail(LE_ERR_ARG, "gram_tile_cols must be 0, 32 or 64");
    if (mc->gram_tile_cols == 32 && (mc->precision == LE_PREC_BF16X3 || mc->precision == LE_PREC_BF16X6))
        return fail(LE_ERR_ARG, "the BF16 split Gram runs 64-column tiles only");
    if (mc->sweep_registers == 1 && mc->lanes_per_walker > 0 &&
        !le::sweep_reg_supported(mc->lanes_per_walker, sys->particles_number, mc->estimator == LE_EST_MIXTURE))
        return fail(LE_ERR_ARG, "sweep_registers needs lanes_per_walker 8 or 16 with ceil(N / lanes) in [2, 25] "
                                "(13 or 25 for the mixture estimator)");
    std::vector<int> dls(sectors, sectors + n_sectors);
    for (int dl : dls)
        if (dl < 0 || dl > 40) return fail(LE_ERR_ARG, "sector dL must be in [0, 40]");
    if (mc->estimator != LE_EST_PLAIN && mc->estimator != LE_EST_ROWWISE && mc->estimator != LE_EST_MIXTURE)
        return fail(LE_ERR_ARG, "estimator must be LE_EST_PLAIN, LE_EST_ROWWISE or LE_EST_MIXTURE");
    if (mc->estimator == LE_EST_MIXTURE) {
        if (mc->symmetry_moves > 1 || (obs && (obs->n_harmonics || obs->n_user_weights)))
            return fail(LE_ERR_ARG, "the mixture estimator takes no dilation moves or extra observables");
        if (!(mc->mixture_rho > 0.0 && mc->mixture_rho < 1.0)) return fail(LE_ERR_ARG, "mixture_rho must be in (0, 1)");
        int L = 0;
        for (int i = 0; i < n_sectors; ++i) L = std::max(L, sectors[i]);
        if (L < 1 || L > le::kMixMax) return fail(LE_ERR_ARG, "the mixture estimator needs a largest sector dL in [1, 24]");
    }
    if (mc->estimator == LE_EST_ROWWISE &&
        (mc->symmetry_moves || mc->sweep_registers || (obs && (obs->n_harmonics || obs->n_user_weights))))
        return fail(LE_ERR_ARG, "the row-wise estimator takes no symmetry moves, register sweeps or extra observables");
    if (mc->fused_step != 0 && mc->fused_step != 1) return fail(LE_ERR_ARG, "fused_step must be 0 or 1");
    if (mc->fused_rows < 0 || mc->fused_rows % 32) return fail(LE_ERR_ARG, "fused_rows must be 0 or a positive multiple of 32");
    if (mc->fused_step &&
        (mc->estimator != LE_EST_PLAIN || mc->precision != LE_PREC_FP32 || mc->symmetry_moves ||
         (mc->lanes_per_walker != 0 && mc->lanes_per_walker != -1 && mc->lanes_per_walker != 8) ||
         (obs && (obs->n_harmonics || obs->n_user_weights))))
        return fail(LE_ERR_ARG, "fused_step needs the plain estimator, FP32, no symmetry moves, 8 lanes per walker "
                                "and S, H only");
    auto e = new le_engine();
    e->sys = *sys;
    e->mc = *mc;
    e->prec = mc->precision == LE_PREC_FP64 ? LE_PREC_FP64 : LE_PREC_FP32;
    e->gram_prec = mc->precision == LE_PREC_FP64     ? le::kFP64
                  : mc->precision == LE_PREC_BF16X3 ? le::kBF16X3
                  : mc->precision == LE_PREC_BF16X6 ? le::kBF16X6
                                                    : le::kFP32;
    e->basis.build(sys->particles_number, dls);
    if (mc->estimator == LE_EST_MIXTURE) {
        e->mixture = true;
        for (int dl : dls) e->mix_L = std::max(e->mix_L, dl);
    }
    // capacity checks (reference: em_spectrum.cu:235-252 exits); here they are codes
    if (e->basis.K > 24) {
        delete e;
        return fail(LE_ERR_CAPACITY, "largest power-sum degree " + std::to_string(e->basis.K) + " exceeds 24");
    }
    if (e->basis.longest > le::kMaxFactors) {
        delete e;
        return fail(LE_ERR_CAPACITY, "compressed partition longer than 8");
    }
    if (e->basis.D > 4096) {
        delete e;
        return fail(LE_ERR_CAPACITY, "basis larger than 4096 states");
    }
    for (int s = 0; s < e->basis.nsec(); ++s)
        if (e->basis.dim(s) > 640) {
            delete e;
            return fail(LE_ERR_CAPACITY, "sector larger than 640 states");
        }
    if (mc->estimator == LE_EST_ROWWISE) {
        const long chains = (long)e->basis.D * mc->walkers;
        if (chains > (1L << 26)) {
            delete e;
            return fail(LE_ERR_CAPACITY, "row-wise estimator: basis states x walkers above 2^26 chains");
        }
        e->rowwise = true;
        e->row_W = mc->walkers;
        e->mc.walkers = (int)chains;
        e->mc.samples_per_launch = 1;
        e->mc.graphs = 0;
        if (e->mc.lanes_per_walker < 0) e->mc.lanes_per_walker = 0;   // the timed autotune would advance the chains
    }
    int rc = e->setup(obs);
    if (rc != LE_OK) {
        std::string msg = g_err;
        delete e;
        g_err = msg;
        return rc;
    }
    e->gram_spread = mc->gram_xcd_spread > 0 ? mc->gram_xcd_spread : 8;
    e->gram_tc = mc->gram_tile_cols > 0 ? mc->gram_tile_cols : 64;
    e->need_spread_tune = mc->gram_xcd_spread == 0 ||
                          (mc->gram_tile_cols == 0 && e->gram_prec != le::kBF16X3 && e->gram_prec != le::kBF16X6);
    e->G = mc->lanes_per_walker > 0 ? mc->lanes_per_walker : default_lanes(sys->particles_number, e->prec);
    if (mc->lanes_per_walker > 0 && mc->sweep_registers == 1) e->G |= kRegSweep;
    if (mc->fused_step) {
        // the fused kernel's sweep: particles in registers at 8 lanes per walker
        if (!le::step_supported(8, sys->particles_number, e->basis.K)) {
            delete e;
            return fail(LE_ERR_ARG, "fused_step needs ceil(N / 8) = 13 or 25 and power sums up to K <= 16");
        }
        e->G = 8 | kRegSweep;
        e->mc.lanes_per_walker = 8;
        e->mc.sweep_registers = 1;
    }
    *out = e;
    return LE_OK;
}

extern "C" int le_basis_describe(int N, const int* sectors, int n_sectors, int capacity, int* D, int* K,
                                 int* dls, int* pairs, int* n_pairs)
{
    if (N < 1 || n_sectors <= 0 || !sectors || !D) return fail(LE_ERR_ARG, "le_basis_describe: bad argument");
    for (int q = 0; q < n_sectors; ++q)
        if (sectors[q] < 0 || sectors[q] > 40) return fail(LE_ERR_ARG, "sector dL must be in [0, 40]");
    le::Basis b;
    b.build(N, std::vector<int>(sectors, sectors + n_sectors));
    *D = b.D;
    if (K) *K = b.K;
    for (int t = 0; t < b.D && t < capacity; ++t) {
        if (dls) dls[t] = b.states[t].dL;
        if (n_pairs) n_pairs[t] = (int)b.states[t].pairs.size();
        if (pairs)
            for (size_t q = 0; q < b.states[t].pairs.size() && q < 8; ++q) {
                pairs[16 * t + 2 * q] = b.states[t].pairs[q].number;
                pairs[16 * t + 2 * q + 1] = b.states[t].pairs[q].repetitions;
            }
    }
    return LE_OK;
}

extern "C" int le_engine_destroy(le_engine* e)
{
    delete e;
    return LE_OK;
}

extern "C" int le_engine_get_info(const le_engine* e, le_engine_info* info)
{
    if (!e || !info) return fail(LE_ERR_ARG, "null argument");
    info->D = e->basis.D;
    info->K = e->basis.K;
    info->n_sectors = e->basis.nsec();
    info->longest_partition = e->basis.longest;
    info->block_elems = e->basis.block_elems();
    info->lanes_per_walker = map_lanes(e->G);
    info->sweep_in_registers = map_reg(e->G) ? 1 : 0;
    info->samples_per_launch = e->spl;
    info->n_tiles = (e->gram_tc == 32 && !e->jobs.empty()) ? e->jobs[0].ntiles32 : e->ntiles;
    info->gram_xcd_spread = e->gram_spread;
    info->gram_tile_cols = e->gram_tc;
    info->chains = e->mc.walkers;
    info->x_ld = e->ldx;
    info->gram_mfma_flop_per_row = e->gram_tc == 32 ? e->gram_flop_row32 : e->gram_flop_row;
    return LE_OK;
}

extern "C" int le_set_mixture(le_engine* e, const double* nk, int n_k, double rho)
{
    if (!e) return fail(LE_ERR_ARG, "null engine");
    return e->set_mixture(nk, n_k, rho);
}

extern "C" int le_get_mixture(const le_engine* e, int* L, double* c0, double* a, double* b)
{
    if (!e || !L) return fail(LE_ERR_ARG, "null argument");
    *L = e->mixp.kmix;
    if (c0) *c0 = e->mixp.c0;
    for (int k = 0; k < e->mixp.kmix; ++k) {
        if (a) a[k] = e->mixp.a[k];
        if (b) b[k] = e->mixp.b[k];
    }
    return LE_OK;
}

extern "C" int le_sector_dims(const le_engine* e, int* dims, int* offsets)
{
    if (!e) return fail(LE_ERR_ARG, "null engine");
    for (int s = 0; s < e->basis.nsec(); ++s) {
        if (dims) dims[s] = e->basis.dim(s);
    }
    if (offsets)
        for (int s = 0; s <= e->basis.nsec(); ++s) offsets[s] = e->basis.sector_offset[s];
    return LE_OK;
}

extern "C" int le_basis_state(const le_engine* e, int state, int* dL, int* pairs, int* n_pairs)
{
    if (!e || state < 0 || state >= e->basis.D) return fail(LE_ERR_ARG, "state out of range");
    const auto& s = e->basis.states[state];
    if (dL) *dL = s.dL;
    if (n_pairs) *n_pairs = (int)s.pairs.size();
    if (pairs)
        for (size_t q = 0; q < s.pairs.size() && q < 8; ++q) {
            pairs[2 * q] = s.pairs[q].number;
            pairs[2 * q + 1] = s.pairs[q].repetitions;
        }
    return LE_OK;
}

int le_engine::run_sweeps(int nsweeps, unsigned long long sweep0, void* positions,
                          const unsigned long long* sweep_base, bool mix_fresh)
{
    le::SweepArgs a{};
    a.sweep_base = sweep_base;
    a.pos = positions;
    a.accepted = acc;
    a.seed = mc.seed;
    a.sweep0 = sweep0;
    a.N = sys.particles_number;
    a.m = sys.inverse_filling;
    a.walkers = mc.walkers;
    a.nsweeps = nsweeps;
    a.step = step;
    if (rowwise) {
        a.walker_id0 = row_W;
        a.row_walkers = row_W;
        a.row_fac = fac;
        a.row_pk = row_pk;
        a.R = R;
        HIPCHK(le::launch_sweep_row(a, prec, map_lanes(G), stream));
        return LE_OK;
    }
    if (two_sided) {
        if (map_reg(G)) HIPCHK(le::launch_sweep2_reg(a, prec, map_lanes(G), stream));
        else HIPCHK(le::launch_sweep2(a, prec, G, stream));
    } else if (proposal == LE_PROPOSAL_LEARNED) {
        le::LearnedTable lt;
        lt.table = table;
        lt.T = table_T;
        for (int q = 0; q < 3; ++q) {
            lt.lo[q] = table_lo[q];
            lt.hi[q] = table_hi[q];
        }
        lt.invR = 1.0 / R;
        if (map_reg(G) && le::learned_reg_supported(map_lanes(G), sys.particles_number))
            HIPCHK(le::launch_sweep_learned_reg(a, lt, map_lanes(G), stream));
        else
            HIPCHK(le::launch_sweep_learned(a, lt, std::max(map_lanes(G), 8), stream));
    } else if (mixture && mixp.kmix > 0) {
        // launches of at most kMixChunk sweeps, each starting its running p_k from the positions
        a.mix = mixp;
        a.mix_pk = mix_pk;
        a.R = R;
        for (int done = 0; done < nsweeps; done += kMixChunk) {
            a.sweep0 = sweep0 + (unsigned long long)done;
            a.nsweeps = std::min(kMixChunk, nsweeps - done);
            if (done > 0 || !mix_fresh) HIPCHK(le::launch_mix_init(a, prec, stream));
            if (map_reg(G)) HIPCHK(le::launch_sweep_reg(a, prec, map_lanes(G), stream));
            else HIPCHK(le::launch_sweep(a, prec, G, stream));
        }
    } else if (map_reg(G)) {
        HIPCHK(le::launch_sweep_reg(a, prec, map_lanes(G), stream));
    } else {
        HIPCHK(le::launch_sweep(a, prec, G, stream));
    }
    return LE_OK;
}

// Mixture density from its normalisers n_k = <|p_k|^2> (k = 1..L): term k is
// c_k (|p_k|^2 / n_k)^(L/k) with c_k = (1 - rho) / (L Gamma(L/k + 1)), stored as
// exp(a_k ln|p_k|^2 + b_k).
int le_engine::set_mixture(const double* nk, int n, double rho)
{
    if (!mixture) return fail(LE_ERR_STATE, "le_set_mixture needs an engine created with LE_EST_MIXTURE");
    if (!nk || n != mix_L) return fail(LE_ERR_ARG, "le_set_mixture: n_k must hold L = " + std::to_string(mix_L) + " values");
    if (!(rho > 0.0 && rho < 1.0)) return fail(LE_ERR_ARG, "mixture rho must be in (0, 1)");
    for (int k = 0; k < n; ++k)
        if (!(nk[k] > 0.0) || !std::isfinite(nk[k])) return fail(LE_ERR_ARG, "mixture normalisers n_k must be finite and > 0");
    le::MixParams m{};
    m.kmix = n;
    m.c0 = rho;
    for (int k = 1; k <= n; ++k) {
        const double ak = (double)n / (double)k;
        m.a[k - 1] = ak;
        m.b[k - 1] = std::log((1.0 - rho) / n) - std::lgamma(ak + 1.0) - ak * std::log(nk[k - 1]);
    }
    mixp = m;
    drop_graphs();   // the captured rounds hold the previous density
    return LE_OK;
}

// n_k = mean |p_k|^2 over the walkers at their current positions
int le_engine::calibrate_mixture()
{
    le::SweepArgs a{};
    a.pos = pos;
    a.N = sys.particles_number;
    a.walkers = mc.walkers;
    a.R = R;
    a.mix.kmix = mix_L;
    a.mix_pk = mix_pk;
    HIPCHK(le::launch_mix_init(a, prec, stream));
    std::vector<double2> pk((size_t)mc.walkers * le::kMixMax);
    HIPCHK(hipMemcpyAsync(pk.data(), mix_pk, sizeof(double2) * pk.size(), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    std::vector<double> nk(mix_L, 0.0);
    for (int w = 0; w < mc.walkers; ++w)
        for (int k = 0; k < mix_L; ++k) {
            const double2 v = pk[(size_t)w * le::kMixMax + k];
            nk[k] += v.x * v.x + v.y * v.y;
        }
    for (auto& v : nk) v /= mc.walkers;
    return set_mixture(nk.data(), mix_L, mc.mixture_rho);
}

void le_engine::mix_features(le::FeatureArgs& f, bool walkers) const
{
    if (!mixture) return;
    f.mix = mixp;
    f.mix_row = mix_row;
    f.mix_acc = mix_acc;
    f.mix_pk_out = walkers ? mix_pk : nullptr;
}

// Exact centre-of-mass heat bath + global rotation (+ dilation of the relative coordinates at
// symmetry_moves = 2) of every walker (edge engine only).
int le_engine::symmetry(unsigned long long sweep0, const unsigned long long* sweep_base)
{
    if (!mc.symmetry_moves || two_sided) return LE_OK;
    le::SweepArgs a{};
    a.pos = pos;
    a.seed = mc.seed;
    a.sweep0 = sweep0;
    a.sweep_base = sweep_base;
    a.N = sys.particles_number;
    a.m = sys.inverse_filling;
    a.walkers = mc.walkers;
    a.dilate = mc.symmetry_moves >= 2;
    if (mixture) {
        if (mixp.kmix == 0) return LE_OK;   // before the density is set the chains are plain
        a.mix = mixp;
        a.R = R;
        HIPCHK(le::launch_symmetry_mix(a, prec, stream));
        return LE_OK;
    }
    HIPCHK(le::launch_symmetry(a, prec, stream));
    return LE_OK;
}

// Time two sweeps per candidate mapping on a scratch copy of the walkers.
int le_engine::autotune_lanes()
{
    const int N = sys.particles_number;
    // the two-sided chain has its own kernel (and no register variant): its own entries
    auto key = std::make_tuple(N, mc.walkers, prec, two_sided ? 1 : 0);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned_lanes.find(key);
        if (it != g_tuned_lanes.end()) {
            G = it->second;
            return LE_OK;
        }
    }
    void* scratch = nullptr;
    const size_t bytes = (size_t)mc.walkers * npos() * 2 * rsz;
    HIPCHK(hipMalloc(&scratch, bytes));
    unsigned* acc_save = acc;
    unsigned* acc_tmp = nullptr;
    HIPCHK(dalloc(&acc_tmp, (size_t)mc.walkers));
    acc = acc_tmp;
    float best = 1e30f;
    int bestG = 8;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    std::vector<int> cands;
    for (int c : {4, 8, 16, 32, 64}) {
        if ((N + c - 1) / c < 2 && c > 4) continue;
        if ((two_sided ? le::sweep2_lds_bytes(prec, c, N) : le::sweep_lds_bytes(prec, c, N)) > 160 * 1024) continue;
        cands.push_back(c);
    }
    for (int c : {8, 16})
        if (two_sided ? le::sweep2_reg_supported(c, N) : le::sweep_reg_supported(c, N, mixture)) cands.push_back(c | kRegSweep);
    for (int cand : cands) {
        G = cand;
        HIPCHK(hipMemcpyAsync(scratch, pos, bytes, hipMemcpyDeviceToDevice, stream));
        int rc = run_sweeps(1, 1ull << 46, scratch);   // warm-up, separate RNG range
        if (rc) break;
        HIPCHK(hipEventRecord(e0, stream));
        rc = run_sweeps(2, (1ull << 46) + 8, scratch);
        if (rc) break;
        HIPCHK(hipEventRecord(e1, stream));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
            best = ms;
            bestG = cand;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(scratch);
    (void)hipFree(acc_tmp);
    acc = acc_save;
    G = bestG;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned_lanes[key] = G;
    return LE_OK;
}

// Fresh walkers (uniform in the disc); row-wise chains also take their row state's p_k.
hipError_t le_engine::init_positions()
{
    hipError_t h = le::launch_init_positions(pos, prec, npos(), mc.walkers, R, mc.seed, stream, rowwise ? row_W : 0);
    if (h != hipSuccess || !rowwise) return h;
    le::SweepArgs a{};
    a.pos = pos;
    a.N = sys.particles_number;
    a.walkers = mc.walkers;
    a.row_walkers = row_W;
    a.row_fac = fac;
    a.row_pk = row_pk;
    a.R = R;
    return le::launch_row_init(a, prec, stream);
}

extern "C" int le_init_walkers(le_engine* e)
{
    if (!e) return fail(LE_ERR_ARG, "null engine");
    HIPCHK(e->init_positions());
    HIPCHK(hipMemsetAsync(e->acc, 0, sizeof(unsigned) * e->mc.walkers, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    e->walkers_ready = true;
    e->sweep_counter = 0;
    e->sweeps_done = 0;
    if (e->mc.lanes_per_walker < 0) {
        int rc = e->autotune_lanes();
        if (rc) return rc;
    }
    return LE_OK;
}

double le_engine::measure_acceptance(double trial_step, int attempt, int* rc)
{
    // fresh positions and tune_sweeps sweeps, as mcmc_acceptanceBisection (em_bisection.cu:7-47)
    *rc = LE_OK;
    step = trial_step;
    hipError_t h = init_positions();
    if (h == hipSuccess) h = hipMemsetAsync(acc, 0, sizeof(unsigned) * mc.walkers, stream);
    if (h != hipSuccess) {
        *rc = fail(LE_ERR_HIP, hipGetErrorString(h));
        return 0;
    }
    *rc = run_sweeps(mc.tune_sweeps, (1ull << 47) + ((unsigned long long)attempt << 20), pos);
    if (*rc) return 0;
    h = hipMemcpyAsync(acc_h, acc, sizeof(unsigned) * mc.walkers, hipMemcpyDeviceToHost, stream);
    if (h == hipSuccess) h = hipStreamSynchronize(stream);
    if (h != hipSuccess) {
        *rc = fail(LE_ERR_HIP, hipGetErrorString(h));
        return 0;
    }
    double tot = 0;
    for (int w = 0; w < mc.walkers; ++w) tot += acc_h[w];
    return tot / ((double)mc.walkers * mc.tune_sweeps * npos());
}

// edgeMC::optimal_step (em_bisection.cu:49-153): le_step_search (le_tune.cpp) over
// acceptances measured on fresh walkers, tune_sweeps sweeps each (mcmc_acceptanceBisection,
// em_bisection.cu:7-47).
namespace {
struct TuneCtx {
    le_engine* e;
    int attempt;
};
int tune_measure(void* ctx, double step, double* acceptance)
{
    auto* t = static_cast<TuneCtx*>(ctx);
    int rc = LE_OK;
    *acceptance = t->e->measure_acceptance(step, t->attempt++, &rc);
    return rc;
}
}  // namespace

extern "C" int le_tune_step(le_engine* e, double* step_out)
{
    if (!e) return fail(LE_ERR_ARG, "null engine");
    if (e->mc.tune_sweeps <= 0) return fail(LE_ERR_ARG, "tune_sweeps must be > 0");
    TuneCtx ctx{e, 0};
    double s = e->step;
    const double s_in = e->step;
    int rc = le_step_search(&tune_measure, &ctx, s_in, e->mc.target_acceptance, e->mc.acceptance_tol, &s, nullptr);
    if (rc != LE_OK) {
        e->step = s_in;
        // leave fresh walkers whatever the outcome
        const std::string msg = g_err;
        (void)le_init_walkers(e);
        g_err = msg;
        return rc;
    }
    e->step = s;
    e->st.step = s;
    if (step_out) *step_out = s;
    // leave fresh walkers for the burn-in
    return le_init_walkers(e);
}

extern "C" int le_burnin(le_engine* e, int sweeps)
{
    if (!e) return fail(LE_ERR_ARG, "null engine");
    if (!e->walkers_ready) {
        int rc = le_init_walkers(e);
        if (rc) return rc;
    }
    if (sweeps <= 0) return LE_OK;
    int first = sweeps;
    if (e->mixture && e->mixp.kmix == 0) first = sweeps / 2;   // |Psi_L|^2 first, then w |Psi_L|^2
    int rc = e->run_sweeps(first, e->sweep_counter, e->pos);
    if (rc) return rc;
    if (first < sweeps) {
        rc = e->calibrate_mixture();
        if (rc) return rc;
        rc = e->run_sweeps(sweeps - first, e->sweep_counter + (unsigned long long)first, e->pos);
        if (rc) return rc;
    }
    e->sweep_counter += (unsigned long long)sweeps;
    e->sweeps_done += sweeps;
    e->st.sweeps = e->sweeps_done;
    HIPCHK(hipStreamSynchronize(e->stream));
    return LE_OK;
}

hipEvent_t le_engine::next_event()
{
    if (evused == evpool.size()) {
        hipEvent_t q = nullptr;
        if (hipEventCreate(&q) != hipSuccess) return nullptr;
        evpool.push_back(q);
    }
    return evpool[evused++];
}

int le_engine::gram_rows(long rows, int buf, hipStream_t s)
{
    if (!s) s = stream;
    if (weights_gram) {
        le::WeightGramArgs g{};
        g.X = Xb[buf];
        g.ldx = ldx;
        g.W = weight_ptr(LE_W_USER(0), buf);
        g.wld = wrows;
        g.nweights = n_user;
        g.jobs = jobtab[buf];
        g.r0 = colb[0];
        g.nr = 2 * basis.dim(0);
        g.c0 = colb[1];
        g.nc = 2 * basis.dim(1);
        g.rows = rows;
        g.ldc = ldx;
        const long tiles = (long)((n_user + 63) / 64) * ((g.nr * g.nc + 63) / 64);
        long splits = std::max<long>(1, (6144 + tiles - 1) / tiles);
        long rps = std::max<long>((rows + splits - 1) / splits, 1024);
        rps = ((rps + 31) / 32) * 32;
        g.rows_per_split = rps;
        g.splits = (int)((rows + rps - 1) / rps);
        HIPCHK(le::launch_gram_weights(g, gram_prec, s));
        return LE_OK;
    }
    for (auto& bt : batches) {
        const Job& j = jobs[bt.first];
        le::GramArgs g{};
        g.X = Xb[buf];
        g.ldx = ldx;
        g.w0 = weight_ptr(j.wid[0], buf);
        g.w1 = j.nw > 1 ? weight_ptr(j.wid[1], buf) : nullptr;
        if (bt.count > 1) {
            g.jobs = jobtab[buf] + bt.first;
            g.njobs = bt.count;
        }
        g.nw = j.nw;
        g.rows = rows;
        // ~6144 workgroups (24 per CU) per launch: with 2-3 resident 512-thread workgroups per CU
        // the last, partial round of workgroups is then a small part of the launch (at 2048
        // workgroups N = 200, dL 1..16 needed 4.02 rounds and lost ~20% to the fifth)
        const bool narrow = gram_tc == 32;
        const long want_wg = narrow ? 12288 : 6144;   // the same threads with 256-thread workgroups
        const long per_split = (long)std::max(1, narrow ? j.ntiles32 : j.ntiles) * bt.count;
        long splits = std::max<long>(1, (want_wg + per_split - 1) / per_split);
        // with the tiles of a split on K XCDs, splits come in multiples of the 8 / K groups
        const int K = gram_spread;
        const long groups = (K == 1 || K == 2 || K == 4) ? 8 / K : 1;
        splits = ((splits + groups - 1) / groups) * groups;
        long rps = (rows + splits - 1) / splits;
        rps = std::max<long>(rps, 1024);
        rps = ((rps + 31) / 32) * 32;
        splits = (rows + rps - 1) / rps;
        g.xcd_spread = K;
        g.rows_per_split = rps;
        g.splits = (int)splits;
        g.tiles = narrow ? j.tiles32 : j.tiles;
        g.ntiles = narrow ? j.ntiles32 : j.ntiles;
        g.tile_cols = gram_tc;
        g.C0 = j.C[0];
        g.C1 = j.C[1];
        g.ldc = ldx;
        HIPCHK(le::launch_gram(g, gram_prec, (int)splits, s));
    }
    return LE_OK;
}

int le_engine::close_block(long long samples)
{
    // blocks are means over the samples, or for the mixture estimator sums over the sum of 1 / w
    double norm = (double)samples;
    if (mixture) {
        HIPCHK(hipMemcpyAsync(mix_acc_h, mix_acc, sizeof(double) * wrows, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        norm = 0.0;
        for (long r = 0; r < wrows; ++r) norm += mix_acc_h[r];
        HIPCHK(le::launch_zero(mix_acc, wrows, stream));
    }
    le::ExtractArgs x{};
    x.items = items_d;
    x.nitems = (int)items.size();
    x.total = rec_elems;
    x.ldc = ldx;
    x.scale = norm > 0.0 ? 1.0 / norm : 0.0;
    x.out = rec_d;
    HIPCHK(le::launch_extract(x, stream));
    double* slot = nullptr;
    HIPCHK(rec.next(rec_h, &slot));
    HIPCHK(hipMemcpyAsync(slot, rec_d, sizeof(double) * 2 * rec_elems, hipMemcpyDeviceToHost, stream));
    HIPCHK(le::launch_zero(cpool, (long)cpool_elems, stream));
    HIPCHK(hipStreamSynchronize(stream));
    rec.push(slot);
    block_grams = 0;
    blk_samples.push_back(samples);
    blk_norm.push_back(norm);
    return LE_OK;
}

// Capture (once per set of baked-in parameters) the sampler work of one full round: spl x
// (thin sweeps + features) on the sampler stream.  The Philox sweep base and the trace row
// base are read from bases_d, which a one-thread kernel sets before each replay.
int le_engine::sample_graph(le::FeatureArgs f, int buf, bool traced, int tw, hipGraphExec_t* out)
{
    auto find = [&](int b) -> hipGraphExec_t {
        for (auto& g : graphs)
            if (g.buf == b && g.traced == traced && g.thin == mc.thin && g.step == step && g.proposal == proposal &&
                g.G == G && g.table == table && g.table_T == table_T && g.n == spl)
                return g.exec;
        return nullptr;
    };
    if ((*out = find(buf))) return LE_OK;
    // Capture the round for both sample buffers at once, so that a later call that starts on
    // the other buffer replays instead of capturing inside its own time.
    for (int b : {1 - buf, buf}) {
        if (b != buf && find(b)) continue;
        le::FeatureArgs fb = f;
        fb.X = Xb[b];
        fb.W = Wbb[b];
        hipGraphExec_t ex = nullptr;
        int rc = capture_round(fb, b, traced, tw, &ex);
        if (rc) return rc;
        if (b == buf) *out = ex;
    }
    return LE_OK;
}

int le_engine::capture_round(le::FeatureArgs f, int buf, bool traced, int tw, hipGraphExec_t* out)
{
    for (size_t q = 0; q < graphs.size();)
        if (graphs[q].buf == buf && graphs[q].traced == traced) {
            (void)hipGraphExecDestroy(graphs[q].exec);
            graphs.erase(graphs.begin() + (long)q);
        } else {
            ++q;
        }
    if (!bases_d) HIPCHK(dalloc(&bases_d, 2));
    HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    int rc = LE_OK;
    hipError_t h = hipSuccess;
    for (int q = 0; q < spl && rc == LE_OK && h == hipSuccess; ++q) {
        rc = run_sweeps(mc.thin, (unsigned long long)q * (unsigned long long)mc.thin, pos, bases_d, mixture);
        if (rc == LE_OK) rc = symmetry((unsigned long long)(q + 1) * (unsigned long long)mc.thin, bases_d);
        f.row0 = (long)q * mc.walkers;
        f.trace = traced ? trace_d : nullptr;
        f.trace_walkers = traced ? tw : 0;
        f.trace_row0 = (long)q * tw;
        f.trace_base = traced ? reinterpret_cast<const long long*>(bases_d + 1) : nullptr;
        if (rc == LE_OK) h = le::launch_features(f, prec, stream);
    }
    hipGraph_t graph = nullptr;
    const hipError_t he = hipStreamEndCapture(stream, &graph);
    if (rc != LE_OK || h != hipSuccess || he != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        if (rc != LE_OK) return rc;
        return fail(LE_ERR_HIP, hipGetErrorString(h != hipSuccess ? h : he));
    }
    hipGraphExec_t ex = nullptr;
    h = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (h != hipSuccess) return fail(LE_ERR_HIP, hipGetErrorString(h));
    // upload now, so that the graph's first replay (possibly inside a timed call) does not pay it
    h = hipGraphUpload(ex, stream);
    if (h != hipSuccess) {
        (void)hipGraphExecDestroy(ex);
        return fail(LE_ERR_HIP, hipGetErrorString(h));
    }
    graphs.push_back({ex, buf, traced, mc.thin, proposal, G, spl, table_T, step, table});
    *out = ex;
    return LE_OK;
}

// Time each placement of the Gram tiles over the XCDs (gram_xcd_spread 8, 4, 2, 1) with each
// tile width (64 columns, and 32 for the FP32 / FP64 tiles), those not fixed by the Monte-Carlo
// parameters: two launches each on the round in buffer `buf`, the shortest counts, on the Gram
// stream before the round's own accumulation; keep the fastest for this layout and clear the
// accumulators the timing launches wrote.  Once per (rows per launch, Gram kind, sector layout).
int le_engine::tune_gram_spread(int buf)
{
    need_spread_tune = false;
    const long rows = (long)spl * mc.walkers;
    std::vector<int> layout(basis.sectors);
    layout.push_back(n_user);
    layout.push_back(nh);
    layout.push_back(mc.gram_xcd_spread);
    layout.push_back(mc.gram_tile_cols);
    const auto key = std::make_tuple(rows, gram_prec, layout, (int)jobs.size());
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned_spread.find(key);
        if (it != g_tuned_spread.end()) {
            gram_spread = it->second.first;
            gram_tc = it->second.second;
            return LE_OK;
        }
    }
    hipEvent_t t0 = next_event(), t1 = next_event();
    if (!t0 || !t1) return fail(LE_ERR_HIP, "event creation failed");
    std::vector<int> Ks, TCs;
    if (mc.gram_xcd_spread > 0) Ks = {mc.gram_xcd_spread};
    else Ks = {8, 4, 2, 1};
    const bool split_form = gram_prec == le::kBF16X3 || gram_prec == le::kBF16X6;
    if (mc.gram_tile_cols > 0) TCs = {mc.gram_tile_cols};
    else if (split_form) TCs = {64};
    else TCs = {64, 32};
    int best = Ks[0], best_tc = TCs[0];
    float best_ms = 1e30f;
    for (int TC : TCs)
        for (int K : Ks) {
            gram_spread = K;
            gram_tc = TC;
            // two timed launches per candidate; one when a launch takes over 10 ms (the
            // excitation response's 2714-state Gram: ~140 ms each), where the launch-to-launch
            // spread is small against the differences between candidates
            for (int rep = 0; rep < 2; ++rep) {
                HIPCHK(hipEventRecord(t0, gstream));
                int rc = gram_rows(rows, buf, gstream);
                if (rc) return rc;
                HIPCHK(hipEventRecord(t1, gstream));
                HIPCHK(hipEventSynchronize(t1));
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, t0, t1));
                if (ms < best_ms) {
                    best_ms = ms;
                    best = K;
                    best_tc = TC;
                }
                if (ms > 10.f) break;
            }
        }
    gram_spread = best;
    gram_tc = best_tc;
    HIPCHK(le::launch_zero(cpool, (long)cpool_elems, gstream));
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned_spread[key] = {best, best_tc};
    return LE_OK;
}

// Row-wise estimator (em_spectrum.cu:16-137): per kept sample, thin sweeps of every chain on
// |p_f Psi_L|^2, the monomials of every chain, and the per-row sums of g/f, V g/f, |g/f|^2 over
// f's sector; one block of sums per call.
int le_engine::sample_rowwise(int samples)
{
    const auto t0 = std::chrono::steady_clock::now();
    HIPCHK(hipMemsetAsync(acc, 0, sizeof(unsigned) * mc.walkers, stream));
    le::FeatureArgs f{};
    f.ps = ps_d;
    f.ps_ld = kPsLd;
    f.pos = pos;
    f.X = Xb[0];
    f.ldx = ldx;
    f.W = Wbb[0];
    f.wld = wrows;
    f.N = sys.particles_number;
    f.walkers = mc.walkers;
    f.D = basis.D;
    f.K = basis.K;
    f.fac = fac;
    f.col_re = col_re;
    f.col_dim = col_dim;
    f.prod = prod_d;
    f.pw_k = pwk_d;
    f.pw_r = pwr_d;
    f.npw = npw;
    f.invR = 1.0 / R;
    f.pot = {sys.conf_w, sys.conf_eps, sys.conf_strength, sys.conf_exponent, R};
    le::RowAccumArgs ra{};
    ra.X = Xb[0];
    ra.ldx = ldx;
    ra.V = Wbb[0];
    ra.col_re = col_re;
    ra.col_dim = col_dim;
    ra.row_sector = row_sector;
    ra.sec_off = sec_off_d;
    ra.sec_bo = sec_bo_d;
    ra.block_elems = basis.block_elems();
    ra.D = basis.D;
    ra.row_walkers = row_W;
    ra.acc = row_acc;
    const int tw = std::min(mc.trace_walkers, mc.walkers);
    for (int q = 0; q < samples; ++q) {
        int rc = run_sweeps(mc.thin, sweep_counter, pos);
        if (rc) return rc;
        sweep_counter += (unsigned long long)mc.thin;
        if (trace_d && trace_rows + tw <= trace_cap_rows) {
            f.trace = trace_d;
            f.trace_walkers = tw;
            f.trace_row0 = trace_rows;
            trace_rows += tw;
        } else {
            f.trace = nullptr;
            f.trace_walkers = 0;
        }
        HIPCHK(le::launch_features(f, prec, stream));
        HIPCHK(le::launch_row_accum(ra, prec, stream));
    }
    sweeps_done += (long long)samples * mc.thin;
    // close the block: raw sums to the host, accumulators cleared
    const long n5 = 5 * basis.block_elems();
    HIPCHK(hipMemcpyAsync(acc_h, acc, sizeof(unsigned) * mc.walkers, hipMemcpyDeviceToHost, stream));
    double* slot = nullptr;
    HIPCHK(rec.next(rec_h, &slot));
    HIPCHK(hipMemcpyAsync(slot, row_acc, sizeof(double) * n5, hipMemcpyDeviceToHost, stream));
    HIPCHK(le::launch_zero(row_acc, n5, stream));
    HIPCHK(hipStreamSynchronize(stream));
    rec.push(slot);
    blk_samples.push_back((long long)samples * row_W);     // samples per row
    double tot = 0;
    for (int w = 0; w < mc.walkers; ++w) tot += acc_h[w];
    acc_sampling += (unsigned long long)tot;
    acc_moves += (long long)samples * mc.thin * npos() * mc.walkers;
    st.acceptance = (double)acc_sampling / (double)std::max<long long>(1, acc_moves);
    st.samples += (long long)samples * mc.walkers;
    st.sweeps = sweeps_done;
    st.step = step;
    st.sweep_launches += samples;
    st.ms_sample_total += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return LE_OK;
}

// What the fused kernel leaves out, checked at every le_sample (the proposal can change after
// creation).
int le_engine::fused_check() const
{
    if (proposal != LE_PROPOSAL_PLAIN) return fail(LE_ERR_ARG, "fused_step runs the plain proposal only");
    if (wfn || n_user || nh || weights_gram || two_sided || rowwise || mixture)
        return fail(LE_ERR_ARG, "fused_step accumulates S and H of the plain estimator only");
    if (jobs.size() != 1 || jobs[0].nw != 2 || batches.size() != 1 || gram_prec != le::kFP32 || prec != LE_PREC_FP32)
        return fail(LE_ERR_ARG, "fused_step needs one FP32 S/H Gram job");
    if (G != (8 | kRegSweep) || !step_q) return fail(LE_ERR_STATE, "fused_step engine without its sweep mapping");
    return LE_OK;
}

// Kept samples by fused launches (launch_step): launch q sweeps the walkers and writes sample q
// into buffer q mod 2 while its Gram items accumulate sample q - 1 from the other buffer; a
// plain Gram launch closes the call with the last sample.  Each sample fills rows 0..walkers-1
// of its buffer.
int le_engine::sample_fused(int samples, le::FeatureArgs f, int tw)
{
    int rc = fused_check();
    if (rc) return rc;
    const Job& j = jobs[0];
    le::StepArgs a{};
    a.sw.pos = pos;
    a.sw.accepted = acc;
    a.sw.seed = mc.seed;
    a.sw.N = sys.particles_number;
    a.sw.m = sys.inverse_filling;
    a.sw.walkers = mc.walkers;
    a.sw.nsweeps = mc.thin;
    a.sw.step = step;
    a.units = (mc.walkers + 63) / 64;
    a.queue = step_q;
    a.g.ldx = ldx;
    a.g.nw = 2;
    a.g.tiles = j.tiles;
    a.g.ntiles = j.ntiles;
    a.g.C0 = j.C[0];
    a.g.C1 = j.C[1];
    a.g.ldc = ldx;
    const long rps = mc.fused_rows > 0 ? mc.fused_rows : 2048;
    a.g.rows_per_split = rps;
    a.g.splits = (int)((mc.walkers + rps - 1) / rps);
    f.row0 = 0;
    evused = 0;
    hipEvent_t e0 = next_event(), e1 = next_event();
    if (!e0 || !e1) return fail(LE_ERR_HIP, "event creation failed");
    HIPCHK(hipEventRecord(e0, stream));
    for (int q = 0; q < samples; ++q) {
        const int b = q & 1;
        a.sw.sweep0 = sweep_counter;
        a.f = f;
        a.f.X = Xb[b];
        a.f.W = Wbb[b];
        if (trace_d && trace_rows + tw <= trace_cap_rows) {
            a.f.trace = trace_d;
            a.f.trace_walkers = tw;
            a.f.trace_row0 = trace_rows;
            trace_rows += tw;
        } else {
            a.f.trace = nullptr;
            a.f.trace_walkers = 0;
        }
        a.g.rows = q > 0 ? mc.walkers : 0;
        a.g.X = Xb[1 - b];
        a.g.w0 = weight_ptr(j.wid[0], 1 - b);
        a.g.w1 = weight_ptr(j.wid[1], 1 - b);
        HIPCHK(le::launch_step(a, 8, stream));
        sweep_counter += (unsigned long long)mc.thin;
        st.step_launches += 1;
        st.sweep_launches += 1;
        if (q > 0) st.gram_launches += 1;
    }
    rc = gram_rows(mc.walkers, (samples - 1) & 1, stream);
    if (rc) return rc;
    st.gram_launches += 1;
    ++block_grams;
    HIPCHK(hipEventRecord(e1, stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    st.ms_step += ms;
    sweeps_done += (long long)samples * mc.thin;
    const long long nsamp = (long long)samples * mc.walkers;
    HIPCHK(hipMemcpyAsync(acc_h, acc, sizeof(unsigned) * mc.walkers, hipMemcpyDeviceToHost, stream));
    rc = close_block(nsamp);
    if (rc) return rc;
    double tot = 0;
    for (int w = 0; w < mc.walkers; ++w) tot += acc_h[w];
    acc_sampling += (unsigned long long)tot;
    acc_moves += (long long)samples * mc.thin * npos() * mc.walkers;
    st.acceptance = (double)acc_sampling / (double)std::max<long long>(1, acc_moves);
    st.samples += nsamp;
    st.sweeps = sweeps_done;
    st.step = step;
    return LE_OK;
}

extern "C" int le_sample(le_engine* e, int samples)
{
    if (!e) return fail(LE_ERR_ARG, "null engine");
    if (samples <= 0) return fail(LE_ERR_ARG, "samples must be > 0");
    if (!e->walkers_ready) {
        int rc = le_init_walkers(e);
        if (rc) return rc;
    }
    if (e->rowwise) return e->sample_rowwise(samples);
    if (e->mixture && e->mixp.kmix == 0) {
        int rc = e->calibrate_mixture();
        if (rc) return rc;
    }
    const auto t0 = std::chrono::steady_clock::now();
    HIPCHK(hipMemsetAsync(e->acc, 0, sizeof(unsigned) * e->mc.walkers, e->stream));
    le::FeatureArgs f{};
    f.ps = e->ps_d;
    f.ps_ld = kPsLd;
    f.pos = e->pos;
    f.pos_is_double = 0;
    f.X = e->X;
    f.ldx = e->ldx;
    f.W = e->Wb;
    f.wld = e->wrows;
    f.nh = e->nh;
    for (int h = 0; h < e->nh; ++h) f.harm[h] = e->harm[h];
    f.N = e->sys.particles_number;
    f.walkers = e->mc.walkers;
    f.D = e->basis.D;
    f.K = e->basis.K;
    f.fac = e->fac;
    f.col_re = e->col_re;
    f.col_dim = e->col_dim;
    f.prod = e->prod_d;
    f.pw_k = e->pwk_d;
    f.pw_r = e->pwr_d;
    f.npw = e->npw;
    f.invR = 1.0 / e->R;
    f.pot = {e->sys.conf_w, e->sys.conf_eps, e->sys.conf_strength, e->sys.conf_exponent, e->R};
    e->mix_features(f, true);
    if (e->mixture) {
        // the first sweeps start from the walkers' p_k; every later one from its features
        le::SweepArgs a{};
        a.pos = e->pos;
        a.N = e->sys.particles_number;
        a.walkers = e->mc.walkers;
        a.R = e->R;
        a.mix = e->mixp;
        a.mix_pk = e->mix_pk;
        HIPCHK(le::launch_mix_init(a, e->prec, e->stream));
    }
    const int tw = std::min(e->mc.trace_walkers, e->mc.walkers);
    if (e->mc.fused_step) {
        int rc = e->sample_fused(samples, f, tw);
        if (rc) return rc;
        const auto t1 = std::chrono::steady_clock::now();
        e->st.ms_sample_total += std::chrono::duration<double, std::milli>(t1 - t0).count();
        return LE_OK;
    }
    long long kept = 0;
    // Rounds alternate between two sample buffers: the sampler stream fills buffer b while
    // the Gram stream accumulates the other one; events order the reuse of each buffer.
    struct Span { hipEvent_t a, b; int kind; bool full = false; };   // 0 sweep, 1 features, 2 gram
    std::vector<Span> spans;
    e->evused = 0;
    auto ev = [&]() {
        hipEvent_t q = e->next_event();
        return q;
    };
    int round = 0;
    for (int done = 0; done < samples; ++round) {
        const int n = std::min(e->spl, samples - done);
        const int b = round & 1;
        HIPCHK(hipStreamWaitEvent(e->stream, e->drained[b], 0));
        f.X = e->Xb[b];
        f.W = e->Wbb[b];
        const bool traced_all = e->trace_d && e->trace_rows + (long)n * tw <= e->trace_cap_rows;
        const bool trace_full = !e->trace_d || e->trace_rows + tw > e->trace_cap_rows;
        const bool use_graph = e->mc.graphs && !e->two_sided && !(e->wfn && e->n_user) && n == e->spl &&
                               (traced_all || trace_full);
        if (use_graph) {
            hipGraphExec_t gx = nullptr;
            int rc = e->sample_graph(f, b, traced_all, tw, &gx);
            if (rc) return rc;
            hipEvent_t e0 = ev(), e1 = ev();
            if (!e0 || !e1) return fail(LE_ERR_HIP, "event creation failed");
            HIPCHK(hipEventRecord(e0, e->stream));
            HIPCHK(le::launch_set_bases(e->bases_d, e->sweep_counter, traced_all ? e->trace_rows : 0, e->stream));
            HIPCHK(hipGraphLaunch(gx, e->stream));
            HIPCHK(hipEventRecord(e1, e->stream));
            spans.push_back({e0, e1, 3});
            e->sweep_counter += (unsigned long long)n * (unsigned long long)e->mc.thin;
            if (traced_all) e->trace_rows += (long)n * tw;
        }
        for (int q = 0; q < n && !use_graph; ++q) {
            hipEvent_t e0 = ev(), e1 = ev(), e2 = ev();
            if (!e0 || !e1 || !e2) return fail(LE_ERR_HIP, "event creation failed");
            HIPCHK(hipEventRecord(e0, e->stream));
            int rc = e->run_sweeps(e->mc.thin, e->sweep_counter, e->pos, nullptr, e->mixture);
            if (rc) return rc;
            e->sweep_counter += (unsigned long long)e->mc.thin;
            rc = e->symmetry(e->sweep_counter, nullptr);
            if (rc) return rc;
            HIPCHK(hipEventRecord(e1, e->stream));
            f.row0 = (long)q * e->mc.walkers;
            if (e->trace_d && e->trace_rows + tw <= e->trace_cap_rows) {
                f.trace = e->trace_d;
                f.trace_walkers = tw;
                f.trace_row0 = e->trace_rows;
                e->trace_rows += tw;
            } else {
                f.trace = nullptr;
                f.trace_walkers = 0;
            }
            if (e->two_sided) {
                e->X = e->Xb[b];
                e->Wb = e->Wbb[b];
                HIPCHK(e->features2(e->pos, 0, e->mc.walkers, f.row0));
                e->X = e->Xb[0];
                e->Wb = e->Wbb[0];
            } else {
                HIPCHK(le::launch_features(f, e->prec, e->stream));
            }
            if (e->wfn && e->n_user && !e->two_sided) {
                void* wdst = (char*)e->Wbb[b] + ((size_t)(1 + 2 * e->nh) * e->wrows + f.row0) * e->rsz;
                if (e->wfn(e->wctx, e->pos, e->prec == LE_PREC_FP64, e->mc.walkers, e->sys.particles_number, wdst,
                           e->wrows, e->prec == LE_PREC_FP64, (void*)e->stream) != 0)
                    return fail(LE_ERR_STATE, "user weight callback failed");
            }
            HIPCHK(hipEventRecord(e2, e->stream));
            spans.push_back({e0, e1, 0});
            spans.push_back({e1, e2, 1});
        }
        HIPCHK(hipEventRecord(e->filled[b], e->stream));
        HIPCHK(hipStreamWaitEvent(e->gstream, e->filled[b], 0));
        if (e->need_spread_tune && n == e->spl && !e->weights_gram && e->block_grams == 0) {
            int rc = e->tune_gram_spread(b);
            if (rc) return rc;
        }
        hipEvent_t g0 = ev(), g1 = ev();
        if (!g0 || !g1) return fail(LE_ERR_HIP, "event creation failed");
        HIPCHK(hipEventRecord(g0, e->gstream));
        int rc = e->gram_rows((long)n * e->mc.walkers, b, e->gstream);
        if (rc) return rc;
        ++e->block_grams;
        HIPCHK(hipEventRecord(g1, e->gstream));
        HIPCHK(hipEventRecord(e->drained[b], e->gstream));
        spans.push_back({g0, g1, 2, n == e->spl});
        e->st.gram_launches += 1;
        e->st.sweep_launches += n;
        done += n;
        kept += n;
    }
    HIPCHK(hipStreamSynchronize(e->gstream));
    HIPCHK(hipStreamSynchronize(e->stream));
    for (auto& sp : spans) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, sp.a, sp.b));
        if (sp.kind == 3) e->st.ms_sampler_graph += ms;
        else if (sp.kind == 0) e->st.ms_sweep += ms;
        else if (sp.kind == 1) e->st.ms_features += ms;
        else {
            e->st.ms_gram += ms;
            if (sp.full && (e->st.ms_gram_launch_min == 0.0 || ms < e->st.ms_gram_launch_min))
                e->st.ms_gram_launch_min = ms;
        }
    }
    e->sweeps_done += (long long)samples * e->mc.thin;
    const long long nsamp = kept * e->mc.walkers;
    // the accepted-move counts ride along with the block record (close_block synchronises)
    HIPCHK(hipMemcpyAsync(e->acc_h, e->acc, sizeof(unsigned) * e->mc.walkers, hipMemcpyDeviceToHost, e->stream));
    int rc = e->close_block(nsamp);
    if (rc) return rc;
    double tot = 0;
    for (int w = 0; w < e->mc.walkers; ++w) tot += e->acc_h[w];
    e->acc_sampling += (unsigned long long)tot;
    e->acc_moves += (long long)samples * e->mc.thin * e->npos() * e->mc.walkers;
    e->st.acceptance = (double)e->acc_sampling / (double)std::max<long long>(1, e->acc_moves);
    e->st.samples += nsamp;
    e->st.sweeps = e->sweeps_done;
    e->st.step = e->step;
    e->st.ms_sample_total +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return LE_OK;
}

extern "C" int le_run(le_engine* e)
{
    if (!e) return fail(LE_ERR_ARG, "null engine");
    int rc = le_init_walkers(e);
    if (rc) return rc;
    if (e->mc.tune_sweeps > 0) {
        rc = le_tune_step(e, nullptr);
        if (rc) return rc;
    }
    rc = le_burnin(e, e->mc.burnin_sweeps);
    if (rc) return rc;
    const int nb = e->mc.n_blocks;
    if (e->mc.samples < nb) return fail(LE_ERR_ARG, "samples must be >= n_blocks");
    const int per = e->mc.samples / nb;
    for (int b = 0; b < nb; ++b) {
        rc = le_sample(e, per);
        if (rc) return rc;
    }
    return LE_OK;
}

extern "C" int le_get_stats(const le_engine* e, le_run_stats* st)
{
    if (!e || !st) return fail(LE_ERR_ARG, "null argument");
    *st = e->st;
    return LE_OK;
}

extern "C" int le_n_blocks_done(const le_engine* e) { return e ? (int)e->rec.size() : 0; }

extern "C" int le_set_weight_callback(le_engine* e, le_weight_fn fn, void* ctx)
{
    if (!e) return fail(LE_ERR_ARG, "null engine");
    e->wfn = fn;
    e->wctx = ctx;
    return LE_OK;
}

// copy `n` complex elements at `off` of a record (block >= 0) or the sample-weighted mean
static int record_slice(const le_engine* e, int block, long off, long n, double* out)
{
    const int nb = (int)e->rec.size();
    if (nb == 0) return fail(LE_ERR_STATE, "no closed sample blocks");
    if (block >= nb || block < -1) return fail(LE_ERR_ARG, "block out of range");
    if (block >= 0) {
        std::memcpy(out, e->rec[block] + 2 * off, sizeof(double) * 2 * n);
        return LE_OK;
    }
    // blocks weigh by their samples (mixture: by their sums of 1 / w)
    double tot = 0;
    for (auto c : e->blk_norm) tot += c;
    for (long q = 0; q < 2 * n; ++q) {
        double s = 0;
        for (int b = 0; b < nb; ++b) {
            const double w = tot > 0.0 ? e->blk_norm[b] / tot : 1.0 / nb;
            s += w * e->rec[b][2 * off + q];
        }
        out[q] = s;
    }
    return LE_OK;
}

// Row-wise estimator: normalised, symmetrised M and H (em_spectrum.cu:318-335, 382-390) from the
// raw sums of one block or of all blocks (sums add; each row's sample count is blk_samples).
static int rowwise_moments(const le_engine* e, int block, double* S, double* H)
{
    const int nb = (int)e->rec.size();
    if (nb == 0) return fail(LE_ERR_STATE, "no closed sample blocks");
    if (block >= nb || block < -1) return fail(LE_ERR_ARG, "block out of range");
    const long be = e->basis.block_elems();
    std::vector<double> sum(5 * be, 0.0);
    double n = 0;
    for (int b = 0; b < nb; ++b) {
        if (block >= 0 && b != block) continue;
        for (long q = 0; q < 5 * be; ++q) sum[q] += e->rec[b][q];
        n += (double)e->blk_samples[b];
    }
    for (int s = 0; s < e->basis.nsec(); ++s) {
        const int d = e->basis.dim(s);
        const long bo = e->basis.block_offset(s);
        // X(f, g) = <x g/f>_f / sqrt(<|g/f|^2>_f), then (X + X^H) / 2
        auto norm_entry = [&](int q0, int f, int g, double& re, double& im) {
            const long ix = bo + (long)f * d + g;
            const double nrm = std::sqrt(sum[4 * be + ix] / n);
            re = sum[q0 * be + ix] / n / nrm;
            im = sum[(q0 + 1) * be + ix] / n / nrm;
        };
        for (int a = 0; a < d; ++a)
            for (int c = 0; c < d; ++c) {
                double ar, ai, cr, ci;
                const long out = 2 * (bo + (long)a * d + c);
                if (S) {
                    norm_entry(0, a, c, ar, ai);
                    norm_entry(0, c, a, cr, ci);
                    S[out] = 0.5 * (ar + cr);
                    S[out + 1] = 0.5 * (ai - ci);
                }
                if (H) {
                    norm_entry(2, a, c, ar, ai);
                    norm_entry(2, c, a, cr, ci);
                    H[out] = 0.5 * (ar + cr);
                    H[out + 1] = 0.5 * (ai - ci);
                }
            }
    }
    return LE_OK;
}

extern "C" int le_get_moments(const le_engine* e, int block, double* S, double* H)
{
    if (!e) return fail(LE_ERR_ARG, "null engine");
    if (e->rowwise) return rowwise_moments(e, block, S, H);
    long o = 0;
    for (int s = 0; s < e->basis.nsec(); ++s) {
        const long n = (long)e->basis.dim(s) * e->basis.dim(s);
        if (n) {
            if (S) {
                int rc = record_slice(e, block, e->item_offset(LE_W_ONE, s, s), n, S + 2 * o);
                if (rc) return rc;
            }
            if (H) {
                int rc = record_slice(e, block, e->item_offset(LE_W_V, s, s), n, H + 2 * o);
                if (rc) return rc;
            }
        }
        o += n;
    }
    return LE_OK;
}

extern "C" int le_get_block(const le_engine* e, int block, int weight, int s1, int s2, double* out)
{
    if (e && e->rowwise) {
        // the row-wise estimator keeps the normalised M (weight 1) and H (weight V) of each sector
        if (!out || s1 != s2 || s1 < 0 || s1 >= e->basis.nsec() || (weight != LE_W_ONE && weight != LE_W_V))
            return fail(LE_ERR_ARG, "the row-wise estimator keeps only each sector's M and H blocks");
        const long be = e->basis.block_elems();
        std::vector<double> S(2 * be), H(2 * be);
        const int rc = rowwise_moments(e, block, S.data(), H.data());
        if (rc) return rc;
        const long off = 2 * e->basis.block_offset(s1), n = 2L * e->basis.dim(s1) * e->basis.dim(s1);
        std::memcpy(out, (weight == LE_W_ONE ? S.data() : H.data()) + off, sizeof(double) * n);
        return LE_OK;
    }
    if (!e || !out) return fail(LE_ERR_ARG, "null argument");
    const int ns = e->basis.nsec();
    if (s1 < 0 || s2 < 0 || s1 >= ns || s2 >= ns) return fail(LE_ERR_ARG, "sector out of range");
    const int a = std::min(s1, s2), b = std::max(s1, s2);
    const long off = e->item_offset(weight, a, b);
    if (off < 0) return fail(LE_ERR_ARG, "this weight was not accumulated between these sectors");
    const int da = e->basis.dim(a), db = e->basis.dim(b);
    if (s1 <= s2) return record_slice(e, block, off, (long)da * db, out);
    std::vector<double> tmp(2 * (size_t)da * db);
    int rc = record_slice(e, block, off, (long)da * db, tmp.data());
    if (rc) return rc;
    for (int i = 0; i < db; ++i)       // out is db x da: conj transpose
        for (int j = 0; j < da; ++j) {
            out[2 * ((size_t)i * da + j)] = tmp[2 * ((size_t)j * db + i)];
            out[2 * ((size_t)i * da + j) + 1] = -tmp[2 * ((size_t)j * db + i) + 1];
        }
    return LE_OK;
}

extern "C" int le_get_trace(const le_engine* e, float* out, long capacity, long* n_rows)
{
    if (!e) return fail(LE_ERR_ARG, "null engine");
    if (n_rows) *n_rows = e->trace_rows;
    if (!out) return LE_OK;
    const long n = std::min(capacity, e->trace_rows);
    if (n > 0) HIPCHK(hipMemcpy(out, e->trace_d, sizeof(float) * 9 * n, hipMemcpyDeviceToHost));
    return LE_OK;
}

extern "C" int le_get_positions(le_engine* e, double* out)
{
    if (!e || !out) return fail(LE_ERR_ARG, "null argument");
    const size_t n = (size_t)e->mc.walkers * e->npos() * 2;
    if (e->prec == LE_PREC_FP64) {
        HIPCHK(hipMemcpy(out, e->pos, n * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        std::vector<float> tmp(n);
        HIPCHK(hipMemcpy(tmp.data(), e->pos, n * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t q = 0; q < n; ++q) out[q] = tmp[q];
    }
    return LE_OK;
}

extern "C" int le_set_positions(le_engine* e, const double* in)
{
    if (!e || !in) return fail(LE_ERR_ARG, "null argument");
    const size_t n = (size_t)e->mc.walkers * e->npos() * 2;
    if (e->prec == LE_PREC_FP64) {
        HIPCHK(hipMemcpy(e->pos, in, n * sizeof(double), hipMemcpyHostToDevice));
    } else {
        std::vector<float> tmp(n);
        for (size_t q = 0; q < n; ++q) tmp[q] = (float)in[q];
        HIPCHK(hipMemcpy(e->pos, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice));
    }
    if (e->rowwise) {
        // the chains' running p_k of their row states follow the new positions
        le::SweepArgs a{};
        a.pos = e->pos;
        a.N = e->sys.particles_number;
        a.walkers = e->mc.walkers;
        a.row_walkers = e->row_W;
        a.row_fac = e->fac;
        a.row_pk = e->row_pk;
        a.R = e->R;
        HIPCHK(le::launch_row_init(a, e->prec, e->stream));
        HIPCHK(hipStreamSynchronize(e->stream));
    }
    e->walkers_ready = true;
    return LE_OK;
}

extern "C" int le_reset_moments(le_engine* e)
{
    if (!e) return fail(LE_ERR_ARG, "null engine");
    e->rec.clear();
    e->blk_samples.clear();
    e->blk_norm.clear();
    if (e->mixture) HIPCHK(le::launch_zero(e->mix_acc, e->wrows, e->stream));
    e->block_grams = 0;
    if (e->rowwise) HIPCHK(le::launch_zero(e->row_acc, 5 * e->basis.block_elems(), e->stream));
    HIPCHK(le::launch_zero(e->cpool, (long)e->cpool_elems, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return LE_OK;
}

extern "C" int le_set_thin(le_engine* e, int thin)
{
    if (!e) return fail(LE_ERR_ARG, "null engine");
    if (thin < 1) return fail(LE_ERR_ARG, "thin must be >= 1");
    e->mc.thin = thin;
    return LE_OK;
}

extern "C" int le_accumulate_positions(le_engine* e, const double* pos, long n)
{
    if (!e || !pos || n <= 0) return fail(LE_ERR_ARG, "null argument or n <= 0");
    if (e->rowwise) return fail(LE_ERR_STATE, "fixed samples do not apply to the row-wise estimator's chains");
    if (e->mixture && e->mixp.kmix == 0)
        return fail(LE_ERR_STATE, "fixed samples of a mixture engine need its density: call le_set_mixture first");
    const int N = e->npos();
    const long chunk = (long)e->spl * e->mc.walkers;
    double* stage = nullptr;
    HIPCHK(dalloc(&stage, (size_t)std::min(chunk, n) * N * 2));
    le::FeatureArgs f{};
    f.ps = e->ps_d;
    f.ps_ld = kPsLd;
    f.pos = stage;
    f.pos_is_double = 1;
    f.X = e->X;
    f.ldx = e->ldx;
    f.W = e->Wb;
    f.wld = e->wrows;
    f.nh = e->nh;
    for (int h = 0; h < e->nh; ++h) f.harm[h] = e->harm[h];
    f.row0 = 0;
    f.N = N;
    f.D = e->basis.D;
    f.K = e->basis.K;
    f.fac = e->fac;
    f.col_re = e->col_re;
    f.col_dim = e->col_dim;
    f.prod = e->prod_d;
    f.pw_k = e->pwk_d;
    f.pw_r = e->pwr_d;
    f.npw = e->npw;
    f.invR = 1.0 / e->R;
    f.pot = {e->sys.conf_w, e->sys.conf_eps, e->sys.conf_strength, e->sys.conf_exponent, e->R};
    e->mix_features(f, false);
    int rc = LE_OK;
    for (long s0 = 0; s0 < n && rc == LE_OK; s0 += chunk) {
        const long m = std::min(chunk, n - s0);
        hipError_t h = hipMemcpyAsync(stage, pos + (size_t)s0 * N * 2, sizeof(double) * m * N * 2,
                                      hipMemcpyHostToDevice, e->stream);
        if (h != hipSuccess) { rc = fail(LE_ERR_HIP, hipGetErrorString(h)); break; }
        f.walkers = (int)m;
        h = e->two_sided ? e->features2(stage, 1, (int)m, 0) : le::launch_features(f, e->prec, e->stream);
        if (h != hipSuccess) { rc = fail(LE_ERR_HIP, hipGetErrorString(h)); break; }
        if (e->wfn && e->n_user && !e->two_sided) {
            void* wdst = (char*)e->Wb + (size_t)(1 + 2 * e->nh) * e->wrows * e->rsz;
            if (e->wfn(e->wctx, stage, 1, (int)m, N, wdst, e->wrows, e->prec == LE_PREC_FP64, (void*)e->stream) != 0) {
                rc = fail(LE_ERR_STATE, "user weight callback failed");
                break;
            }
        }
        e->evused = 0;
        hipEvent_t g0 = e->next_event(), g1 = e->next_event();
        if (!g0 || !g1) { rc = fail(LE_ERR_HIP, "event creation failed"); break; }
        h = hipEventRecord(g0, e->stream);
        if (h != hipSuccess) { rc = fail(LE_ERR_HIP, hipGetErrorString(h)); break; }
        rc = e->gram_rows(m);
        ++e->block_grams;
        if (rc == LE_OK) {
            h = hipEventRecord(g1, e->stream);
            if (h == hipSuccess) h = hipStreamSynchronize(e->stream);
            float ms = 0.f;
            if (h == hipSuccess) h = hipEventElapsedTime(&ms, g0, g1);
            if (h != hipSuccess) rc = fail(LE_ERR_HIP, hipGetErrorString(h));
            e->st.ms_gram += ms;          // the Gram launches alone: nothing else runs meanwhile
            e->st.gram_launches += 1;
        }
    }
    (void)hipFree(stage);
    if (rc) return rc;
    return e->close_block(n);
}

// ---------------------------------------------------------------------------
// FP64 generalized eigensolves on the GPU
// ---------------------------------------------------------------------------
static int geneig_batch(const std::vector<int>& dims, const double* S, const double* H, double* evals,
                        double* evecs, std::vector<int>* status_out)
{
    const int np = (int)dims.size();
    std::vector<le::GenEigProblem> probs(np);
    long in_off = 0, ev_off = 0, w_off = 0;
    for (int p = 0; p < np; ++p) {
        probs[p] = {dims[p], in_off, ev_off, w_off};
        in_off += (long)dims[p] * dims[p];
        ev_off += dims[p];
        w_off += 8L * dims[p] * dims[p] + 4L * dims[p];
    }
    double *dS = nullptr, *dH = nullptr, *dW = nullptr, *dE = nullptr, *dV = nullptr;
    le::GenEigProblem* dP = nullptr;
    int* dSt = nullptr;
    hipError_t h = hipSuccess;
    auto chk = [&](hipError_t x) { if (h == hipSuccess) h = x; };
    chk(dalloc(&dS, (size_t)std::max<long>(2, 2 * in_off)));
    chk(dalloc(&dH, (size_t)std::max<long>(2, 2 * in_off)));
    chk(dalloc(&dW, (size_t)std::max<long>(2, 2 * w_off)));
    chk(dalloc(&dE, (size_t)std::max<long>(1, ev_off)));
    if (evecs) chk(dalloc(&dV, (size_t)std::max<long>(2, 2 * in_off)));
    chk(dalloc(&dP, (size_t)std::max(1, np)));
    chk(dalloc(&dSt, (size_t)std::max(1, np)));
    if (h == hipSuccess) {
        chk(hipMemcpy(dS, S, sizeof(double) * 2 * in_off, hipMemcpyHostToDevice));
        chk(hipMemcpy(dH, H, sizeof(double) * 2 * in_off, hipMemcpyHostToDevice));
        chk(hipMemcpy(dP, probs.data(), sizeof(le::GenEigProblem) * np, hipMemcpyHostToDevice));
        chk(hipMemset(dSt, 0xff, sizeof(int) * np));
    }
    const int dmax = dims.empty() ? 1 : *std::max_element(dims.begin(), dims.end());
    if (h == hipSuccess) chk(le::launch_geneig(dS, dH, dW, dP, np, dmax, dE, dV, dSt, nullptr));
    std::vector<int> st(np, -1);
    if (h == hipSuccess) chk(hipDeviceSynchronize());
    if (h == hipSuccess) {
        chk(hipMemcpy(evals, dE, sizeof(double) * ev_off, hipMemcpyDeviceToHost));
        if (evecs) chk(hipMemcpy(evecs, dV, sizeof(double) * 2 * in_off, hipMemcpyDeviceToHost));
        chk(hipMemcpy(st.data(), dSt, sizeof(int) * np, hipMemcpyDeviceToHost));
    }
    (void)hipFree(dS);
    (void)hipFree(dH);
    (void)hipFree(dW);
    (void)hipFree(dE);
    (void)hipFree(dV);
    (void)hipFree(dP);
    (void)hipFree(dSt);
    if (h != hipSuccess) return fail(LE_ERR_HIP, hipGetErrorString(h));
    if (status_out) *status_out = st;
    for (int p = 0; p < np; ++p)
        if (st[p] == 1) return fail(LE_ERR_NOT_PD, "metric is not positive definite (problem " + std::to_string(p) + ")");
    return LE_OK;
}

extern "C" int le_geneig(int d, const double* S, const double* H, double* evals, double* evecs)
{
    if (d <= 0 || d > 640 || !S || !H || !evals) return fail(LE_ERR_ARG, "le_geneig: bad argument");
    return geneig_batch({d}, S, H, evals, evecs, nullptr);
}

extern "C" int le_engine_spectrum(le_engine* e, int sector, int block, double* evals)
{
    if (!e || !evals || sector < 0 || sector >= e->basis.nsec()) return fail(LE_ERR_ARG, "bad argument");
    const long be2 = 2 * e->basis.block_elems();
    std::vector<double> S(be2), H(be2);
    int rc = le_get_moments(e, block, S.data(), H.data());
    if (rc) return rc;
    const long off = 2 * e->basis.block_offset(sector);
    return geneig_batch({e->basis.dim(sector)}, S.data() + off, H.data() + off, evals, nullptr, nullptr);
}

extern "C" int le_engine_spectra(le_engine* e, int block, double* evals, int* status)
{
    if (!e || !evals) return fail(LE_ERR_ARG, "bad argument");
    const long be2 = 2 * e->basis.block_elems();
    std::vector<double> S(be2), H(be2);
    int rc = le_get_moments(e, block, S.data(), H.data());
    if (rc) return rc;
    std::vector<int> dims;
    for (int s = 0; s < e->basis.nsec(); ++s) dims.push_back(e->basis.dim(s));
    std::vector<int> st;
    rc = geneig_batch(dims, S.data(), H.data(), evals, nullptr, &st);
    if (status)
        for (size_t q = 0; q < st.size(); ++q) status[q] = st[q];
    if (rc == LE_ERR_NOT_PD && status) return LE_OK;   // per-sector status reported
    return rc;
}

// Particle-removal spectral weights of the Laughlin state.  Psi_L^{N+1}(z_1..z_N, z_0) =
// Psi_L^N(z_1..z_N) prod_j (z_j - z_0)^m, and the z_0^l coefficient of prod_j (z_j - z_0)^m is
// (-1)^l e_k over the multiset {z_j, each m times}, k = m N - l.  So c_l Psi_L^{N+1} is
// proportional to e^(m)_k Psi_L^N, a state of this engine's sector dL = k; by Newton's
// identities e^(m)_k = sum_{lambda |- k} (-1)^(k - len) m^len / z_lambda p_lambda, with
// z_lambda = prod_i i^(m_i) m_i!.  Its weight on the S-normalized eigenvector v_j is
// |v_j^H S a|^2 / (a^H S a).
extern "C" int le_removal_weights(le_engine* e, int sector, int block, int* orbital, double* energies,
                                  double* weights)
{
    if (!e || !energies || !weights || sector < 0 || sector >= e->basis.nsec())
        return fail(LE_ERR_ARG, "le_removal_weights: bad argument");
    if (e->rowwise) return fail(LE_ERR_ARG, "le_removal_weights needs the plain estimator's unnormalised S");
    if (e->two_sided) return fail(LE_ERR_STATE, "le_removal_weights: needs an edge engine, not the occupation chain");
    const int N = e->sys.particles_number, m = e->sys.inverse_filling;
    const int k = e->basis.sectors[sector];
    if (k > N) return fail(LE_ERR_ARG, "le_removal_weights: needs dL <= particles_number");
    const int d = e->basis.dim(sector);
    const long be2 = 2 * e->basis.block_elems();
    std::vector<double> S(be2), H(be2);
    int rc = le_get_moments(e, block, S.data(), H.data());
    if (rc) return rc;
    const long off = 2 * e->basis.block_offset(sector);
    std::vector<double> Ss(2 * (size_t)d * d), Hs(2 * (size_t)d * d);
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < d; ++j) {
            const size_t ij = 2 * ((size_t)i * d + j), ji = 2 * ((size_t)j * d + i);
            Ss[ij] = 0.5 * (S[off + ij] + S[off + ji]);
            Ss[ij + 1] = 0.5 * (S[off + ij + 1] - S[off + ji + 1]);
            Hs[ij] = 0.5 * (H[off + ij] + H[off + ji]);
            Hs[ij + 1] = 0.5 * (H[off + ij + 1] - H[off + ji + 1]);
        }
    // Newton coefficients, scaled by their largest magnitude (the weights do not depend on it)
    std::vector<double> lc(d), sg(d), a(d);
    double lmax = -1e300;
    for (int q = 0; q < d; ++q) {
        const auto& st = e->basis.states[e->basis.sector_offset[sector] + q];
        int len = 0;
        double lz = 0.0;
        for (const auto& pp : st.pairs) {
            if (pp.number == 0) continue;   // the dL = 0 state is the constant 1
            len += pp.repetitions;
            lz += pp.repetitions * std::log((double)pp.number) + std::lgamma(pp.repetitions + 1.0);
        }
        lc[q] = len * std::log((double)m) - lz;
        sg[q] = ((k - len) % 2) ? -1.0 : 1.0;
        lmax = std::max(lmax, lc[q]);
    }
    for (int q = 0; q < d; ++q) a[q] = sg[q] * std::exp(lc[q] - lmax);
    std::vector<double> vec(2 * (size_t)d * d);
    rc = geneig_batch({d}, Ss.data(), Hs.data(), energies, vec.data(), nullptr);
    if (rc) return rc;
    // S a and a^T S a (a is real)
    std::vector<double> sa(2 * (size_t)d, 0.0);
    double asa = 0.0;
    for (int i = 0; i < d; ++i) {
        double re = 0, im = 0;
        for (int j = 0; j < d; ++j) {
            re += Ss[2 * ((size_t)i * d + j)] * a[j];
            im += Ss[2 * ((size_t)i * d + j) + 1] * a[j];
        }
        sa[2 * i] = re;
        sa[2 * i + 1] = im;
        asa += a[i] * re;
    }
    if (!(asa > 0)) return fail(LE_ERR_NOT_PD, "le_removal_weights: a^T S a is not positive");
    for (int j = 0; j < d; ++j) {
        double re = 0, im = 0;   // v_j^H (S a)
        for (int i = 0; i < d; ++i) {
            const double vr = vec[2 * ((size_t)i * d + j)], vi = vec[2 * ((size_t)i * d + j) + 1];
            re += vr * sa[2 * i] + vi * sa[2 * i + 1];
            im += vr * sa[2 * i + 1] - vi * sa[2 * i];
        }
        weights[j] = (re * re + im * im) / asa;
    }
    if (orbital) *orbital = m * N - k;
    return LE_OK;
}

// ---------------------------------------------------------------------------
// FP64 time evolution (interaction picture in the generalized eigenbasis)
// ---------------------------------------------------------------------------
extern "C" int le_evolve(int D, int n_blocks, const int* block_dims, const double* M, const double* H, int n_u,
                         const double* U, int n_rho, const double* rho, le_time_fn F, void* ctx, double end_time,
                         double save_step, double eps_abs, double eps_rel, double* out_rho, double* out_norm,
                         long* n_steps)
{
    if (D <= 0 || n_blocks <= 0 || !block_dims || !M || !H || n_u <= 0 || !U || n_rho < 0 || (n_rho && !rho) || !F ||
        !(save_step > 0) || !(end_time > 0) || !out_rho)
        return fail(LE_ERR_ARG, "le_evolve: bad argument");
    long dsum = 0;
    for (int b = 0; b < n_blocks; ++b) {
        if (block_dims[b] <= 0 || block_dims[b] > 640) return fail(LE_ERR_ARG, "le_evolve: block size out of range");
        dsum += block_dims[b];
    }
    if (dsum != D) return fail(LE_ERR_ARG, "le_evolve: block sizes do not add up to D");
    const size_t DD = (size_t)D * D;
    // 1. per-block generalized eigenproblems (FP64, GPU)
    std::vector<double> Sb, Hb;
    std::vector<int> dims(block_dims, block_dims + n_blocks), offs(n_blocks + 1, 0);
    for (int b = 0; b < n_blocks; ++b) offs[b + 1] = offs[b] + dims[b];
    for (int b = 0; b < n_blocks; ++b)
        for (int i = 0; i < dims[b]; ++i)
            for (int j = 0; j < dims[b]; ++j) {
                const size_t q = (size_t)(offs[b] + i) * D + offs[b] + j;
                Sb.push_back(M[2 * q]);
                Sb.push_back(M[2 * q + 1]);
                Hb.push_back(H[2 * q]);
                Hb.push_back(H[2 * q + 1]);
            }
    std::vector<double> lam(D), vecs(Sb.size());
    int rc = geneig_batch(dims, Sb.data(), Hb.data(), lam.data(), vecs.data(), nullptr);
    if (rc) return rc;
    std::vector<double> V(2 * DD, 0.0);     // block-diagonal, column j = eigenvector j
    {
        size_t o = 0;
        for (int b = 0; b < n_blocks; ++b) {
            const int d = dims[b];
            for (int i = 0; i < d; ++i)
                for (int j = 0; j < d; ++j) {
                    const size_t q = (size_t)(offs[b] + i) * D + offs[b] + j;
                    V[2 * q] = vecs[2 * (o + (size_t)i * d + j)];
                    V[2 * q + 1] = vecs[2 * (o + (size_t)i * d + j) + 1];
                }
            o += (size_t)d * d;
        }
    }
    // a(0) = V^H M e_0
    std::vector<double> a0(2 * (size_t)D, 0.0);
    for (int j = 0; j < D; ++j) {
        double sr = 0, si = 0;
        for (int k = 0; k < D; ++k) {
            const double vr = V[2 * ((size_t)k * D + j)], vi = -V[2 * ((size_t)k * D + j) + 1];
            const double mr = M[2 * ((size_t)k * D)], mi = M[2 * ((size_t)k * D) + 1];
            sr += vr * mr - vi * mi;
            si += vr * mi + vi * mr;
        }
        a0[2 * j] = sr;
        a0[2 * j + 1] = si;
    }
    hipStream_t st = nullptr;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    const long nt = (long)std::ceil(end_time / save_step);
    double *dV = nullptr, *dT = nullptr, *dX = nullptr, *dW = nullptr, *dR = nullptr, *dlam = nullptr, *db = nullptr,
           *dbn = nullptr, *dks = nullptr, *dy = nullptr, *dcoef = nullptr, *dout = nullptr, *dA = nullptr;
    unsigned long long* demax = nullptr;
    unsigned long long* hemax = nullptr;   // pinned: the per-step readback of the error maximum
    hipError_t h = hipSuccess;
    auto chk = [&](hipError_t x) { if (h == hipSuccess) h = x; };
    chk(dalloc(&dV, 2 * DD));
    chk(dalloc(&dT, 2 * DD));
    chk(dalloc(&dX, 2 * DD));
    chk(dalloc(&dW, 2 * DD * n_u));
    chk(dalloc(&dR, 2 * DD * std::max(1, n_rho)));
    chk(dalloc(&dlam, (size_t)D));
    chk(dalloc(&db, 2 * (size_t)D * n_u));
    chk(dalloc(&dbn, 2 * (size_t)D * n_u));
    chk(dalloc(&dks, 12 * (size_t)D * n_u));
    chk(dalloc(&dy, 2 * (size_t)D * n_u));
    chk(dalloc(&dA, 2 * (size_t)D * n_u));
    chk(dalloc(&dcoef, 64));
    chk(dalloc(&dout, 2 * (size_t)nt * n_u * (n_rho + 1)));
    chk(dalloc(&demax, 1));
    chk(hipHostMalloc((void**)&hemax, sizeof(unsigned long long), hipHostMallocDefault));
    // 2. W_u = V^H U_u V, R_l = V^H rho_l V
    if (h == hipSuccess) chk(hipMemcpy(dV, V.data(), sizeof(double) * 2 * DD, hipMemcpyHostToDevice));
    auto sandwich = [&](const double* host, double* dst) {
        chk(hipMemcpyAsync(dX, host, sizeof(double) * 2 * DD, hipMemcpyHostToDevice, st));
        chk(le::launch_zgemm(dX, dV, dT, D, D, D, D, D, D, 0, st));     // T = X V
        chk(le::launch_zgemm(dV, dT, dst, D, D, D, D, D, D, 1, st));    // V^H T
        chk(hipStreamSynchronize(st));
    };
    for (int u = 0; u < n_u && h == hipSuccess; ++u) sandwich(U + 2 * DD * u, dW + 2 * DD * u);
    // the stage products read only the upper tiles of each W_u: use its Hermitian part
    if (h == hipSuccess && n_u > 0) {
        chk(le::launch_hermitian_part(dW, D, n_u, st));
        chk(hipStreamSynchronize(st));
    }
    for (int l = 0; l < n_rho && h == hipSuccess; ++l) sandwich(rho + 2 * DD * l, dR + 2 * DD * l);
    // 3. Cash-Karp 5(4) tableau
    static const double cc[6] = {0.0, 0.2, 0.3, 0.6, 1.0, 0.875};
    static const double aa[6][5] = {{0, 0, 0, 0, 0},
                                    {0.2, 0, 0, 0, 0},
                                    {3.0 / 40, 9.0 / 40, 0, 0, 0},
                                    {0.3, -0.9, 1.2, 0, 0},
                                    {-11.0 / 54, 2.5, -70.0 / 27, 35.0 / 27, 0},
                                    {1631.0 / 55296, 175.0 / 512, 575.0 / 13824, 44275.0 / 110592, 253.0 / 4096}};
    static const double b5[6] = {37.0 / 378, 0, 250.0 / 621, 125.0 / 594, 0, 512.0 / 1771};
    static const double b4[6] = {2825.0 / 27648, 0, 18575.0 / 48384, 13525.0 / 55296, 277.0 / 14336, 0.25};
    std::vector<double> coef(64, 0.0);
    for (int sI = 0; sI < 6; ++sI)
        for (int r = 0; r < 5; ++r) coef[8 * sI + r] = aa[sI][r];
    for (int sI = 0; sI < 6; ++sI) {
        coef[48 + sI] = b5[sI];
        coef[56 + sI] = b5[sI] - b4[sI];
    }
    if (h == hipSuccess) {
        chk(hipMemcpy(dcoef, coef.data(), sizeof(double) * 64, hipMemcpyHostToDevice));
        chk(hipMemcpy(dlam, lam.data(), sizeof(double) * D, hipMemcpyHostToDevice));
        for (int u = 0; u < n_u; ++u)
            chk(hipMemcpy(db + 2 * (size_t)D * u, a0.data(), sizeof(double) * 2 * D, hipMemcpyHostToDevice));
        chk(hipMemset(dout, 0, sizeof(double) * 2 * (size_t)nt * n_u * (n_rho + 1)));
    }
    // 4. adaptive integration, observing at t_k (host drives the controller; F is a host callback)
    const int n = D * n_u;
    double t = 0.0, dt = save_step;
    long steps = 0;
    const long max_steps = 50000000;
    for (long kt = 0; kt < nt && h == hipSuccess; ++kt) {
        const double tk = kt * save_step;
        while (t < tk - 1e-14 * (1.0 + std::fabs(tk)) && h == hipSuccess) {
            double hs = dt;
            bool last = false;
            if (t + hs >= tk) {
                hs = tk - t;
                last = true;
            }
            for (int tries = 0;; ++tries) {
                for (int sI = 0; sI < 6 && h == hipSuccess; ++sI) {
                    const double ts = t + cc[sI] * hs;
                    double* kz = dks + 2 * (size_t)sI * n;
                    chk(le::launch_stage_prep(db, dks, sI, dcoef + 8 * sI, hs, ts, dlam, dy, kz, D, n_u,
                                              sI == 0 ? demax : nullptr, st));
                    chk(le::launch_stage_hemv(dW, dy, kz, F(ctx, ts), ts, dlam, D, n_u, st));
                }
                chk(le::launch_rk_combine(db, dks, dcoef + 48, dcoef + 56, hs, eps_abs, eps_rel, dbn, demax, n, st));
                chk(hipMemcpyAsync(hemax, demax, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
                chk(hipStreamSynchronize(st));
                if (h != hipSuccess) break;
                const unsigned long long bits = *hemax;
                double emax;
                std::memcpy(&emax, &bits, sizeof(emax));
                ++steps;
                if (!(emax <= 1.0)) {
                    if (tries > 200 || steps > max_steps) {
                        h = hipErrorUnknown;
                        break;
                    }
                    const double f = std::isfinite(emax) ? 0.9 * std::pow(emax, -1.0 / 3.0) : 0.2;
                    hs *= std::max(f, 0.2);
                    last = false;
                    continue;
                }
                std::swap(db, dbn);
                t += hs;
                if (last) t = tk;
                if (emax < 0.5) {
                    const double e = std::max(emax, std::pow(5.0, -5.0));
                    hs *= 0.9 * std::pow(e, -1.0 / 5.0);
                }
                if (!last || hs > dt) dt = hs;
                break;
            }
        }
        if (h == hipSuccess)
            chk(le::launch_observe(db, dR, dlam, tk, D, n_u, n_rho, dA, dout + 2 * (size_t)kt * n_u * (n_rho + 1), st));
    }
    std::vector<double> res(2 * (size_t)nt * n_u * (n_rho + 1));
    if (h == hipSuccess) {
        chk(hipMemcpyAsync(res.data(), dout, sizeof(double) * res.size(), hipMemcpyDeviceToHost, st));
        chk(hipStreamSynchronize(st));
    }
    for (double* p : {dV, dT, dX, dW, dR, dlam, db, dbn, dks, dy, dcoef, dout, dA}) (void)hipFree(p);
    (void)hipFree(demax);
    if (hemax) (void)hipHostFree(hemax);
    (void)hipStreamDestroy(st);
    if (h != hipSuccess) return fail(LE_ERR_HIP, std::string("le_evolve: ") + hipGetErrorString(h));
    for (int u = 0; u < n_u; ++u)
        for (long kt = 0; kt < nt; ++kt) {
            const size_t src = ((size_t)kt * n_u + u) * (n_rho + 1);
            for (int l = 0; l < n_rho; ++l) {
                out_rho[2 * (((size_t)u * nt + kt) * n_rho + l)] = res[2 * (src + l)];
                out_rho[2 * (((size_t)u * nt + kt) * n_rho + l) + 1] = res[2 * (src + l) + 1];
            }
            if (out_norm) out_norm[(size_t)u * nt + kt] = res[2 * (src + n_rho)];
        }
    if (n_steps) *n_steps = steps;
    return LE_OK;
}

extern "C" int le_engine_create_occupation(const le_system_params* sys, const le_mc_params* mc, int dL, int lcut,
                                           le_engine** out)
{
    if (!sys || !mc || !out) return fail(LE_ERR_ARG, "null argument");
    *out = nullptr;
    if (lcut < 1 || lcut > 4096) return fail(LE_ERR_ARG, "lcut must be in [1, 4096]");
    if (sys->particles_number < 2 || sys->inverse_filling < 1)
        return fail(LE_ERR_ARG, "need particles_number >= 2 and inverse_filling >= 1");
    if (mc->walkers < 1 || mc->thin < 1 || mc->n_blocks < 1 || mc->samples < 0 || mc->burnin_sweeps < 0)
        return fail(LE_ERR_ARG, "walkers, thin and n_blocks must be >= 1; samples, burn-in >= 0");
    if (mc->precision < LE_PREC_FP32 || mc->precision > LE_PREC_BF16X6)
        return fail(LE_ERR_ARG, "unknown precision");
    if (mc->lanes_per_walker > 0 && !le::sweep_supported(mc->lanes_per_walker))
        return fail(LE_ERR_ARG, "lanes_per_walker must be -1, 0, 2, 4, 8, 16, 32 or 64");
    if (mc->graphs != 0 && mc->graphs != 1) return fail(LE_ERR_ARG, "graphs must be 0 or 1");
    if (mc->symmetry_moves < 0 || mc->symmetry_moves > 2) return fail(LE_ERR_ARG, "symmetry_moves must be 0, 1 or 2");
    if (mc->sweep_registers != 0 && mc->sweep_registers != 1) return fail(LE_ERR_ARG, "sweep_registers must be 0 or 1");
    if (mc->gram_xcd_spread != 0 && mc->gram_xcd_spread != 1 && mc->gram_xcd_spread != 2 && mc->gram_xcd_spread != 4 &&
        mc->gram_xcd_spread != 8)
        return fail(LE_ERR_ARG, "gram_xcd_spread must be 0, 1, 2, 4 or 8");
    if (mc->gram_tile_cols != 0 && mc->gram_tile_cols != 32 && mc->gram_tile_cols != 64)
        return fail(LE_ERR_ARG, "gram_tile_cols must be 0, 32 or 64");
    if (mc->gram_tile_cols == 32 && (mc->precision == LE_PREC_BF16X3 || mc->precision == LE_PREC_BF16X6))
        return fail(LE_ERR_ARG, "the BF16 split Gram runs 64-column tiles only");
    if (mc->sweep_registers == 1 && mc->lanes_per_walker > 0 &&
        !le::sweep2_reg_supported(mc->lanes_per_walker, sys->particles_number))
        return fail(LE_ERR_ARG, "sweep_registers needs lanes_per_walker 8 or 16 with ceil((N - 1) / lanes) 13 or 25");
    if (dL < 0 || dL > 40) return fail(LE_ERR_ARG, "sector dL must be in [0, 40]");
    if (mc->estimator != LE_EST_PLAIN) return fail(LE_ERR_ARG, "the occupation engine takes the plain estimator only");
    auto e = new le_engine();
    e->sys = *sys;
    e->mc = *mc;
    e->prec = mc->precision == LE_PREC_FP64 ? LE_PREC_FP64 : LE_PREC_FP32;
    e->gram_prec = mc->precision == LE_PREC_FP64     ? le::kFP64
                  : mc->precision == LE_PREC_BF16X3 ? le::kBF16X3
                  : mc->precision == LE_PREC_BF16X6 ? le::kBF16X6
                                                    : le::kFP32;
    e->two_sided = true;
    e->lcut = lcut;
    e->n_user = 2 * lcut;
    e->basis.build(sys->particles_number, {dL, dL});
    if (e->basis.K > 24 || e->basis.longest > le::kMaxFactors || e->basis.dim(0) > 640) {
        delete e;
        return fail(LE_ERR_CAPACITY, "sector beyond the engine's capacity (degree 24, 8 factors, 640 states)");
    }
    int rc = e->setup(nullptr);
    if (rc != LE_OK) {
        std::string msg = g_err;
        delete e;
        g_err = msg;
        return rc;
    }
    e->gram_spread = mc->gram_xcd_spread > 0 ? mc->gram_xcd_spread : 8;
    e->gram_tc = mc->gram_tile_cols > 0 ? mc->gram_tile_cols : 64;
    e->G = mc->lanes_per_walker > 0 ? mc->lanes_per_walker : default_lanes(sys->particles_number, e->prec);
    if (mc->lanes_per_walker > 0 && mc->sweep_registers == 1) e->G |= kRegSweep;
    *out = e;
    return LE_OK;
}

extern "C" int le_get_occupation(const le_engine* e, int block, double* out)
{
    if (!e || !out) return fail(LE_ERR_ARG, "null argument");
    if (!e->two_sided) return fail(LE_ERR_STATE, "not an occupation-number engine");
    const int d = e->basis.dim(0);
    const size_t dd = (size_t)d * d;
    std::vector<double> br(2 * dd), bi(2 * dd);
    for (int l = 0; l < e->lcut; ++l) {
        int rc = le_get_block(e, block, LE_W_USER(2 * l), 0, 1, br.data());
        if (rc) return rc;
        rc = le_get_block(e, block, LE_W_USER(2 * l + 1), 0, 1, bi.data());
        if (rc) return rc;
        for (size_t q = 0; q < dd; ++q) {   // O_l = B[Re w_l] + i B[Im w_l]
            out[2 * (q * e->lcut + l)] = br[2 * q] - bi[2 * q + 1];
            out[2 * (q * e->lcut + l) + 1] = br[2 * q + 1] + bi[2 * q];
        }
    }
    return LE_OK;
}

// ---------------------------------------------------------------------------
// Learned proposal
// ---------------------------------------------------------------------------
extern "C" void le_train_params_default(le_train_params* p)
{
    std::memset(p, 0, sizeof(*p));
    p->steps = 2000;
    p->lr = 1e-3;
    p->batch = 4096;
    p->n_configs = 100000;
    p->snapshot_sweeps = 4;
    p->seed = 99;
    p->sigma0 = 0.0;
    p->table_size = 12;
}

static int build_table(le_engine* e, int T)
{
    if (T < 2 || T > 64) return fail(LE_ERR_ARG, "table_size must be in [2, 64]");
    if (e->table_T != T) {
        (void)hipFree(e->table);
        e->table = nullptr;
        HIPCHK(dalloc(&e->table, (size_t)2 * T * T * T));
        e->table_T = T;
    }
    float* lohi = nullptr;
    HIPCHK(dalloc(&lohi, 6));
    float h[6] = {e->table_lo[0], e->table_lo[1], e->table_lo[2], e->table_hi[0], e->table_hi[1], e->table_hi[2]};
    hipError_t x = hipMemcpy(lohi, h, sizeof(h), hipMemcpyHostToDevice);
    if (x == hipSuccess) x = le::launch_mlp_table(e->mlp, T, lohi, lohi + 3, (float)e->sigma0, e->table, e->stream);
    if (x == hipSuccess) x = hipStreamSynchronize(e->stream);
    (void)hipFree(lohi);
    if (x != hipSuccess) return fail(LE_ERR_HIP, hipGetErrorString(x));
    return LE_OK;
}

extern "C" int le_set_proposal(le_engine* e, int kind)
{
    if (!e) return fail(LE_ERR_ARG, "null engine");
    if (kind == LE_PROPOSAL_PLAIN) {
        e->proposal = kind;
        return LE_OK;
    }
    if (kind != LE_PROPOSAL_LEARNED) return fail(LE_ERR_ARG, "unknown proposal");
    if (e->mc.fused_step) return fail(LE_ERR_ARG, "fused_step runs the plain proposal only");
    if (!e->table) return fail(LE_ERR_STATE, "no trained proposal (le_train_proposal or le_set_proposal_params)");
    if (e->prec != LE_PREC_FP32 || e->two_sided || e->rowwise || e->mixture)
        return fail(LE_ERR_ARG, "the learned proposal needs an FP32 edge engine with the plain estimator");
    e->proposal = kind;
    return LE_OK;
}

extern "C" int le_get_proposal_params(const le_engine* e, float* out)
{
    if (!e || !out) return fail(LE_ERR_ARG, "null argument");
    if (!e->mlp) return fail(LE_ERR_STATE, "no proposal parameters");
    HIPCHK(hipMemcpy(out, e->mlp, sizeof(float) * le::kMlpParams, hipMemcpyDeviceToHost));
    return LE_OK;
}

extern "C" int le_set_proposal_params(le_engine* e, const float* in, double sigma0, int table_size)
{
    if (!e || !in || !(sigma0 > 0)) return fail(LE_ERR_ARG, "bad argument");
    if (!e->mlp) HIPCHK(dalloc(&e->mlp, (size_t)le::kMlpParams));
    HIPCHK(hipMemcpy(e->mlp, in, sizeof(float) * le::kMlpParams, hipMemcpyHostToDevice));
    e->sigma0 = sigma0;
    return build_table(e, table_size);
}

extern "C" int le_proposal_loss_grad(le_engine* e, const float* configs, int n_cfg, const int* cfg, const int* part,
                                     const float* noise, int B, double sigma0, double* loss, float* grad)
{
    if (!e || !configs || !cfg || !part || !noise || B <= 0 || n_cfg <= 0 || !(sigma0 > 0))
        return fail(LE_ERR_ARG, "bad argument");
    if (!e->mlp) return fail(LE_ERR_STATE, "no proposal parameters");
    const int N = e->sys.particles_number;
    for (int b = 0; b < B; ++b)
        if (cfg[b] < 0 || cfg[b] >= n_cfg || part[b] < 0 || part[b] >= N) return fail(LE_ERR_ARG, "sample index out of range");
    float *dc = nullptr, *dn = nullptr, *drec = nullptr, *dg = nullptr;
    int *dcfg = nullptr, *dpart = nullptr;
    double* dst = nullptr;
    hipError_t h = hipSuccess;
    auto chk = [&](hipError_t x) { if (h == hipSuccess) h = x; };
    chk(dalloc(&dc, (size_t)n_cfg * N * 2));
    chk(dalloc(&dn, (size_t)2 * B));
    chk(dalloc(&drec, (size_t)2 * B * le::kMlpRecord));
    chk(dalloc(&dg, (size_t)le::kMlpParams));
    chk(dalloc(&dcfg, (size_t)B));
    chk(dalloc(&dpart, (size_t)B));
    chk(dalloc(&dst, 4));
    if (h == hipSuccess) {
        chk(hipMemcpy(dc, configs, sizeof(float) * n_cfg * N * 2, hipMemcpyHostToDevice));
        chk(hipMemcpy(dn, noise, sizeof(float) * 2 * B, hipMemcpyHostToDevice));
        chk(hipMemcpy(dcfg, cfg, sizeof(int) * B, hipMemcpyHostToDevice));
        chk(hipMemcpy(dpart, part, sizeof(int) * B, hipMemcpyHostToDevice));
        chk(hipMemset(dst, 0, sizeof(double) * 4));
    }
    if (h == hipSuccess)
        chk(le::launch_mlp_sample(e->mlp, dc, N, e->sys.inverse_filling, (float)(1.0 / e->R), (float)sigma0, dcfg, dpart,
                                  dn, B, drec, dst, e->stream));
    if (h == hipSuccess) chk(le::launch_mlp_grad(drec, B, dg, e->stream));
    double hs[4] = {0, 0, 0, 0};
    if (h == hipSuccess) chk(hipStreamSynchronize(e->stream));
    if (h == hipSuccess) {
        chk(hipMemcpy(hs, dst, sizeof(hs), hipMemcpyDeviceToHost));
        if (grad) chk(hipMemcpy(grad, dg, sizeof(float) * le::kMlpParams, hipMemcpyDeviceToHost));
    }
    for (void* q : {(void*)dc, (void*)dn, (void*)drec, (void*)dg, (void*)dcfg, (void*)dpart, (void*)dst}) (void)hipFree(q);
    if (h != hipSuccess) return fail(LE_ERR_HIP, hipGetErrorString(h));
    if (loss) *loss = hs[0];
    return LE_OK;
}

extern "C" int le_train_proposal(le_engine* e, const le_train_params* tp, le_train_report* rep)
{
    if (!e || !tp) return fail(LE_ERR_ARG, "null argument");
    if (e->prec != LE_PREC_FP32 || e->two_sided || e->rowwise)
        return fail(LE_ERR_ARG, "the learned proposal needs an FP32 edge engine with the plain estimator");
    if (tp->steps < 0 || tp->batch <= 0 || tp->n_configs <= 0 || tp->snapshot_sweeps < 0)
        return fail(LE_ERR_ARG, "bad training parameters");
    const auto t0 = std::chrono::steady_clock::now();
    if (!e->walkers_ready) {
        int rc = le_init_walkers(e);
        if (rc) return rc;
    }
    const int N = e->sys.particles_number;
    const int W = e->mc.walkers;
    const int ncfg = tp->n_configs;
    e->sigma0 = tp->sigma0 > 0 ? tp->sigma0 : e->step;
    // configurations of the plain chain (the engine's own walkers, advanced between snapshots)
    float* cfgs = nullptr;
    HIPCHK(dalloc(&cfgs, (size_t)ncfg * N * 2));
    const int saved = e->proposal;
    e->proposal = LE_PROPOSAL_PLAIN;
    int rc = LE_OK;
    for (int have = 0; have < ncfg && rc == LE_OK;) {
        if (tp->snapshot_sweeps > 0) {
            rc = e->run_sweeps(tp->snapshot_sweeps, e->sweep_counter, e->pos);
            e->sweep_counter += (unsigned long long)tp->snapshot_sweeps;
            e->sweeps_done += tp->snapshot_sweeps;
        }
        const int take = std::min(W, ncfg - have);
        hipError_t h = hipMemcpyAsync(cfgs + (size_t)have * N * 2, e->pos, sizeof(float) * take * N * 2,
                                      hipMemcpyDeviceToDevice, e->stream);
        if (h != hipSuccess) rc = fail(LE_ERR_HIP, hipGetErrorString(h));
        have += take;
    }
    e->proposal = saved;
    const int B = tp->batch;
    float *M1 = nullptr, *M2 = nullptr, *Gd = nullptr, *recs = nullptr, *noise = nullptr;
    int *cfg = nullptr, *part = nullptr;
    double* stats = nullptr;
    hipError_t h = hipSuccess;
    auto chk = [&](hipError_t x) { if (h == hipSuccess) h = x; };
    if (!e->mlp) chk(dalloc(&e->mlp, (size_t)le::kMlpParams));
    chk(dalloc(&M1, (size_t)le::kMlpParams));
    chk(dalloc(&M2, (size_t)le::kMlpParams));
    chk(dalloc(&Gd, (size_t)le::kMlpParams));
    chk(dalloc(&recs, (size_t)2 * B * le::kMlpRecord));
    chk(dalloc(&noise, (size_t)2 * B));
    chk(dalloc(&cfg, (size_t)B));
    chk(dalloc(&part, (size_t)B));
    chk(dalloc(&stats, (size_t)4 * (tp->steps + 1)));
    if (h == hipSuccess && rc == LE_OK) {
        chk(hipMemsetAsync(M1, 0, sizeof(float) * le::kMlpParams, e->stream));
        chk(hipMemsetAsync(M2, 0, sizeof(float) * le::kMlpParams, e->stream));
        chk(hipMemsetAsync(stats, 0, sizeof(double) * 4 * (tp->steps + 1), e->stream));
        chk(le::launch_mlp_init(e->mlp, tp->seed, e->stream));
        const float invR = (float)(1.0 / e->R);
        for (int t = 0; t <= tp->steps && h == hipSuccess; ++t) {
            chk(le::launch_train_samples(B, ncfg, N, tp->seed, t, cfg, part, noise, e->stream));
            chk(le::launch_mlp_sample(e->mlp, cfgs, N, e->sys.inverse_filling, invR, (float)e->sigma0, cfg, part, noise,
                                      B, recs, stats + 4 * t, e->stream));
            if (t == tp->steps) break;      // last pass only measures
            chk(le::launch_mlp_grad(recs, B, Gd, e->stream));
            chk(le::launch_adam(e->mlp, M1, M2, Gd, (float)tp->lr, t + 1, e->stream));
        }
    }
    std::vector<double> hs(4 * (tp->steps + 1), 0.0);
    if (h == hipSuccess) chk(hipMemcpyAsync(hs.data(), stats, sizeof(double) * hs.size(), hipMemcpyDeviceToHost, e->stream));
    if (h == hipSuccess) chk(hipStreamSynchronize(e->stream));
    for (void* q : {(void*)M1, (void*)M2, (void*)Gd, (void*)recs, (void*)noise, (void*)cfg, (void*)part, (void*)stats,
                    (void*)cfgs})
        (void)hipFree(q);
    if (rc) return rc;
    if (h != hipSuccess) return fail(LE_ERR_HIP, hipGetErrorString(h));
    rc = build_table(e, tp->table_size);
    if (rc) return rc;
    e->proposal = LE_PROPOSAL_LEARNED;
    if (rep) {
        rep->loss_first = hs[0];
        rep->accept_first = hs[1];
        rep->loss_last = hs[4 * tp->steps];
        rep->accept_last = hs[4 * tp->steps + 1];
        rep->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return LE_OK;
}
