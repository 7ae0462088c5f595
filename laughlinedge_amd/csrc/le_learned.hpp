// Launchers of the learned-proposal kernels (le_learned.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "le_kernels.hpp"

namespace le {

constexpr int kMlpParams = 4546;
constexpr int kMlpRecord = 264;    // floats per (sample, evaluation) training record

struct LearnedTable {
    const float* table;   // [T][T][T][2] log widths
    int T;
    float lo[3], hi[3];
    double invR;
};

hipError_t launch_mlp_init(float* P, unsigned long long seed, hipStream_t s);
hipError_t launch_train_samples(int B, int ncfg, int N, unsigned long long seed, int step, int* cfg, int* part,
                                float* noise, hipStream_t s);
hipError_t launch_mlp_sample(const float* P, const float* cfgs, int N, int m, float invR, float sigma0, const int* cfg,
                             const int* part, const float* noise, int B, float* recs, double* stats, hipStream_t s);
hipError_t launch_mlp_grad(const float* recs, int B, float* grad, hipStream_t s);
hipError_t launch_adam(float* P, float* M1, float* M2, const float* G, float lr, int t, hipStream_t s);
hipError_t launch_mlp_table(const float* P, int T, const float* lo, const float* hi, float sigma0, float* table,
                            hipStream_t s);
hipError_t launch_sweep_learned(const SweepArgs& a, const LearnedTable& lt, int G, hipStream_t s);
// the same sweeps with each lane's particles in registers: G in {8, 16}, ceil(N / G) in {13, 25}
bool learned_reg_supported(int G, int N);
hipError_t launch_sweep_learned_reg(const SweepArgs& a, const LearnedTable& lt, int G, hipStream_t s);
size_t learned_lds_bytes(int G, int N);

}  // namespace le
