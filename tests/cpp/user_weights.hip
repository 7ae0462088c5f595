// Test helper: a user weight callback for the C ABI (le_weight_fn), built as a small
// shared library.  U0 = sum_i (x_i / R)^2 (the reference's example excitation profile,
// main.cu:72-80) and U1 = sum_i r_i^2 cos(2 theta_i) / R^2.
#include <hip/hip_runtime.h>

#include <cmath>

template <typename P, typename W>
__global__ void k_user(const P* pos, int walkers, int N, W* w, long ld, double invR2)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= walkers) return;
    double u0 = 0.0, u1 = 0.0;
    for (int i = 0; i < N; ++i) {
        const double x = pos[(size_t)s * N + i].x, y = pos[(size_t)s * N + i].y;
        u0 += x * x * invR2;
        u1 += (x * x - y * y) * invR2;   // r^2 cos(2 theta)
    }
    w[s] = (W)u0;
    w[ld + s] = (W)u1;
}

static double g_invR2 = 1.0;

extern "C" void le_test_set_radius(double R) { g_invR2 = 1.0 / (R * R); }

extern "C" int le_test_weight_fn(void*, const void* pos, int pos_double, int walkers, int N, void* w, long ld,
                                 int w_double, void* stream)
{
    dim3 grid((walkers + 255) / 256), block(256);
    hipStream_t st = (hipStream_t)stream;
    if (pos_double && w_double)
        hipLaunchKernelGGL((k_user<double2, double>), grid, block, 0, st, (const double2*)pos, walkers, N, (double*)w, ld, g_invR2);
    else if (pos_double)
        hipLaunchKernelGGL((k_user<double2, float>), grid, block, 0, st, (const double2*)pos, walkers, N, (float*)w, ld, g_invR2);
    else if (w_double)
        hipLaunchKernelGGL((k_user<float2, double>), grid, block, 0, st, (const float2*)pos, walkers, N, (double*)w, ld, g_invR2);
    else
        hipLaunchKernelGGL((k_user<float2, float>), grid, block, 0, st, (const float2*)pos, walkers, N, (float*)w, ld, g_invR2);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

// Edge kicks: U_k = sum_i cos(k theta_i) exp(-(r_i - R)^2 / 2), k = 1..K, the spatial profile
// of the perturbations A cos(k theta) exp(-(r - R)^2 / 2) f(t).
template <typename P, typename W>
__global__ void k_kicks(const P* pos, int walkers, int N, W* w, long ld, int K, double R)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= walkers) return;
    double u[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < N; ++i) {
        const double x = pos[(size_t)s * N + i].x, y = pos[(size_t)s * N + i].y;
        const double r = sqrt(x * x + y * y), th = atan2(y, x);
        const double g = exp(-0.5 * (r - R) * (r - R));
        for (int k = 0; k < K; ++k) u[k] += cos((k + 1) * th) * g;
    }
    for (int k = 0; k < K; ++k) w[(size_t)k * ld + s] = (W)u[k];
}

static int g_kicks = 1;
static double g_kick_R = 1.0;

extern "C" void le_test_set_kicks(int K, double R)
{
    g_kicks = K < 1 ? 1 : (K > 8 ? 8 : K);
    g_kick_R = R;
}

extern "C" int le_test_kick_fn(void*, const void* pos, int pos_double, int walkers, int N, void* w, long ld,
                               int w_double, void* stream)
{
    dim3 grid((walkers + 127) / 128), block(128);
    hipStream_t st = (hipStream_t)stream;
    if (pos_double && w_double)
        hipLaunchKernelGGL((k_kicks<double2, double>), grid, block, 0, st, (const double2*)pos, walkers, N, (double*)w, ld, g_kicks, g_kick_R);
    else if (pos_double)
        hipLaunchKernelGGL((k_kicks<double2, float>), grid, block, 0, st, (const double2*)pos, walkers, N, (float*)w, ld, g_kicks, g_kick_R);
    else if (w_double)
        hipLaunchKernelGGL((k_kicks<float2, double>), grid, block, 0, st, (const float2*)pos, walkers, N, (double*)w, ld, g_kicks, g_kick_R);
    else
        hipLaunchKernelGGL((k_kicks<float2, float>), grid, block, 0, st, (const float2*)pos, walkers, N, (float*)w, ld, g_kicks, g_kick_R);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}
