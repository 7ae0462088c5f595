// Device helpers shared by the sampler, two-sided chain and learned-proposal kernels:
// the Philox-4x32-10 counter RNG, uniform and Gaussian draws, and in-group reductions.
#ifndef LE_DEVICE_HPP
#define LE_DEVICE_HPP

#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

namespace le {

// Philox-4x32-10 (Salmon et al., SC'11).  Stream layout used by every kernel:
//   key = {seed_lo, seed_hi}; ctr = {particle, sweep_lo, sweep_hi16 | tag << 16, walker}
__device__ __forceinline__ void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                                       uint32_t out[4])
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

template <typename Real> struct V2T;
template <> struct V2T<float> { using T = float2; };
template <> struct V2T<double> { using T = double2; };

// uniform in (0, 1]: (x + 1) 2^-32
__device__ __forceinline__ double u01d(uint32_t x) { return ((double)x + 1.0) * 2.3283064365386963e-10; }
__device__ __forceinline__ float u01f(uint32_t x) { return (float)(((double)x + 1.0) * 2.3283064365386963e-10); }

// Box-Muller pair: FP32 with the hardware log2 / sin / cos (angle in revolutions), FP64 exact
__device__ __forceinline__ void gauss2(uint32_t a, uint32_t b, float& n1, float& n2)
{
    float r = __fsqrt_rn(-1.3862943611198906f * __builtin_amdgcn_logf(u01f(a)));
    float t = u01f(b);
    n1 = r * __builtin_amdgcn_cosf(t);
    n2 = r * __builtin_amdgcn_sinf(t);
}

__device__ __forceinline__ void gauss2(uint32_t a, uint32_t b, double& n1, double& n2)
{
    double r = sqrt(-2.0 * log(u01d(a)));
    double s, c;
    sincos(6.283185307179586 * u01d(b), &s, &c);
    n1 = r * c;
    n2 = r * s;
}

// DPP lane exchanges within a 16-lane row: quad xor 1 (0xB1), quad xor 2 (0x4E),
// half-row mirror (0x141), row mirror (0x140), row rotate by 8 = xor 8 (0x128).
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v)
{
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the G lanes of a group (G a power of two, groups aligned to G), identical in
// every lane: each step pairs lanes holding equal partial sums of disjoint lane sets, and
// floating-point addition is commutative.
template <int G, typename Real>
__device__ __forceinline__ Real group_sum(Real v)
{
    if constexpr (sizeof(Real) == 4) {
        if constexpr (G >= 2) v += dpp_f<0xB1>(v);
        if constexpr (G >= 4) v += dpp_f<0x4E>(v);
        if constexpr (G >= 8) v += dpp_f<0x141>(v);
        if constexpr (G >= 16) v += dpp_f<0x140>(v);
    } else {
        if constexpr (G >= 2) v += dpp_d<0xB1>(v);
        if constexpr (G >= 4) v += dpp_d<0x4E>(v);
        if constexpr (G >= 8) v += dpp_d<0x141>(v);
        if constexpr (G >= 16) v += dpp_d<0x140>(v);
    }
    if constexpr (G >= 32) v += __shfl_xor(v, 16, 64);
    if constexpr (G >= 64) v += __shfl_xor(v, 32, 64);
    return v;
}

// explicit fused multiply-add: the sweep kernels spell out every contraction, so their
// arithmetic (and their chains) do not depend on the compiler's contraction choices
__device__ __forceinline__ float mad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double mad(double a, double b, double c) { return __builtin_fma(a, b, c); }

// logarithms of pair-distance products: FP32 takes log2 of 8 products at a time (their
// exponents kept in an integer), FP64 the natural log of 32
template <typename Real> struct LogOps;
template <> struct LogOps<float> {
    static constexpr int chunk = 8;
    __device__ static float lg(float x) { return __builtin_amdgcn_logf(fmaxf(x, 1e-30f)); }
    static constexpr float to_ln = 0.69314718055994531f;
};
template <> struct LogOps<double> {
    static constexpr int chunk = 32;
    __device__ static double lg(double x) { return log(fmax(x, 1e-300)); }
    static constexpr double to_ln = 1.0;
};

// compile-time loop: f(integral_constant<int, I>) for each I of the sequence
template <typename F, int... Is>
__device__ __forceinline__ void static_for(std::integer_sequence<int, Is...>, F&& f)
{
    (f(std::integral_constant<int, Is>{}), ...);
}

}  // namespace le

#endif
