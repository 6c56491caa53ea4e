// Shared helpers for the RVGP HIP kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "../../include/rvgp.h"

namespace rvgp {

void set_error(const char* fmt, ...);

// Return-on-error helpers for the C-ABI layer.
#define RVGP_CHECK_ARG(cond, ...)                     \
    do {                                              \
        if (!(cond)) {                                \
            ::rvgp::set_error(__VA_ARGS__);           \
            return RVGP_EINVAL;                       \
        }                                             \
    } while (0)

#define RVGP_CHECK_HIP(expr)                                                     \
    do {                                                                         \
        hipError_t _e = (expr);                                                  \
        if (_e != hipSuccess) {                                                  \
            ::rvgp::set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr,          \
                              hipGetErrorString(_e));                            \
            return RVGP_EHIP;                                                    \
        }                                                                        \
    } while (0)

#define RVGP_LAUNCH_CHECK() RVGP_CHECK_HIP(hipGetLastError())

static inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

static inline unsigned int cdiv(size_t a, size_t b) { return (unsigned int)((a + b - 1) / b); }
static inline size_t align_up(size_t a, size_t b) { return (a + b - 1) / b * b; }

constexpr int kWave = 64;

// Wave-wide reductions via cross-lane shuffles.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T u = __shfl_xor(v, o, 64);
        v = u > v ? u : v;
    }
    return v;
}

// Broadcast lane `l` of v (l wave-uniform) -- v_readlane into an SGPR.
__device__ __forceinline__ float bcast(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double bcast(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Direct squared distance with a fixed operation order:
//   acc = fma(t_k, t_k, acc), t_k = a_k - b_k, k = 0..D-1.
// The CPU oracle uses the same sequence, so results agree bit for bit.
__device__ __forceinline__ double d2_direct(const double* __restrict__ a,
                                            const double* __restrict__ b, int D) {
    double acc = 0.0;
    for (int k = 0; k < D; ++k) {
        double t = __dsub_rn(a[k], b[k]);
        acc = __fma_rn(t, t, acc);
    }
    return acc;
}

}  // namespace rvgp
