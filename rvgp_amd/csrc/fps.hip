// a1  Furthest point sampling (RVGP/geometry.py:126-162).
//
// The reference materialises the dense n x n distance matrix and runs a
// numpy argmax/minimum loop.  Here no distance matrix exists: every
// distance is recomputed on the fly with the fixed direct formula of
// common.h, and the whole greedy loop runs in ONE persistent workgroup of
// 1024 threads (16 waves) so that an iteration costs one pass over the
// point set plus one block-wide (value, index) argmax -- no kernel launch
// and no host round trip per selected point.  The diameter (D.max(),
// geometry.py:144) is a separate all-pairs max over upper-triangle tiles.
#include "common.h"

namespace rvgp {

// ---- diameter: max over i<j of d2 --------------------------------------
constexpr int kDiamRows = 256;  // points per block (one per thread)

// candidate tile staged in LDS: 32 KB whatever the dimension
template <int DMAX>
struct DiamTile {
    static constexpr int cols = 4096 / DMAX < 128 ? 4096 / DMAX : 128;
};

template <int DMAX>
__global__ __launch_bounds__(kDiamRows) void diam2_kernel(const double* __restrict__ x, int n,
                                                          int D,
                                                          unsigned long long* __restrict__ out) {
    constexpr int kDiamCols = DiamTile<DMAX>::cols;
    __shared__ double tile[kDiamCols * DMAX];
    const int i0 = blockIdx.x * kDiamRows;
    const int j0 = blockIdx.y * kDiamCols;
    if (j0 + kDiamCols - 1 <= i0) return;  // tile entirely at or below the diagonal
    const int i = i0 + threadIdx.x;
    double xi[DMAX];
#pragma unroll
    for (int k = 0; k < DMAX; ++k) xi[k] = (i < n && k < D) ? x[(size_t)i * D + k] : 0.0;
    for (int e = threadIdx.x; e < kDiamCols * D; e += kDiamRows) {
        int jj = e / D, kk = e - jj * D;
        int j = j0 + jj;
        tile[jj * DMAX + kk] = (j < n) ? x[(size_t)j * D + kk] : 0.0;
    }
    __syncthreads();
    double best = 0.0;
    if (i < n) {
        const int jend = min(kDiamCols, n - j0);
        for (int jj = 0; jj < jend; ++jj) {
            if (j0 + jj <= i) continue;
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < DMAX; ++k)
                if (k < D) {
                    double t = __dsub_rn(xi[k], tile[jj * DMAX + k]);
                    acc = __fma_rn(t, t, acc);
                }
            best = acc > best ? acc : best;
        }
    }
    best = wave_max(best);
    __shared__ double wbest[kDiamRows / kWave];
    if ((threadIdx.x & 63) == 0) wbest[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = wbest[0];
        for (int w = 1; w < kDiamRows / kWave; ++w) b = wbest[w] > b ? wbest[w] : b;
        // Non-negative doubles order like their bit patterns.
        atomicMax(out, (unsigned long long)__double_as_longlong(b));
    }
}

// ---- the greedy loop ---------------------------------------------------
constexpr int kFpsThreads = 1024;
constexpr int kFpsMaxD = 8192;

struct ArgMax {
    double v;
    int i;
};

// (value descending, index ascending): numpy.argmax returns the first max.
__device__ __forceinline__ bool better(double v, int i, double bv, int bi) {
    return v > bv || (v == bv && i < bi);
}

__device__ __forceinline__ ArgMax block_argmax(double v, int i, double* sv, int* si) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        double ov = __shfl_xor(v, o, 64);
        int oi = __shfl_xor(i, o, 64);
        if (better(ov, oi, v, i)) { v = ov; i = oi; }
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sv[w] = v; si[w] = i; }
    __syncthreads();
    ArgMax r{sv[0], si[0]};
    for (int k = 1; k < kFpsThreads / kWave; ++k)
        if (better(sv[k], si[k], r.v, r.i)) { r.v = sv[k]; r.i = si[k]; }
    __syncthreads();  // sv/si are reused by the next call
    return r;
}

__global__ __launch_bounds__(kFpsThreads) void fps_kernel(
    const double* __restrict__ x, int n, int D, int n_iter, int start, double spacing,
    int use_spacing, const double* __restrict__ diam2, double* __restrict__ ds,
    int* __restrict__ perm, double* __restrict__ lambdas, int* __restrict__ n_sel) {
    extern __shared__ double sx[];  // the selected point's D coordinates
    __shared__ double sv[kFpsThreads / kWave];
    __shared__ int si[kFpsThreads / kWave];
    const double diam = use_spacing ? sqrt(diam2[0]) : 1.0;

    for (int c = threadIdx.x; c < D; c += kFpsThreads) sx[c] = x[(size_t)start * D + c];
    __syncthreads();
    double bv = -1.0;
    int bi = 0x7fffffff;
    for (int j = threadIdx.x; j < n; j += kFpsThreads) {
        double v = sqrt(d2_direct(sx, x + (size_t)j * D, D));
        ds[j] = v;
        if (better(v, j, bv, bi)) { bv = v; bi = j; }
    }
    if (threadIdx.x == 0) {
        perm[0] = start;
        lambdas[0] = 0.0;
    }
    int count = n_iter;
    for (int it = 1; it < n_iter; ++it) {
        ArgMax am = block_argmax(bv, bi, sv, si);
        if (threadIdx.x == 0) {
            perm[it] = am.i;
            lambdas[it] = am.v;
        }
        if (use_spacing && am.v / diam < spacing) {  // geometry.py:156-160
            count = it;
            break;
        }
        for (int c = threadIdx.x; c < D; c += kFpsThreads) sx[c] = x[(size_t)am.i * D + c];
        __syncthreads();
        bv = -1.0;
        bi = 0x7fffffff;
        for (int j = threadIdx.x; j < n; j += kFpsThreads) {
            double v = sqrt(d2_direct(sx, x + (size_t)j * D, D));
            double o = ds[j];
            v = v < o ? v : o;  // np.minimum
            ds[j] = v;
            if (better(v, j, bv, bi)) { bv = v; bi = j; }
        }
        __syncthreads();  // sx is rewritten next iteration
    }
    if (threadIdx.x == 0) n_sel[0] = count;
}

}  // namespace rvgp

using namespace rvgp;

extern "C" int rvgp_fps_diam2(const double* x, int n, int D, double* diam2, void* stream) {
    RVGP_CHECK_ARG(n >= 1 && D >= 1 && D <= 64, "rvgp_fps_diam2: need n >= 1 and 1 <= D <= 64");
    auto* out = reinterpret_cast<unsigned long long*>(diam2);
#define RVGP_DIAM(DM)                                                                    \
    hipLaunchKernelGGL(diam2_kernel<DM>, dim3(cdiv(n, kDiamRows), cdiv(n, DiamTile<DM>::cols)), \
                       dim3(kDiamRows), 0, as_stream(stream), x, n, D, out)
    if (D <= 4) RVGP_DIAM(4);
    else if (D <= 16) RVGP_DIAM(16);
    else if (D <= 32) RVGP_DIAM(32);
    else RVGP_DIAM(64);
#undef RVGP_DIAM
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}

extern "C" int rvgp_fps(const double* x, int n, int D, int n_iter, int start_idx, double spacing,
                        int use_spacing, const double* diam2, double* ds, int* perm,
                        double* lambdas, int* n_sel, void* stream) {
    // the selected point is staged in LDS: D <= 8192 (64 KiB); the spacing
    // mode's diameter kernel takes D <= 64
    RVGP_CHECK_ARG(n >= 1 && D >= 1 && D <= kFpsMaxD, "rvgp_fps: need n >= 1 and 1 <= D <= %d",
                   kFpsMaxD);
    RVGP_CHECK_ARG(n_iter >= 1, "rvgp_fps: n_iter must be >= 1");
    RVGP_CHECK_ARG(start_idx >= 0 && start_idx < n, "rvgp_fps: start_idx out of range");
    RVGP_CHECK_ARG(!use_spacing || diam2 != nullptr, "rvgp_fps: spacing mode needs diam2");
    hipLaunchKernelGGL(fps_kernel, dim3(1), dim3(kFpsThreads), sizeof(double) * D, as_stream(stream),
                       x, n, D,
                       n_iter, start_idx, spacing, use_spacing, diam2, ds, perm, lambdas, n_sel);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}
