// a2  kNN graph (RVGP/geometry.py:100-123) and its symmetrised CSR
// (RVGP/lib/ptu_dijkstra.pyx:84-103).
//
// kNN: brute force.  A workgroup owns 64 queries (one per lane) and its four
// waves split the candidate range into four contiguous slices, so n = 1e5
// queries still put ~6 waves on every SIMD.  Each wave stages its candidate
// tile in LDS and reads each coordinate as a broadcast; the running top-k
// lives in registers as an unrolled sorted list (compare-and-shift network,
// no dynamic register indexing) with the current k-th distance cached in a
// register, so a rejected candidate costs one compare.  Coordinates past D
// are zero in both the query and the tile: fma(0, 0, acc) == acc, so the
// padded sum is bit-identical to the D-term one.  Within a slice candidates
// arrive in increasing index order and a strict "<" keeps the lower index
// among equal distances; the slices are then merged into wave 0 in slice
// order with the same insertion, which keeps that order across slices.
//
// Symmetrisation: degree count (atomics) -> exclusive scan -> scatter of
// forward + reverse edges + self loop -> per-row sort + de-duplicate ->
// scan -> compaction.  The per-row sort makes the result independent of the
// order in which atomics landed.
#include "common.h"

namespace rvgp {

constexpr int kKnnSlices = 4;             // waves per workgroup = candidate slices
constexpr int kKnnThreads = 64 * kKnnSlices;
constexpr int kKnnTileDoubles = 1024;     // LDS doubles per wave tile

// The running list is right-aligned in KMAX slots: slots [0, KMAX-k) hold a
// -inf sentinel that never moves, slots [KMAX-k, KMAX) hold the k best so far
// in ascending order (+inf while empty), so the k-th distance is always
// bd[KMAX-1].  Insert (d, j), d < bd[KMAX-1]: entries <= d stay in front,
// larger ones shift back by one and the last one drops out.
template <int KMAX>
__device__ __forceinline__ void knn_insert(double (&bd)[KMAX], int (&bi)[KMAX], double d, int j) {
#pragma unroll
    for (int p = KMAX - 1; p > 0; --p) {
        const bool shift = bd[p - 1] > d;
        const bool here = !shift && bd[p] > d;
        bi[p] = shift ? bi[p - 1] : (here ? j : bi[p]);
        bd[p] = shift ? bd[p - 1] : (here ? d : bd[p]);
    }
    if (bd[0] > d) {
        bd[0] = d;
        bi[0] = j;
    }
}

template <int DMAX, int KMAX>
__global__ __launch_bounds__(kKnnThreads) void knn_kernel(const double* __restrict__ x, int n,
                                                          int D, int k, int* __restrict__ nbr) {
    constexpr int TILE = kKnnTileDoubles / DMAX;  // candidates per tile
    constexpr int TILE_B = kKnnSlices * kKnnTileDoubles * 8;
    constexpr int MERGE_B = 64 * KMAX * 12;
    __shared__ __align__(16) unsigned char lds[TILE_B > MERGE_B ? TILE_B : MERGE_B];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    double* tile = reinterpret_cast<double*>(lds) + w * kKnnTileDoubles;

    double q[DMAX];
#pragma unroll
    for (int c = 0; c < DMAX; ++c) q[c] = (i < n && c < D) ? x[(size_t)i * D + c] : 0.0;
    double bd[KMAX];
    int bi[KMAX];
#pragma unroll
    for (int p = 0; p < KMAX; ++p) {
        // -inf sentinels in front, +inf (empty) in the k live slots
        bd[p] = __longlong_as_double(p < KMAX - k ? 0xfff0000000000000LL : 0x7ff0000000000000LL);
        bi[p] = -1;
    }

    // Slice w = candidates [s0, s1); every wave runs the same trip count.
    const int per = ((n + kKnnSlices - 1) / kKnnSlices + TILE - 1) / TILE * TILE;
    const int s0 = w * per, s1 = min(n, s0 + per);
    for (int t = 0; t < per; t += TILE) {
        const int j0 = s0 + t;
        __syncthreads();
        for (int e = lane; e < TILE * DMAX; e += 64) {
            const int jj = e / DMAX, c = e % DMAX, j = j0 + jj;
            tile[e] = (j < s1 && c < D) ? x[(size_t)j * D + c] : 0.0;
        }
        __syncthreads();
        const int jend = min(TILE, s1 - j0);
        // two candidates per pass: their DMAX-long FMA chains are independent
        // and interleave; they are then offered in index order
        int jj = 0;
        for (; jj + 1 < jend; jj += 2) {
            double acc0 = 0.0, acc1 = 0.0;
#pragma unroll
            for (int c = 0; c < DMAX; ++c) {
                const double t0 = __dsub_rn(q[c], tile[jj * DMAX + c]);
                const double t1 = __dsub_rn(q[c], tile[(jj + 1) * DMAX + c]);
                acc0 = __fma_rn(t0, t0, acc0);
                acc1 = __fma_rn(t1, t1, acc1);
            }
            const int j = j0 + jj;
            if (j != i && acc0 < bd[KMAX - 1]) knn_insert<KMAX>(bd, bi, acc0, j);
            if (j + 1 != i && acc1 < bd[KMAX - 1]) knn_insert<KMAX>(bd, bi, acc1, j + 1);
        }
        if (jj < jend) {
            const int j = j0 + jj;
            double acc = 0.0;
#pragma unroll
            for (int c = 0; c < DMAX; ++c) {
                const double tt = __dsub_rn(q[c], tile[jj * DMAX + c]);
                acc = __fma_rn(tt, tt, acc);
            }
            if (j != i && acc < bd[KMAX - 1]) knn_insert<KMAX>(bd, bi, acc, j);
        }
    }

    // Merge slices 1..3 into wave 0, in slice order.
    double* md = reinterpret_cast<double*>(lds);
    int* mi = reinterpret_cast<int*>(lds + 64 * KMAX * 8);
    for (int src = 1; src < kKnnSlices; ++src) {
        __syncthreads();
        if (w == src) {
#pragma unroll
            for (int p = 0; p < KMAX; ++p) {
                md[p * 64 + lane] = bd[p];
                mi[p * 64 + lane] = bi[p];
            }
        }
        __syncthreads();
        if (w == 0) {
            for (int p = KMAX - k; p < KMAX; ++p) {
                const double d = md[p * 64 + lane];
                if (d < bd[KMAX - 1]) knn_insert<KMAX>(bd, bi, d, mi[p * 64 + lane]);
            }
        }
    }
    if (w == 0 && i < n) {
#pragma unroll
        for (int p = 0; p < KMAX; ++p)
            if (p >= KMAX - k) nbr[(size_t)i * k + p - (KMAX - k)] = bi[p];
    }
}

// ---- symmetrisation ------------------------------------------------------

__global__ void sym_count_kernel(const int* __restrict__ nbr, int n, int k, int* __restrict__ cnt) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n * k) atomicAdd(&cnt[nbr[e]], 1);
    if (e < n) atomicAdd(&cnt[e], k + 1);
}

// Exclusive scan of cnt[0..n) into out[0..n], out[n] = total, in one block.
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void scan_kernel(const int* __restrict__ cnt, int n,
                                                            int* __restrict__ out) {
    __shared__ int part[kScanThreads];
    const int per = (n + kScanThreads - 1) / kScanThreads;
    const int b = threadIdx.x * per;
    const int e = min(n, b + per);
    int s = 0;
    for (int i = b; i < e; ++i) s += cnt[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < kScanThreads; o <<= 1) {
        int v = (threadIdx.x >= o) ? part[threadIdx.x - o] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int run = part[threadIdx.x] - s;
    for (int i = b; i < e; ++i) {
        out[i] = run;
        run += cnt[i];
    }
    if (threadIdx.x == kScanThreads - 1) out[n] = part[kScanThreads - 1];
}

__global__ void sym_fill_kernel(const int* __restrict__ nbr, int n, int k,
                                const int* __restrict__ start, int* __restrict__ cursor,
                                int* __restrict__ tmp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int pos = atomicAdd(&cursor[i], k + 1);
    int* dst = tmp + start[i] + pos;
    dst[0] = i;
    for (int q = 0; q < k; ++q) {
        int j = nbr[(size_t)i * k + q];
        dst[1 + q] = j;
        int pj = atomicAdd(&cursor[j], 1);
        tmp[start[j] + pj] = i;
    }
}

// Sort each row segment in place (insertion sort; rows are short) and
// count the unique entries.
__global__ void sym_sort_kernel(const int* __restrict__ start, int n, int* __restrict__ tmp,
                                int* __restrict__ ucnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int* seg = tmp + start[i];
    const int len = start[i + 1] - start[i];
    for (int a = 1; a < len; ++a) {
        int v = seg[a];
        int p = a - 1;
        while (p >= 0 && seg[p] > v) {
            seg[p + 1] = seg[p];
            --p;
        }
        seg[p + 1] = v;
    }
    int u = 0;
    for (int a = 0; a < len; ++a) u += (a == 0 || seg[a] != seg[a - 1]);
    ucnt[i] = u;
}

__global__ void sym_compact_kernel(const int* __restrict__ start, int n,
                                   const int* __restrict__ tmp, const int* __restrict__ rowptr,
                                   int* __restrict__ cols) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int* seg = tmp + start[i];
    const int len = start[i + 1] - start[i];
    int o = rowptr[i];
    for (int a = 0; a < len; ++a)
        if (a == 0 || seg[a] != seg[a - 1]) cols[o++] = seg[a];
}

}  // namespace rvgp

using namespace rvgp;

template <int DMAX>
static void launch_knn(dim3 g, hipStream_t s, const double* x, int n, int D, int k, int* nbr) {
    if (k <= 16)
        hipLaunchKernelGGL((knn_kernel<DMAX, 16>), g, dim3(kKnnThreads), 0, s, x, n, D, k, nbr);
    else if (k <= 24)
        hipLaunchKernelGGL((knn_kernel<DMAX, 24>), g, dim3(kKnnThreads), 0, s, x, n, D, k, nbr);
    else if (k <= 32)
        hipLaunchKernelGGL((knn_kernel<DMAX, 32>), g, dim3(kKnnThreads), 0, s, x, n, D, k, nbr);
    else
        hipLaunchKernelGGL((knn_kernel<DMAX, 64>), g, dim3(kKnnThreads), 0, s, x, n, D, k, nbr);
}

extern "C" int rvgp_knn(const double* x, int n, int D, int k, int* nbr, void* stream) {
    RVGP_CHECK_ARG(k > 0 && k < n, "rvgp_knn: need 0 < k < n (k=%d, n=%d)", k, n);
    RVGP_CHECK_ARG(k <= 64, "rvgp_knn: k <= 64 supported");
    RVGP_CHECK_ARG(D >= 1 && D <= 64, "rvgp_knn: 1 <= D <= 64 supported");
    dim3 g(cdiv(n, 64));  // 64 queries per workgroup
    hipStream_t s = as_stream(stream);
    if (D <= 4) launch_knn<4>(g, s, x, n, D, k, nbr);
    else if (D <= 8) launch_knn<8>(g, s, x, n, D, k, nbr);
    else if (D <= 16) launch_knn<16>(g, s, x, n, D, k, nbr);
    else if (D <= 32) launch_knn<32>(g, s, x, n, D, k, nbr);
    else launch_knn<64>(g, s, x, n, D, k, nbr);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}

extern "C" size_t rvgp_graph_symmetrize_ws_bytes(int n, int k) {
    // cnt(n+1) start(n+1) cursor(n) ucnt(n+1) tmp(n*(2k+1))
    return sizeof(int) * ((size_t)4 * n + 3 + (size_t)n * (2 * (size_t)k + 1));
}

extern "C" int rvgp_graph_symmetrize(const int* nbr, int n, int k, int* rowptr, int* cols,
                                     void* ws, size_t ws_bytes, void* stream) {
    RVGP_CHECK_ARG(n >= 1 && k >= 1, "rvgp_graph_symmetrize: need n >= 1, k >= 1");
    RVGP_CHECK_ARG(ws_bytes >= rvgp_graph_symmetrize_ws_bytes(n, k),
                   "rvgp_graph_symmetrize: workspace too small");
    hipStream_t s = as_stream(stream);
    int* cnt = static_cast<int*>(ws);
    int* start = cnt + (n + 1);
    int* cursor = start + (n + 1);
    int* ucnt = cursor + n;
    int* tmp = ucnt + (n + 1);
    RVGP_CHECK_HIP(hipMemsetAsync(cnt, 0, sizeof(int) * (size_t)(n + 1), s));
    RVGP_CHECK_HIP(hipMemsetAsync(cursor, 0, sizeof(int) * (size_t)n, s));
    const int T = 256;
    hipLaunchKernelGGL(sym_count_kernel, dim3(cdiv((size_t)n * k, T)), dim3(T), 0, s, nbr, n, k, cnt);
    RVGP_LAUNCH_CHECK();
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(kScanThreads), 0, s, cnt, n, start);
    RVGP_LAUNCH_CHECK();
    hipLaunchKernelGGL(sym_fill_kernel, dim3(cdiv(n, T)), dim3(T), 0, s, nbr, n, k, start, cursor, tmp);
    RVGP_LAUNCH_CHECK();
    hipLaunchKernelGGL(sym_sort_kernel, dim3(cdiv(n, T)), dim3(T), 0, s, start, n, tmp, ucnt);
    RVGP_LAUNCH_CHECK();
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(kScanThreads), 0, s, ucnt, n, rowptr);
    RVGP_LAUNCH_CHECK();
    hipLaunchKernelGGL(sym_compact_kernel, dim3(cdiv(n, T)), dim3(T), 0, s, start, n, tmp, rowptr, cols);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}
