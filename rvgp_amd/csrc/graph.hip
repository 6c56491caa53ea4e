// a2  kNN graph (RVGP/geometry.py:100-123) and its symmetrised CSR
// (RVGP/lib/ptu_dijkstra.pyx:84-103).
//
// kNN: brute force, one query per lane, candidate tiles staged in LDS so a
// wave reads each candidate coordinate as an LDS broadcast; the running top-k
// lives in registers as an unrolled sorted list (compare-and-shift network,
// no dynamic register indexing).  Candidates arrive in increasing index
// order, so a strict "<" keeps the lower index among equal distances.
//
// Symmetrisation: degree count (atomics) -> exclusive scan -> scatter of
// forward + reverse edges + self loop -> per-row sort + de-duplicate ->
// scan -> compaction.  The per-row sort makes the result independent of the
// order in which atomics landed.
#include "common.h"

namespace rvgp {

constexpr int kKnnThreads = 256;
constexpr int kKnnTile = 64;

template <int DMAX, int KMAX>
__global__ __launch_bounds__(kKnnThreads) void knn_kernel(const double* __restrict__ x, int n,
                                                          int D, int k, int* __restrict__ nbr) {
    __shared__ double tile[kKnnTile * DMAX];
    const int i = blockIdx.x * kKnnThreads + threadIdx.x;
    double q[DMAX];
#pragma unroll
    for (int c = 0; c < DMAX; ++c) q[c] = (i < n && c < D) ? x[(size_t)i * D + c] : 0.0;
    double bd[KMAX];
    int bi[KMAX];
#pragma unroll
    for (int p = 0; p < KMAX; ++p) {
        bd[p] = __longlong_as_double(0x7ff0000000000000LL);  // +inf
        bi[p] = -1;
    }
    for (int j0 = 0; j0 < n; j0 += kKnnTile) {
        __syncthreads();
        for (int e = threadIdx.x; e < kKnnTile * D; e += kKnnThreads) {
            int jj = e / D, c = e - jj * D;
            int j = j0 + jj;
            tile[jj * DMAX + c] = (j < n) ? x[(size_t)j * D + c] : 0.0;
        }
        __syncthreads();
        const int jend = min(kKnnTile, n - j0);
        for (int jj = 0; jj < jend; ++jj) {
            const int j = j0 + jj;
            double acc = 0.0;
#pragma unroll
            for (int c = 0; c < DMAX; ++c)
                if (c < D) {
                    double t = __dsub_rn(q[c], tile[jj * DMAX + c]);
                    acc = __fma_rn(t, t, acc);
                }
            // worst kept entry is bd[k-1]; select it without dynamic indexing
            double worst = bd[0];
#pragma unroll
            for (int p = 1; p < KMAX; ++p)
                if (p == k - 1) worst = bd[p];
            if (j != i && acc < worst) {
#pragma unroll
                for (int p = KMAX - 1; p >= 0; --p) {
                    if (p < k && bd[p] > acc) {
                        double prev = (p > 0) ? bd[p > 0 ? p - 1 : 0] : -1.0;
                        if (p > 0 && prev > acc) {
                            bd[p] = prev;
                            bi[p] = bi[p > 0 ? p - 1 : 0];
                        } else {
                            bd[p] = acc;
                            bi[p] = j;
                        }
                    }
                }
            }
        }
    }
    if (i < n) {
#pragma unroll
        for (int p = 0; p < KMAX; ++p)
            if (p < k) nbr[(size_t)i * k + p] = bi[p];
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
    else if (k <= 32)
        hipLaunchKernelGGL((knn_kernel<DMAX, 32>), g, dim3(kKnnThreads), 0, s, x, n, D, k, nbr);
    else
        hipLaunchKernelGGL((knn_kernel<DMAX, 64>), g, dim3(kKnnThreads), 0, s, x, n, D, k, nbr);
}

extern "C" int rvgp_knn(const double* x, int n, int D, int k, int* nbr, void* stream) {
    RVGP_CHECK_ARG(k > 0 && k < n, "rvgp_knn: need 0 < k < n (k=%d, n=%d)", k, n);
    RVGP_CHECK_ARG(k <= 64, "rvgp_knn: k <= 64 supported");
    RVGP_CHECK_ARG(D >= 1 && D <= 64, "rvgp_knn: 1 <= D <= 64 supported");
    dim3 g(cdiv(n, kKnnThreads));
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
