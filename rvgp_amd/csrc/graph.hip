// a2  kNN graph (RVGP/geometry.py:100-123) and its symmetrised CSR
// (RVGP/lib/ptu_dijkstra.pyx:84-103).
//
// kNN: brute force, exact.  Candidates are split into S contiguous slices;
// a workgroup scans one slice for 256 queries (one per lane, four waves that
// share the candidate stream) and writes each query's top-k of the slice; a
// merge kernel combines the S lists per query in slice order.  Each wave
// reads a candidate's coordinates with scalar loads (the index is
// wave-uniform); the running top-k lives in registers as an unrolled sorted
// list (compare-and-shift network, no dynamic register indexing) whose k-th
// distance rejects a candidate with one compare, and the partial sum over
// the first 8 coordinates rejects most candidates before the rest are read.
// Coordinates past D are zero in both the query and the candidate:
// fma(0, 0, acc) == acc, so the padded sum is bit-identical to the D-term
// one.  Within a slice candidates arrive in increasing index order and a
// strict "<" keeps the lower index among equal distances; the merge keeps
// that order across slices.  For n > 8192 a first scan + merge over the
// first 8192 points gives every query a bound on its k-th distance that
// starts the full scan's lists.
//
// Symmetrisation: degree count (atomics) -> exclusive scan -> scatter of
// forward + reverse edges + self loop -> per-row sort + de-duplicate ->
// scan -> compaction.  The per-row sort makes the result independent of the
// order in which atomics landed.
#include "common.h"

namespace rvgp {

constexpr int kKnnWaves = 4;              // query waves per workgroup, one candidate stream
constexpr int kKnnThreads = 64 * kKnnWaves;
constexpr int kKnnMergeThreads = 256;
#define KNN_CHUNK 8                       // coordinates between early-exit tests
#define KNN_GROUP 4                       // candidates whose first chunks load together

// The running list is right-aligned in KMAX slots: slots [0, KMAX-k) hold a
// -inf sentinel that never moves, slots [KMAX-k, KMAX) hold the k best so far
// in ascending order, so the k-th distance is always bd[KMAX-1].  Insert
// (d, j), d < bd[KMAX-1]: entries <= d stay in front, larger ones shift back
// by one and the last one drops out.
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

// Query at position pos: qperm[pos] (a permutation of [0, n)), pos without
// an order; -1 past n or for an entry out of range, which is then skipped.
__device__ __forceinline__ int knn_query(const int* __restrict__ qperm, int pos, int n) {
    if (pos >= n) return -1;
    const int v = qperm != nullptr ? qperm[pos] : pos;
    return (unsigned)v < (unsigned)n ? v : -1;
}

// The k live slots start at +inf, or, given a bound thr_in[i] >= query i's
// k-th squared distance, at the next double above it: then only candidates
// with d2 <= thr_in[i] are taken -- every one that can end in the top k, in
// the same order, so the result is unchanged -- and a slot never filled keeps
// that value, which no later insertion test lets through.
template <int KMAX>
__device__ __forceinline__ void knn_init(double (&bd)[KMAX], int (&bi)[KMAX], int k,
                                         const double* __restrict__ thr_in, int i, int n) {
    double live = __longlong_as_double(0x7ff0000000000000LL);
    if (thr_in != nullptr && i < n && thr_in[i] < live)  // a finite bound
        live = __longlong_as_double(__double_as_longlong(thr_in[i]) + 1);
#pragma unroll
    for (int p = 0; p < KMAX; ++p) {
        bd[p] = p < KMAX - k ? __longlong_as_double(0xfff0000000000000LL) : live;
        bi[p] = -1;
    }
}

// Scan: workgroup b takes candidate slice b % S and the kKnnWaves * 64
// query positions of group b / S (one per lane; position pos is query
// qperm[pos], or pos without a query order).  Its waves read the same candidates
// in the same order, so a line the scalar cache fetches for one wave serves
// the others, and consecutive workgroups -- which the dispatcher spreads over
// the XCDs in turn -- take different slices, so with S = 8 each XCD's L2
// holds one slice.  Each wave keeps its lanes' top-k of the slice in
// registers and writes it to lists[(s * k + p) * n + pos].  A query order
// that puts manifold neighbours in one wave makes the early exit on the
// first chunk (a wave-wide test) fire for almost every candidate; the
// result does not depend on it.
template <int DMAX, int KMAX, bool EXACT>
__global__ __launch_bounds__(kKnnThreads) __attribute__((amdgpu_waves_per_eu(3))) void knn_scan_kernel(
    const double* __restrict__ x, int n, int D, int k, int cand_end, int S,
    const int* __restrict__ qperm, const double* __restrict__ thr_in, double* __restrict__ ld,
    int* __restrict__ li) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int sl = blockIdx.x % S;
    const int q0 = (blockIdx.x / S * kKnnWaves + w) * 64;
    if (q0 >= n) return;  // wave-uniform; the kernel has no barrier
    const int pos = q0 + lane;
    const int i = knn_query(qperm, pos, n);  // -1: no query in this lane

    double q[DMAX];
#pragma unroll
    for (int c = 0; c < DMAX; ++c) q[c] = (i >= 0 && c < D) ? x[(size_t)i * D + c] : 0.0;
    double bd[KMAX];
    int bi[KMAX];
    knn_init<KMAX>(bd, bi, k, thr_in, i >= 0 ? pos : n, n);

    // The candidate index is wave-uniform, so its coordinates are read with
    // scalar loads into SGPRs that feed the FMAs directly: no LDS staging,
    // whose broadcast reads (64 lanes x 8 bytes per coordinate) bound an
    // earlier tiled form.  Coordinates c >= D read as 0 in both the query and
    // the candidate: fma(0, 0, acc) == acc, so the padded sum is bit-identical
    // to the D-term one.
    const int per = (cand_end + S - 1) / S;
    const int s0 = sl * per, s1 = min(cand_end, s0 + per);
    constexpr int CH = KNN_CHUNK < DMAX ? KNN_CHUNK : DMAX;
    // Candidate c's coordinate t, zero past D.  EXACT (D == DMAX) reads a
    // compile-time-sized contiguous run, so a chunk becomes wide scalar loads
    // issued together; otherwise the load uses a clamped index and a select
    // (a branch per coordinate would serialise the loads).
    auto coord = [&](int c, int t) -> double {
        if (EXACT) return x[(size_t)c * DMAX + t];
        const double v = x[(size_t)c * D + (t < D ? t : D - 1)];
        return t < D ? v : 0.0;
    };
    // Finish candidate c's sum from coordinate `from` on, leaving early once
    // the partial sum reaches every lane's k-th distance (the partial sums
    // only grow: fma(t, t, acc) >= acc for a finite acc), then offer it.
    auto finish = [&](int c, double acc, int from) {
#pragma unroll
        for (int b = 0; b < DMAX; b += CH) {
            if (b < from) continue;
            double v[CH];
#pragma unroll
            for (int u = 0; u < CH; ++u) v[u] = coord(c, b + u);
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const double t = __dsub_rn(q[b + u], v[u]);
                acc = __fma_rn(t, t, acc);
            }
            if (b + CH < DMAX && __all(acc >= bd[KMAX - 1])) return;
        }
        if (c != i && acc < bd[KMAX - 1]) knn_insert<KMAX>(bd, bi, acc, c);
    };

    // Groups of KNN_GROUP candidates: the first chunk of every candidate in
    // the group is fetched by one batch of scalar loads (one exposed memory
    // latency per group rather than per candidate), then each candidate
    // whose first-chunk partial sum is below some lane's k-th distance is
    // finished, in index order.
    int j = s0;
    for (; j + KNN_GROUP <= s1; j += KNN_GROUP) {
        double v[KNN_GROUP][CH];
#pragma unroll
        for (int g = 0; g < KNN_GROUP; ++g)
#pragma unroll
            for (int u = 0; u < CH; ++u) v[g][u] = coord(j + g, u);
        // keep the group's loads issued back to back, ahead of the arithmetic
        __builtin_amdgcn_sched_barrier(0);
        double acc[KNN_GROUP];
#pragma unroll
        for (int g = 0; g < KNN_GROUP; ++g) {
            acc[g] = 0.0;
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const double t = __dsub_rn(q[u], v[g][u]);
                acc[g] = __fma_rn(t, t, acc[g]);
            }
        }
        // The group's liveness mask is formed before any candidate is
        // finished, so the first-chunk values are dead by then (the finish
        // path reuses their registers).  thr is >= the k-th distance at the
        // time a later candidate of the group is finished, so a candidate all
        // lanes reject against thr is rejected then too.
        const double thr = bd[KMAX - 1];
        unsigned live = 0;
#pragma unroll
        for (int g = 0; g < KNN_GROUP; ++g)
            if (CH == DMAX || !__all(acc[g] >= thr)) live |= 1u << g;
        if (live == 0) continue;
#pragma unroll
        for (int g = 0; g < KNN_GROUP; ++g)
            if (live & (1u << g)) finish(j + g, acc[g], CH);
    }
    for (; j < s1; ++j) finish(j, 0.0, 0);

    if (i >= 0) {
#pragma unroll
        for (int p = 0; p < KMAX; ++p) {
            if (p < KMAX - k) continue;
            const size_t e = ((size_t)sl * k + (p - (KMAX - k))) * n + pos;
            ld[e] = bd[p];
            li[e] = bi[p];
        }
    }
}

// Merge: one thread per query inserts the S slice lists in slice order, each
// in its own ascending order, with the same strict insertion.  The slices are
// increasing index ranges, so entries of equal distance arrive in index order
// and the result is the top k of one scan over all candidates.  A list is
// left at its first entry that the current k-th distance rejects: the rest
// of that list is no smaller.
template <int KMAX>
__global__ __launch_bounds__(kKnnMergeThreads) void knn_merge_kernel(
    int n, int k, int S, const int* __restrict__ qperm, const double* __restrict__ thr_in,
    const double* __restrict__ ld, const int* __restrict__ li, double* __restrict__ thr_out,
    int* __restrict__ nbr) {
    const int pos = blockIdx.x * kKnnMergeThreads + threadIdx.x;
    if (pos >= n) return;
    const int i = knn_query(qperm, pos, n);
    if (i < 0) return;
    double bd[KMAX];
    int bi[KMAX];
    knn_init<KMAX>(bd, bi, k, thr_in, pos, n);
    for (int s = 0; s < S; ++s) {
        for (int p = 0; p < k; ++p) {
            const size_t e = ((size_t)s * k + p) * n + pos;
            const double d = ld[e];
            if (!(d < bd[KMAX - 1])) break;
            knn_insert<KMAX>(bd, bi, d, li[e]);
        }
    }
    if (thr_out != nullptr) {
        thr_out[pos] = bd[KMAX - 1];  // k-th squared distance over the candidates scanned
    } else {
#pragma unroll
        for (int p = 0; p < KMAX; ++p)
            if (p >= KMAX - k) nbr[(size_t)i * k + p - (KMAX - k)] = bi[p];
    }
}

// ---- symmetrisation ------------------------------------------------------

// A neighbour id outside [0, n) (only possible for an nbr array that did not
// come from a complete rvgp_knn, e.g. with a query order that is not a
// permutation) is read as the row itself: it adds a duplicate self loop,
// which the de-duplication removes, and no reverse edge -- every write stays
// inside the row segments the counts reserved.
__global__ void sym_count_kernel(const int* __restrict__ nbr, int n, int k, int* __restrict__ cnt) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n * k) {
        const int j = nbr[e];
        if ((unsigned)j < (unsigned)n) atomicAdd(&cnt[j], 1);
    }
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
        const int j = nbr[(size_t)i * k + q];
        if ((unsigned)j >= (unsigned)n) {
            dst[1 + q] = i;
            continue;
        }
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

template <int DMAX, int KM>
static void launch_knn_scan(hipStream_t st, const double* x, int n, int D, int k, int cand_end,
                            int S, const int* qperm, const double* thr_in, double* ld, int* li) {
    const dim3 g(cdiv(n, 64 * kKnnWaves) * S);
    if (D == DMAX)
        hipLaunchKernelGGL((knn_scan_kernel<DMAX, KM, true>), g, dim3(kKnnThreads), 0, st, x, n,
                           D, k, cand_end, S, qperm, thr_in, ld, li);
    else
        hipLaunchKernelGGL((knn_scan_kernel<DMAX, KM, false>), g, dim3(kKnnThreads), 0, st, x, n,
                           D, k, cand_end, S, qperm, thr_in, ld, li);
}

template <int KM>
static void launch_knn_km(hipStream_t st, const double* x, int n, int D, int k, int cand_end,
                          int S, const int* qperm, const double* thr_in, double* ld, int* li,
                          double* thr_out, int* nbr) {
    if (D <= 4) launch_knn_scan<4, KM>(st, x, n, D, k, cand_end, S, qperm, thr_in, ld, li);
    else if (D <= 8) launch_knn_scan<8, KM>(st, x, n, D, k, cand_end, S, qperm, thr_in, ld, li);
    else if (D <= 16) launch_knn_scan<16, KM>(st, x, n, D, k, cand_end, S, qperm, thr_in, ld, li);
    else if (D <= 32) launch_knn_scan<32, KM>(st, x, n, D, k, cand_end, S, qperm, thr_in, ld, li);
    else launch_knn_scan<64, KM>(st, x, n, D, k, cand_end, S, qperm, thr_in, ld, li);
    hipLaunchKernelGGL((knn_merge_kernel<KM>), dim3(cdiv(n, kKnnMergeThreads)),
                       dim3(kKnnMergeThreads), 0, st, n, k, S, qperm, thr_in, ld, li, thr_out,
                       nbr);
}

// One scan + merge over candidates [0, cand_end).
static void launch_knn(hipStream_t st, const double* x, int n, int D, int k, int cand_end, int S,
                       const int* qperm, const double* thr_in, double* ld, int* li,
                       double* thr_out, int* nbr) {
    if (k <= 16) launch_knn_km<16>(st, x, n, D, k, cand_end, S, qperm, thr_in, ld, li, thr_out, nbr);
    else if (k <= 24) launch_knn_km<24>(st, x, n, D, k, cand_end, S, qperm, thr_in, ld, li, thr_out, nbr);
    else if (k <= 32) launch_knn_km<32>(st, x, n, D, k, cand_end, S, qperm, thr_in, ld, li, thr_out, nbr);
    else launch_knn_km<64>(st, x, n, D, k, cand_end, S, qperm, thr_in, ld, li, thr_out, nbr);
}

// candidates of the bounding pass: the first kKnnBoundCands points
constexpr int kKnnBoundCands = 8192;

// Candidate slices: 8 (one per XCD) while the query groups alone give few
// waves; fewer for very large n, where they bound the list workspace.
static int knn_slices(int n) {
    return n <= 400000 ? 8 : n <= 800000 ? 4 : n <= 1600000 ? 2 : 1;
}

// Dense Gaussian affinity of geometry.manifold_graph(typ='affinity')
// (geometry.py:114-118): A[i][j] = exp(-|x_i - x_j|^2 / (2 sigma^2)), the
// squared distance in the fixed operation order of d2_direct.  16 x 16 tiles,
// the two point blocks staged in LDS.
__global__ __launch_bounds__(256) void affinity_kernel(const double* __restrict__ x, int n, int D,
                                                       double inv2s2, double* __restrict__ A,
                                                       int lda) {
    __shared__ double xi[16][65], xj[16][65];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int i0 = blockIdx.y * 16, j0 = blockIdx.x * 16;
    for (int e = threadIdx.x; e < 16 * D; e += 256) {
        const int r = e / D, c = e - r * D;
        xi[r][c] = i0 + r < n ? x[(size_t)(i0 + r) * D + c] : 0.0;
        xj[r][c] = j0 + r < n ? x[(size_t)(j0 + r) * D + c] : 0.0;
    }
    __syncthreads();
    const int i = i0 + ty, j = j0 + tx;
    if (i >= n || j >= n) return;
    const double d2 = d2_direct(&xi[ty][0], &xj[tx][0], D);
    A[(size_t)i * lda + j] = exp(-d2 * inv2s2);
}

}  // namespace rvgp

using namespace rvgp;

extern "C" int rvgp_affinity(const double* x, int n, int D, double sigma, double* A, int lda,
                             void* stream) {
    RVGP_CHECK_ARG(n >= 1 && D >= 1 && D <= 64 && lda >= n, "rvgp_affinity: need D <= 64, lda >= n");
    RVGP_CHECK_ARG(sigma > 0.0, "rvgp_affinity: sigma must be positive");
    hipLaunchKernelGGL(affinity_kernel, dim3(cdiv(n, 16), cdiv(n, 16)), dim3(256), 0,
                       as_stream(stream), x, n, D, 1.0 / (2.0 * sigma * sigma), A, lda);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}

extern "C" size_t rvgp_knn_ws_bytes(int n, int k) {
    if (n <= 0 || k <= 0) return 0;
    const size_t lists = (size_t)knn_slices(n) * k * n;
    const size_t thr = n > kKnnBoundCands ? (size_t)n : 0;
    return align_up(sizeof(double) * thr, 256) + align_up(sizeof(double) * lists, 256) +
           sizeof(int) * lists;
}

extern "C" int rvgp_knn(const double* x, int n, int D, int k, const int* qperm, int* nbr,
                        void* ws, size_t ws_bytes, void* stream) {
    RVGP_CHECK_ARG(k > 0 && k < n, "rvgp_knn: need 0 < k < n (k=%d, n=%d)", k, n);
    RVGP_CHECK_ARG(k <= 64, "rvgp_knn: k <= 64 supported");
    RVGP_CHECK_ARG(D >= 1 && D <= 64, "rvgp_knn: 1 <= D <= 64 supported");
    RVGP_CHECK_ARG(ws != nullptr && ws_bytes >= rvgp_knn_ws_bytes(n, k),
                   "rvgp_knn: workspace too small");
    hipStream_t st = as_stream(stream);
    const int S = knn_slices(n);
    const size_t lists = (size_t)S * k * n;
    char* p = static_cast<char*>(ws);
    double* thr = nullptr;
    if (n > kKnnBoundCands) {
        thr = reinterpret_cast<double*>(p);
        p += align_up(sizeof(double) * (size_t)n, 256);
    }
    double* ld = reinterpret_cast<double*>(p);
    int* li = reinterpret_cast<int*>(p + align_up(sizeof(double) * lists, 256));
    if (thr != nullptr) {
        // bounding pass: each query's k-th squared distance among the first
        // kKnnBoundCands points (>= its k-th over all points)
        launch_knn(st, x, n, D, k, kKnnBoundCands, S, qperm, nullptr, ld, li, thr, nullptr);
        RVGP_LAUNCH_CHECK();
    }
    launch_knn(st, x, n, D, k, n, S, qperm, thr, ld, li, nullptr, nbr);
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
