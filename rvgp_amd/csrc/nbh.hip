// a3 (part 1)  Geodesic neighbourhoods: the truncated Dijkstra of
// RVGP/lib/ptu_dijkstra.pyx:361-394 driven by the Fibonacci heap of
// pyx:440-690.
//
// Which K+1 nodes form a neighbourhood is decided by the heap's pop order
// among equal hop distances, so the heap is emulated operation for
// operation: the same sibling-list splices, the same linking order, the same
// choice of the new minimum.  One lane owns one source.  Instead of the
// reference's N heap nodes per source (re-initialised N times, an O(N^2)
// total), a lane keeps a private pool of only the nodes its search touches,
// found through a small open-addressing hash table keyed by node id; node
// identity (the reference compares pointers) becomes pool-slot identity.
// Lanes stride over sources, so `batch` lanes handle any n.
#include "common.h"
#include "nbh_heap.h"

namespace rvgp {

__global__ __launch_bounds__(256) void geo_nbh_kernel(const int* __restrict__ rowptr,
                                                      const int* __restrict__ cols,
                                                      const double* __restrict__ w, int n, int K,
                                                      int batch, int pool, int hsize,
                                                      size_t lane_bytes, int* __restrict__ nbh,
                                                      char* __restrict__ ws,
                                                      const int* __restrict__ src_list,
                                                      int* __restrict__ ovf_list,
                                                      int* __restrict__ status) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    // sources: all n, or src_list[1 .. src_list[0]] (the overflow list of a
    // pass with a smaller pool)
    const int nsrc = src_list != nullptr ? src_list[0] : n;
    if (t >= batch || t >= nsrc) return;
    char* base = ws + (size_t)t * lane_bytes;
    NbhLane L;
    L.N = reinterpret_cast<HNode*>(base);
    // hash table entries (node id, pool slot) side by side: one probe, one line
    L.tab = reinterpret_cast<int2*>(L.N + pool);
    L.hposv = reinterpret_cast<int*>(L.tab + hsize);  // table position of each slot
    L.roots = L.hposv + pool;
    L.pool = pool;
    L.mask = (unsigned)hsize - 1u;
    for (int q = 0; q < hsize; ++q) L.tab[q] = make_int2(-1, -1);
    const int Kp1 = K + 1;

    for (int si = t; si < nsrc; si += batch) {
        const int src = src_list != nullptr ? src_list[1 + si] : si;
        int scanned = 0;
        const bool ok = nbh_search(rowptr, cols, w, K, src, L, nbh + (size_t)src * Kp1, &scanned);
        if (!ok) {
            if (ovf_list != nullptr) {
                // redone by the pass with the worst-case pool
                ovf_list[1 + atomicAdd(&ovf_list[0], 1)] = src;
                continue;
            }
            atomicMax(status, RVGP_ST_POOL_OVERFLOW);
            return;
        }
        if (scanned < Kp1) atomicMax(status, RVGP_ST_SHORT_COMPONENT);
    }
}

static int next_pow2(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

static size_t lane_bytes_for(int pool) {
    const int hsize = next_pow2(2 * pool);
    size_t b = sizeof(HNode) * (size_t)pool + sizeof(int2) * (size_t)hsize +
               sizeof(int) * ((size_t)pool + kMaxRank);
    return (b + 255) & ~(size_t)255;
}

__global__ void nbh_clear_count_kernel(int* __restrict__ c) {
    if (threadIdx.x == 0) c[0] = 0;
}

// First pass: every source with a pool of kNbhPool nodes (a search on a
// kNN graph touches a few hundred); the few sources whose search outgrows it
// are listed and redone with the worst-case pool 1 + K * max_degree by
// kNbhBigLanes lanes.  Both passes run the same heap operations, so which
// pass finishes a source does not change its neighbourhood; the small pool
// keeps each lane's state at ~27 KB instead of ~80 KB, which is what the
// latency-bound heap walks spend their time fetching.
constexpr int kNbhPool = 512;
constexpr int kNbhBigLanes = 2048;

static void nbh_plan(int n, int batch, int pool, int& p1, int& b2, size_t& o1, size_t& o2,
                     size_t& o3, size_t& total) {
    p1 = pool < kNbhPool ? pool : kNbhPool;
    b2 = batch < kNbhBigLanes ? batch : kNbhBigLanes;
    const size_t a = lane_bytes_for(p1) * (size_t)batch;
    const size_t b = p1 < pool ? lane_bytes_for(pool) * (size_t)b2 : 0;
    const size_t c = p1 < pool ? sizeof(int) * ((size_t)n + 1) : 0;
    o1 = 0;
    o2 = a;
    o3 = a + b;
    total = a + b + ((c + 255) & ~(size_t)255);
}

}  // namespace rvgp

using namespace rvgp;

extern "C" size_t rvgp_geo_nbh_ws_bytes(int n, int batch, int pool) {
    if (n < 1 || batch < 1 || pool < 1) return 0;
    int p1, b2;
    size_t o1, o2, o3, total;
    nbh_plan(n, batch, pool, p1, b2, o1, o2, o3, total);
    return total;
}

extern "C" int rvgp_geo_nbh(const int* rowptr, const int* cols, const double* w, int n, int K,
                            int batch, int pool, int* nbh, void* ws, size_t ws_bytes, int* status,
                            void* stream) {
    RVGP_CHECK_ARG(n >= 1 && K >= 0, "rvgp_geo_nbh: need n >= 1, K >= 0");
    RVGP_CHECK_ARG(K < n, "Geodesic neighborhood size must be less than the total number of samples");
    RVGP_CHECK_ARG(batch >= 1 && pool >= K + 1, "rvgp_geo_nbh: bad batch/pool");
    RVGP_CHECK_ARG(ws_bytes >= rvgp_geo_nbh_ws_bytes(n, batch, pool),
                   "rvgp_geo_nbh: workspace too small");
    int p1, b2;
    size_t o1, o2, o3, total;
    nbh_plan(n, batch, pool, p1, b2, o1, o2, o3, total);
    char* base = static_cast<char*>(ws);
    hipStream_t st = as_stream(stream);
    const int lanes = batch < n ? batch : n;
    int* ovf = p1 < pool ? reinterpret_cast<int*>(base + o3) : nullptr;
    if (ovf != nullptr) {
        hipLaunchKernelGGL(nbh_clear_count_kernel, dim3(1), dim3(64), 0, st, ovf);
        RVGP_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(geo_nbh_kernel, dim3(cdiv(lanes, 256)), dim3(256), 0, st, rowptr, cols, w,
                       n, K, lanes, p1, next_pow2(2 * p1), lane_bytes_for(p1), nbh, base + o1,
                       nullptr, ovf, status);
    RVGP_LAUNCH_CHECK();
    if (ovf != nullptr) {
        const int l2 = b2 < n ? b2 : n;
        hipLaunchKernelGGL(geo_nbh_kernel, dim3(cdiv(l2, 256)), dim3(256), 0, st, rowptr, cols, w,
                           n, K, l2, pool, next_pow2(2 * pool), lane_bytes_for(pool), nbh,
                           base + o2, ovf, nullptr, status);
        RVGP_LAUNCH_CHECK();
    }
    return RVGP_OK;
}
