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

namespace rvgp {

enum : int { kScanned = 0, kNotInHeap = 1, kInHeap = 2 };
constexpr int kMaxRank = 32;

// 32 bytes, aligned: a node never straddles two cache lines
struct __align__(32) HNode {
    double val;
    int gidx, parent, left, right, children;
    short rank, state;
};

// The root list is the sibling list holding min_node.  Its two ends are kept
// in head / tail, so appending a root and finding the leftmost root are O(1)
// instead of walks over the list: the reference walks it (pyx:526-536,
// 642-660), but the walks only locate its ends, and the list itself -- every
// link, every order -- is the same.
struct Heap {
    HNode* N;
    int* roots;
    int min_node;
    int head, tail;
};

__device__ __forceinline__ int rightmost_sibling(HNode* N, int a) {
    while (N[a].right != -1) a = N[a].right;
    return a;
}

__device__ __forceinline__ int leftmost_sibling(HNode* N, int a) {
    while (N[a].left != -1) a = N[a].left;
    return a;
}

__device__ __forceinline__ void add_sibling(HNode* N, int a, int s) {  // pyx:526-536
    int t = rightmost_sibling(N, a);
    N[t].right = s;
    N[s].left = t;
    N[s].right = -1;
    N[s].parent = N[a].parent;
    if (N[s].parent != -1) N[N[s].parent].rank += 1;
}

__device__ __forceinline__ void add_child(HNode* N, int a, int c) {  // pyx:507-520
    N[c].parent = a;
    if (N[a].children != -1) {
        add_sibling(N, N[a].children, c);
    } else {
        N[a].children = c;
        N[c].right = -1;
        N[c].left = -1;
        N[a].rank = 1;
    }
}

__device__ __forceinline__ void node_remove(Heap& h, int a) {  // pyx:542-560
    HNode* N = h.N;
    int p = N[a].parent;
    if (p == -1) {  // a root: keep the list's ends
        if (a == h.head) h.head = N[a].right;
        if (a == h.tail) h.tail = N[a].left;
    } else {
        N[p].rank -= 1;
        if (N[a].left != -1) N[p].children = N[a].left;
        else if (N[a].right != -1) N[p].children = N[a].right;
        else N[p].children = -1;
    }
    if (N[a].left != -1) N[N[a].left].right = N[a].right;
    if (N[a].right != -1) N[N[a].right].left = N[a].left;
    N[a].left = -1;
    N[a].right = -1;
    N[a].parent = -1;
}

// add_sibling(min_node, s) for the root list (pyx:526-536): s goes after
// the rightmost root; a root's parent is -1, so no rank changes.
__device__ __forceinline__ void root_append(Heap& h, int s) {
    HNode* N = h.N;
    N[h.tail].right = s;
    N[s].left = h.tail;
    N[s].right = -1;
    N[s].parent = -1;
    h.tail = s;
}

__device__ __forceinline__ void insert_node(Heap& h, int a) {  // pyx:579-589
    if (h.min_node != -1) {
        root_append(h, a);
        if (h.N[a].val < h.N[h.min_node].val) h.min_node = a;
    } else {
        h.min_node = a;
        h.head = h.tail = a;
    }
}

__device__ __forceinline__ void decrease_val(Heap& h, int a, double v) {  // pyx:595-608
    HNode* N = h.N;
    N[a].val = v;
    if (N[a].parent != -1 && N[N[a].parent].val >= v) {
        node_remove(h, a);
        insert_node(h, a);
    } else if (N[h.min_node].val > N[a].val) {
        h.min_node = a;
    }
}

// pyx:614-636 (tail recursion as a loop).  Returns false on rank overflow.
__device__ __forceinline__ bool heap_link(Heap& h, int a) {
    HNode* N = h.N;
    for (;;) {
        int r = N[a].rank;
        if (r >= kMaxRank) return false;
        if (h.roots[r] == -1) {
            h.roots[r] = a;
            return true;
        }
        int ln = h.roots[r];
        h.roots[r] = -1;
        if (N[a].val < N[ln].val || a == h.min_node) {
            node_remove(h, ln);
            add_child(N, a, ln);
        } else {
            node_remove(h, a);
            add_child(N, ln, a);
            a = ln;
        }
    }
}

// pyx:642-690.  The reference clears roots_by_rank[0..100); ranks here are
// bounded by kMaxRank (checked), so clearing [0, kMaxRank) is equivalent.
__device__ __forceinline__ int remove_min(Heap& h, bool& ok) {
    HNode* N = h.N;
    int m = h.min_node, t, tr;
    if (N[m].children != -1) {
        t = leftmost_sibling(N, N[m].children);
        while (t != -1) {
            tr = N[t].right;
            node_remove(h, t);
            root_append(h, t);  // add_sibling(m, t): m is a root
            t = tr;
        }
        N[m].children = -1;
    }
    t = h.head;  // leftmost_sibling(m)
    if (t == m) {
        if (N[m].right != -1) {
            t = N[m].right;
        } else {
            h.min_node = -1;
            h.head = h.tail = -1;
            return m;
        }
    }
    node_remove(h, m);
    h.min_node = t;
    for (int i = 0; i < kMaxRank; ++i) h.roots[i] = -1;
    while (t != -1) {
        if (N[t].val < N[h.min_node].val) h.min_node = t;
        tr = N[t].right;
        if (!heap_link(h, t)) { ok = false; return m; }
        t = tr;
    }
    return m;
}

__device__ __forceinline__ unsigned hash_of(int g, unsigned mask) {
    return ((unsigned)g * 2654435761u) & mask;
}

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
    HNode* N = reinterpret_cast<HNode*>(base);
    // hash table entries (node id, pool slot) side by side: one probe, one line
    int2* tab = reinterpret_cast<int2*>(N + pool);
    int* hposv = reinterpret_cast<int*>(tab + hsize);  // table position of each slot
    int* roots = hposv + pool;
    const unsigned mask = (unsigned)hsize - 1u;
    for (int q = 0; q < hsize; ++q) tab[q] = make_int2(-1, -1);
    const int Kp1 = K + 1;

    for (int si = t; si < nsrc; si += batch) {
        const int src = src_list != nullptr ? src_list[1 + si] : si;
        int used = 0;
        bool ok = true;
        auto new_node = [&](int g, unsigned hp) -> int {
            int s = used++;
            HNode& nd = N[s];
            nd.val = 0.0;
            nd.gidx = g;
            nd.rank = 0;
            nd.state = kNotInHeap;
            nd.parent = nd.left = nd.right = nd.children = -1;
            hposv[s] = (int)hp;
            tab[hp] = make_int2(g, s);
            return s;
        };
        Heap h{N, roots, -1, -1, -1};
        {
            unsigned hp = hash_of(src, mask);
            while (tab[hp].x != -1) hp = (hp + 1) & mask;
            insert_node(h, new_node(src, hp));
        }
        int scanned = 0;
        int* out = nbh + (size_t)src * Kp1;
        while (h.min_node != -1 && scanned <= K) {
            int v = remove_min(h, ok);
            if (!ok) break;
            N[v].state = kScanned;
            out[scanned++] = N[v].gidx;
            if (scanned <= K) {
                const int j = N[v].gidx;
                const double vv = N[v].val;
                for (int e = rowptr[j]; e < rowptr[j + 1]; ++e) {
                    const int c = cols[e];
                    unsigned hp = hash_of(c, mask);
                    int sc = -1;
                    for (int2 te = tab[hp]; te.x != -1; te = tab[hp]) {
                        if (te.x == c) { sc = te.y; break; }
                        hp = (hp + 1) & mask;
                    }
                    if (sc != -1 && N[sc].state == kScanned) continue;
                    const double nv = vv + w[e];
                    if (sc == -1 || N[sc].state == kNotInHeap) {
                        if (sc == -1) {
                            if (used >= pool) { ok = false; break; }
                            sc = new_node(c, hp);
                        }
                        N[sc].state = kInHeap;
                        N[sc].val = nv;
                        insert_node(h, sc);
                    } else if (N[sc].val > nv) {
                        decrease_val(h, sc, nv);
                    }
                }
                if (!ok) break;
            }
        }
        for (int s = 0; s < used; ++s) tab[hposv[s]] = make_int2(-1, -1);
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
