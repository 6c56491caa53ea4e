// Fibonacci-heap emulation of RVGP/lib/ptu_dijkstra.pyx:440-690 and the
// per-source truncated Dijkstra of pyx:361-394, shared by the neighbourhood
// kernel (nbh.hip, one lane per source) and the host build that runs the same
// code under AddressSanitizer / UndefinedBehaviorSanitizer
// (tests/sanitize/nbh_host.hip).  Every function compiles for both sides.
#ifndef RVGP_NBH_HEAP_H
#define RVGP_NBH_HEAP_H

#include <hip/hip_runtime.h>

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

__host__ __device__ inline int rightmost_sibling(HNode* N, int a) {
    while (N[a].right != -1) a = N[a].right;
    return a;
}

__host__ __device__ inline int leftmost_sibling(HNode* N, int a) {
    while (N[a].left != -1) a = N[a].left;
    return a;
}

__host__ __device__ inline void add_sibling(HNode* N, int a, int s) {  // pyx:526-536
    int t = rightmost_sibling(N, a);
    N[t].right = s;
    N[s].left = t;
    N[s].right = -1;
    N[s].parent = N[a].parent;
    if (N[s].parent != -1) N[N[s].parent].rank += 1;
}

__host__ __device__ inline void add_child(HNode* N, int a, int c) {  // pyx:507-520
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

__host__ __device__ inline void node_remove(Heap& h, int a) {  // pyx:542-560
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
__host__ __device__ inline void root_append(Heap& h, int s) {
    HNode* N = h.N;
    N[h.tail].right = s;
    N[s].left = h.tail;
    N[s].right = -1;
    N[s].parent = -1;
    h.tail = s;
}

__host__ __device__ inline void insert_node(Heap& h, int a) {  // pyx:579-589
    if (h.min_node != -1) {
        root_append(h, a);
        if (h.N[a].val < h.N[h.min_node].val) h.min_node = a;
    } else {
        h.min_node = a;
        h.head = h.tail = a;
    }
}

__host__ __device__ inline void decrease_val(Heap& h, int a, double v) {  // pyx:595-608
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
__host__ __device__ inline bool heap_link(Heap& h, int a) {
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
__host__ __device__ inline int remove_min(Heap& h, bool& ok) {
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

__host__ __device__ inline unsigned hash_of(int g, unsigned mask) {
    return ((unsigned)g * 2654435761u) & mask;
}


// One lane's private state: `pool` heap nodes, an open-addressing table of
// 2^m >= 2 pool entries (node id, pool slot), each slot's table position and
// the roots-by-rank array.  The table must hold (-1, -1) everywhere on entry;
// nbh_search leaves it that way.
struct NbhLane {
    HNode* N;
    int2* tab;
    int* hposv;
    int* roots;
    int pool;
    unsigned mask;
};

// The K+1 nodes popped by the truncated Dijkstra from `src` (pyx:361-394),
// written to out[0 .. *scanned).  Returns false when the search needs more
// than `pool` nodes (out is then incomplete).
__host__ __device__ inline bool nbh_search(const int* rowptr, const int* cols, const double* w,
                                           int K, int src, const NbhLane& L, int* out,
                                           int* scanned_out) {
    HNode* N = L.N;
    int2* tab = L.tab;
    const unsigned mask = L.mask;
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
        L.hposv[s] = (int)hp;
        tab[hp] = make_int2(g, s);
        return s;
    };
    Heap h{N, L.roots, -1, -1, -1};
    {
        unsigned hp = hash_of(src, mask);
        while (tab[hp].x != -1) hp = (hp + 1) & mask;
        insert_node(h, new_node(src, hp));
    }
    int scanned = 0;
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
                        if (used >= L.pool) { ok = false; break; }
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
    for (int s = 0; s < used; ++s) tab[L.hposv[s]] = make_int2(-1, -1);
    *scanned_out = scanned;
    return ok;
}

}  // namespace rvgp

#endif  // RVGP_NBH_HEAP_H
