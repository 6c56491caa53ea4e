/*
 * CPU restatement of the integer / index parts of RVGP's geometry path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: it is built into
 * oracle/liboracle.so and may be loaded only by tests/, by the smoke check
 * and by the CPU-baseline leg of bench.py.  The product (rvgp_amd) never
 * links, loads or calls it.
 *
 * Functions and the reference code they restate:
 *   oracle_fps        RVGP/geometry.py:126-162   (furthest_point_sampling)
 *   oracle_diam2      RVGP/geometry.py:142-144   (D.max() of pairwise_distances)
 *   oracle_knn        RVGP/geometry.py:100-110   (kneighbors_graph, include_self=False)
 *   oracle_sym_csr    RVGP/geometry.py:111-112 + RVGP/lib/ptu_dijkstra.pyx:84-103
 *                     (A+I, networkx undirected graph, symmetrised sorted CSR)
 *   oracle_geo_nbh    RVGP/lib/ptu_dijkstra.pyx:361-404 (truncated Dijkstra
 *                     driven by the Fibonacci heap of pyx:440-690, then the
 *                     centring of the neighbourhood)
 *
 * Distance formula.  The reference ranks points by sklearn's GEMM-identity
 * distance (|x|^2+|y|^2-2x.y), whose rounding depends on the BLAS build.  This
 * restatement and the HIP kernels both use the direct form
 *     d2(a,b) = fma(t_{D-1},t_{D-1}, ... fma(t_1,t_1, fma(t_0,t_0,0)))
 *     with t_k = a_k - b_k, summed in increasing k,
 * and d = sqrt(d2), so the two agree bit for bit.  Ties in argmax / top-k are
 * broken towards the lower index (numpy.argmax returns the first maximum).
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static inline double d2_direct(const double *a, const double *b, int D) {
    double acc = 0.0;
    for (int k = 0; k < D; ++k) {
        double t = a[k] - b[k];
        acc = fma(t, t, acc);
    }
    return acc;
}

/* max over all pairs of d2 (the square of D.max(), geometry.py:144). */
double oracle_diam2(const double *x, int n, int D) {
    double m = 0.0;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            double v = d2_direct(x + (size_t)i * D, x + (size_t)j * D, D);
            if (v > m) m = v;
        }
    return m;
}

/*
 * geometry.py:126-162.  N < 0 means "N is None": run until
 * lambdas[i]/diam < spacing and drop that i (lines 156-160).  Returns the
 * number of selected points.  perm/lambdas must hold n (N<0) or N entries.
 */
int oracle_fps(const double *x, int n, int D, int N, double spacing,
               int start_idx, int32_t *perm, double *lambdas) {
    int nsel = (N < 0) ? n : N;
    double diam = 0.0;
    if (N < 0) diam = sqrt(oracle_diam2(x, n, D));
    double *ds = (double *)malloc(sizeof(double) * (size_t)n);
    if (!ds) return -3;
    const double *xs = x + (size_t)start_idx * D;
    for (int j = 0; j < n; ++j) ds[j] = sqrt(d2_direct(xs, x + (size_t)j * D, D));
    perm[0] = start_idx;
    lambdas[0] = 0.0;
    int out = nsel;
    for (int i = 1; i < nsel; ++i) {
        int idx = 0;
        double best = ds[0];
        for (int j = 1; j < n; ++j)
            if (ds[j] > best) { best = ds[j]; idx = j; }
        perm[i] = idx;
        lambdas[i] = best;
        const double *xi = x + (size_t)idx * D;
        for (int j = 0; j < n; ++j) {
            double v = sqrt(d2_direct(xi, x + (size_t)j * D, D));
            if (v < ds[j]) ds[j] = v;
        }
        if (N < 0 && lambdas[i] / diam < spacing) { out = i; break; }
    }
    free(ds);
    return out;
}

/*
 * k nearest neighbours of every point, self excluded, ordered by
 * (d2 ascending, index ascending).  nbr is (n, k) row-major.
 */
int oracle_knn(const double *x, int n, int D, int k, int32_t *nbr) {
    if (k <= 0 || k >= n) return -1;
    double *bd = (double *)malloc(sizeof(double) * (size_t)k);
    int32_t *bi = (int32_t *)malloc(sizeof(int32_t) * (size_t)k);
    if (!bd || !bi) { free(bd); free(bi); return -3; }
    for (int i = 0; i < n; ++i) {
        int cnt = 0;
        const double *xi = x + (size_t)i * D;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            double v = d2_direct(xi, x + (size_t)j * D, D);
            if (cnt == k && !(v < bd[k - 1])) continue; /* j increases: equal d2 loses */
            int p = (cnt < k) ? cnt++ : k - 1;
            while (p > 0 && bd[p - 1] > v) { bd[p] = bd[p - 1]; bi[p] = bi[p - 1]; --p; }
            bd[p] = v;
            bi[p] = j;
        }
        memcpy(nbr + (size_t)i * k, bi, sizeof(int32_t) * (size_t)k);
    }
    free(bd);
    free(bi);
    return 0;
}

static int cmp_i32(const void *a, const void *b) {
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/*
 * Symmetrised CSR of (kNN connectivity + I): row i holds the sorted, unique
 * set {i} u nbr(i) u {j : i in nbr(j)}.  This is the pattern that
 * ptu_dijkstra.pyx:99-103 iterates over (networkx turns A+I into an undirected
 * graph with unit weights, geometry.py:111-112).  rowptr has n+1 entries;
 * cols must hold n*(2k+1) entries.  Returns nnz.
 */
int oracle_sym_csr(const int32_t *nbr, int n, int k, int32_t *rowptr, int32_t *cols) {
    int32_t *cnt = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    if (!cnt) return -3;
    for (int i = 0; i < n; ++i) {
        cnt[i] += k + 1;
        for (int q = 0; q < k; ++q) cnt[nbr[(size_t)i * k + q]] += 1;
    }
    int32_t *start = (int32_t *)malloc(sizeof(int32_t) * ((size_t)n + 1));
    int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * (size_t)n * (2 * (size_t)k + 1));
    int32_t *fill = (int32_t *)calloc((size_t)n, sizeof(int32_t));
    if (!start || !tmp || !fill) { free(cnt); free(start); free(tmp); free(fill); return -3; }
    start[0] = 0;
    for (int i = 0; i < n; ++i) start[i + 1] = start[i] + cnt[i];
    for (int i = 0; i < n; ++i) {
        tmp[start[i] + fill[i]++] = i;
        for (int q = 0; q < k; ++q) {
            int j = nbr[(size_t)i * k + q];
            tmp[start[i] + fill[i]++] = j;
            tmp[start[j] + fill[j]++] = i;
        }
    }
    int nnz = 0;
    rowptr[0] = 0;
    for (int i = 0; i < n; ++i) {
        int32_t *seg = tmp + start[i];
        int len = start[i + 1] - start[i];
        qsort(seg, (size_t)len, sizeof(int32_t), cmp_i32);
        for (int p = 0; p < len; ++p)
            if (p == 0 || seg[p] != seg[p - 1]) cols[nnz++] = seg[p];
        rowptr[i + 1] = nnz;
    }
    free(cnt); free(start); free(tmp); free(fill);
    return nnz;
}

/* ------------------------------------------------------------------------
 * Fibonacci heap, restated from ptu_dijkstra.pyx:440-690.  Nodes are the
 * entries of one array indexed by graph node id; "NULL" is -1.  Every
 * operation below performs the same link updates, in the same order, as the
 * pointer version it restates, so the pop order (including the order among
 * equal keys) is the same.
 * ------------------------------------------------------------------------ */
enum { ST_SCANNED = 0, ST_NOT_IN_HEAP = 1, ST_IN_HEAP = 2 };

typedef struct {
    int32_t rank, state;
    double val;
    int32_t parent, left, right, children;
} fnode;

typedef struct {
    int32_t min_node;
    int32_t roots_by_rank[100];
} fheap;

static void fn_init(fnode *nd) { /* pyx:465-479 */
    nd->val = 0.0;
    nd->rank = 0;
    nd->state = ST_NOT_IN_HEAP;
    nd->parent = nd->left = nd->right = nd->children = -1;
}

static int32_t rightmost_sibling(fnode *N, int32_t a) { /* pyx:485-490 */
    while (N[a].right != -1) a = N[a].right;
    return a;
}

static int32_t leftmost_sibling(fnode *N, int32_t a) { /* pyx:496-501 */
    while (N[a].left != -1) a = N[a].left;
    return a;
}

static void add_sibling(fnode *N, int32_t a, int32_t s) { /* pyx:526-536 */
    int32_t t = rightmost_sibling(N, a);
    N[t].right = s;
    N[s].left = t;
    N[s].right = -1;
    N[s].parent = N[a].parent;
    if (N[s].parent != -1) N[N[s].parent].rank += 1;
}

static void add_child(fnode *N, int32_t a, int32_t c) { /* pyx:507-520 */
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

static void fn_remove(fnode *N, int32_t a) { /* pyx:542-560 */
    int32_t p = N[a].parent;
    if (p != -1) {
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

static void insert_node(fheap *h, fnode *N, int32_t a) { /* pyx:579-589 */
    if (h->min_node != -1) {
        add_sibling(N, h->min_node, a);
        if (N[a].val < N[h->min_node].val) h->min_node = a;
    } else {
        h->min_node = a;
    }
}

static void decrease_val(fheap *h, fnode *N, int32_t a, double v) { /* pyx:595-608 */
    N[a].val = v;
    if (N[a].parent != -1 && N[N[a].parent].val >= v) {
        fn_remove(N, a);
        insert_node(h, N, a);
    } else if (N[h->min_node].val > N[a].val) {
        h->min_node = a;
    }
}

static void fh_link(fheap *h, fnode *N, int32_t a) { /* pyx:614-636, recursion as a loop */
    for (;;) {
        int32_t r = N[a].rank;
        if (h->roots_by_rank[r] == -1) {
            h->roots_by_rank[r] = a;
            return;
        }
        int32_t ln = h->roots_by_rank[r];
        h->roots_by_rank[r] = -1;
        if (N[a].val < N[ln].val || a == h->min_node) {
            fn_remove(N, ln);
            add_child(N, a, ln);
        } else {
            fn_remove(N, a);
            add_child(N, ln, a);
            a = ln;
        }
    }
}

static int32_t remove_min(fheap *h, fnode *N) { /* pyx:642-690 */
    int32_t m = h->min_node, t, tr, out;
    if (N[m].children != -1) {
        t = leftmost_sibling(N, N[m].children);
        while (t != -1) {
            tr = N[t].right;
            fn_remove(N, t);
            add_sibling(N, m, t);
            t = tr;
        }
        N[m].children = -1;
    }
    t = leftmost_sibling(N, m);
    if (t == m) {
        if (N[m].right != -1) {
            t = N[m].right;
        } else {
            out = m;
            h->min_node = -1;
            return out;
        }
    }
    out = m;
    fn_remove(N, m);
    h->min_node = t;
    for (int i = 0; i < 100; ++i) h->roots_by_rank[i] = -1;
    while (t != -1) {
        if (N[t].val < N[h->min_node].val) h->min_node = t;
        tr = N[t].right;
        fh_link(h, N, t);
        t = tr;
    }
    return out;
}

/*
 * ptu_dijkstra.pyx:361-404.  For every source i, pop K+1 nodes from the
 * heap-driven truncated Dijkstra over the CSR graph with edge weights w
 * (all 1.0 on the RVGP path), in pop order, into nbh (n, K+1).  If centred is
 * not NULL, also write the neighbourhood coordinates minus their mean,
 * laid out (n, D, K+1), the mean summed in pop order as in pyx:397-404.
 *
 * The reference resets all N heap nodes per source (pyx:364-365); only the
 * nodes the search touched can differ from the initial state, so only those
 * are reset here.  Returns 0, or -(i+1)-1000 when source i reaches fewer than
 * K+1 nodes (the reference then reads indices left over from the previous
 * source; the restatement reports it instead).
 */
int oracle_geo_nbh(const double *x, int n, int D, const int32_t *rowptr,
                   const int32_t *cols, const double *w, int K, int32_t *nbh,
                   double *centred) {
    fnode *N = (fnode *)malloc(sizeof(fnode) * (size_t)n);
    int32_t *touched = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    if (!N || !touched) { free(N); free(touched); return -3; }
    for (int a = 0; a < n; ++a) fn_init(&N[a]);
    fheap h;
    int Kp1 = K + 1;
    for (int i = 0; i < n; ++i) {
        int ntouched = 0;
        h.min_node = -1;
        touched[ntouched++] = i;
        insert_node(&h, N, i);
        int scanned = 0;
        int32_t *out = nbh + (size_t)i * Kp1;
        while (h.min_node != -1 && scanned <= K) {
            int32_t v = remove_min(&h, N);
            N[v].state = ST_SCANNED;
            out[scanned++] = v;
            if (scanned <= K) {
                for (int e = rowptr[v]; e < rowptr[v + 1]; ++e) {
                    int32_t c = cols[e];
                    if (N[c].state != ST_SCANNED) {
                        double nv = N[v].val + w[e];
                        if (N[c].state == ST_NOT_IN_HEAP) {
                            touched[ntouched++] = c;
                            N[c].state = ST_IN_HEAP;
                            N[c].val = nv;
                            insert_node(&h, N, c);
                        } else if (N[c].val > nv) {
                            decrease_val(&h, N, c, nv);
                        }
                    }
                }
            }
        }
        for (int t = 0; t < ntouched; ++t) fn_init(&N[touched[t]]);
        if (scanned < Kp1) { free(N); free(touched); return -1000 - (i + 1); }
        if (centred) {
            double *cm = centred + (size_t)i * D * Kp1;
            for (int p = 0; p < D; ++p) {
                double mean = 0.0;
                for (int q = 0; q < Kp1; ++q) {
                    cm[(size_t)p * Kp1 + q] = x[(size_t)out[q] * D + p];
                    mean += cm[(size_t)p * Kp1 + q];
                }
                mean = mean / Kp1;
                for (int q = 0; q < Kp1; ++q) cm[(size_t)p * Kp1 + q] -= mean;
            }
        }
    }
    free(N);
    free(touched);
    return 0;
}
