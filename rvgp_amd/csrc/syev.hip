// Symmetric eigendecomposition A = V diag(w) V^T (FP64, w ascending) of the
// small projected matrices of the eigensolver's Rayleigh-Ritz step -- the
// dense eigenproblem ARPACK solves inside scipy.sparse.linalg.eigsh, reached
// from RVGP/geometry.py:73.
//
// Three stages, all on the device:
//  1. Householder tridiagonalisation Q^T A Q = T (LAPACK sytd2, lower), one
//     kernel launch per column on up to 64 workgroups that each own a cyclic
//     slab of rows: a launch finishes the previous column's rank-2 update on
//     its rows while forming the next reflector, so the launch boundary is
//     the only synchronisation between workgroups (no device-scope spinning,
//     which on gfx950 means writing back and invalidating L2 that a
//     concurrent stream is using).
//  2. Eigenvalues of T by multisection (one wavefront per eigenvalue, its 64
//     lanes evaluating Sturm counts at 64 interior points, ~9 rounds to full
//     precision); eigenvectors of T by inverse iteration, one wavefront per
//     eigenvalue, with modified Gram-Schmidt inside clusters of eigenvalues
//     closer than 1e-7 |T| (elsewhere the vectors are orthogonal to rounding
//     over the gap).
//  3. V = Q Z: the reflectors are applied in blocks of NB in compact WY form
//     (I - Vb Tb Vb^T) by one kernel, each workgroup carrying a few columns
//     of Z in LDS through all the blocks.
#include "common.h"

namespace rvgp {

constexpr int kSyevNB = 32;        // reflectors per compact-WY block
constexpr int kSyevMaxN = 2048;   // inverse iteration holds 5.5 n doubles in LDS

// Column j of the tridiagonalisation (LAPACK sytd2, lower), one launch per
// column on G workgroups; workgroup g owns the rows r = g (mod G) of the
// work matrix W (full symmetric storage).  The launch boundary is the only
// synchronisation between workgroups.  Launch j
//   1. finishes column j-1: w = p - (tau/2)(p^T v) v from the previous
//      launch's p (each owner wrote its rows) and partial sums of p^T v;
//   2. forms row j of the updated matrix W - v w^T - w v^T and from it the
//      reflector v_j, tau_j, d_j, e_j -- redundantly in every workgroup (the
//      same arithmetic in the same order, so every copy is identical);
//   3. applies column j-1's rank-2 update to its own rows r > j (row j is
//      consumed by step 2 and never read again) and, in the same pass,
//      p_r = tau_j W'[r][j+1:] v_j and its share of p^T v.
// Vectors of column j are indexed from row j+1; buffers alternate by parity.
constexpr int kStepThreads = 256;
constexpr int kStepGroups = 64;
constexpr int kStepGroupsMax = 128;
constexpr int kStepChunks = kSyevMaxN / 64;

__global__ __launch_bounds__(kStepThreads) void sytrd_step_kernel(
    int n, double* __restrict__ W, int ldw, int j, double* __restrict__ vbuf,
    double* __restrict__ pbuf, double* __restrict__ sbuf, double* __restrict__ VhT,
    double* __restrict__ d, double* __restrict__ e, double* __restrict__ tau) {
    extern __shared__ double sm[];
    const int G = gridDim.x, g = blockIdx.x;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    constexpr int NW = kStepThreads / 64;
    const int m0 = n - j;        // rows j..n-1: column j-1's vectors, index i <-> row j+i
    const int m = n - j - 1;     // column j's reflector, index i <-> row j+1+i
    double* wprev = sm;          // m0
    double* vprev = wprev + n;   // m0
    double* row = vprev + n;     // m0: row j of the updated matrix, columns j..
    double* v = row + n;         // m
    double* red = v + n;         // 2 * NW
    const int cur = j & 1, prv = cur ^ 1;
    // 1. column j-1's w
    double kk = 0.0, tprev = 0.0;
    if (j > 0) {
        tprev = tau[j - 1];
        double sp = 0.0;
        for (int q = t; q < G; q += kStepThreads) sp += sbuf[(size_t)prv * G + q];
        sp = wave_sum(sp);
        if (lane == 0) red[wv] = sp;
        __syncthreads();
        for (int q = 0; q < NW; ++q) kk += red[q];
        kk *= 0.5 * tprev;
        const double* vp = vbuf + (size_t)prv * n;
        const double* pp = pbuf + (size_t)prv * n;
        for (int i = t; i < m0; i += kStepThreads) {
            const double vi = vp[i];
            vprev[i] = vi;
            wprev[i] = pp[i] - kk * vi;
        }
    }
    __syncthreads();
    // 2. row j (columns j..n-1) after column j-1's update, and its reflector
    for (int c = t; c < m0; c += kStepThreads) {
        double a = W[(size_t)j * ldw + j + c];
        if (j > 0) a -= vprev[0] * wprev[c] + wprev[0] * vprev[c];
        row[c] = a;
    }
    __syncthreads();
    double s2 = 0.0;
    for (int i = 2 + t; i <= m; i += kStepThreads) s2 += row[i] * row[i];
    s2 = wave_sum(s2);
    if (lane == 0) red[NW + wv] = s2;
    __syncthreads();
    s2 = 0.0;
    for (int q = 0; q < NW; ++q) s2 += red[NW + q];
    const double alpha = m >= 1 ? row[1] : 0.0;
    double beta, tj, scal;
    if (s2 == 0.0) {
        tj = 0.0;
        beta = alpha;
        scal = 0.0;
    } else {
        beta = -copysign(sqrt(alpha * alpha + s2), alpha);
        tj = (beta - alpha) / beta;
        scal = 1.0 / (alpha - beta);
    }
    for (int i = t; i < m; i += kStepThreads) v[i] = i == 0 ? 1.0 : row[1 + i] * scal;
    __syncthreads();
    if (g == 0) {
        for (int i = t; i < m; i += kStepThreads) {
            vbuf[(size_t)cur * n + i] = v[i];
            VhT[(size_t)j * n + j + 1 + i] = v[i];
        }
        if (t == 0) {
            d[j] = row[0];
            e[j] = beta;
            tau[j] = tj;
        }
    }
    // 3. own rows r > j: update with column j-1's vectors, then p_r
    double dot = 0.0;
    const int rfirst = j + 1 + ((g - (j + 1) % G) % G + G) % G;  // first r > j with r = g (mod G)
    // a row's columns are loaded in one batch (kStepChunks per lane), so a
    // row costs one memory latency instead of one per 64 columns
    for (int r = rfirst + wv * G; r < n; r += NW * G) {
        const int ir = r - j;        // index of row r in column j-1's vectors
        const double vr = j > 0 ? vprev[ir] : 0.0, wr = j > 0 ? wprev[ir] : 0.0;
        double* Wr = W + (size_t)r * ldw + (j + 1);
        double a[kStepChunks];
#pragma unroll
        for (int q = 0; q < kStepChunks; ++q) {
            const int c = lane + 64 * q;
            a[q] = c < m ? Wr[c] : 0.0;
        }
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < kStepChunks; ++q) {
            const int c = lane + 64 * q;
            if (c < m) {
                double x = a[q];
                if (j > 0) {
                    x -= vr * wprev[c + 1] + wr * vprev[c + 1];
                    Wr[c] = x;
                }
                acc += x * v[c];
            }
        }
        acc = wave_sum(acc);
        if (lane == 0) {
            const double pr = tj * acc;
            pbuf[(size_t)cur * n + (r - j - 1)] = pr;
            dot += v[r - j - 1] * pr;
        }
        if (j == n - 2 && r == n - 1 && lane == 0) d[n - 1] = Wr[0];
    }
    dot = wave_sum(dot);
    __syncthreads();
    if (lane == 0) red[wv] = dot;
    __syncthreads();
    if (t == 0) {
        double sacc = 0.0;
        for (int q = 0; q < NW; ++q) sacc += red[q];
        sbuf[(size_t)cur * G + g] = sacc;
    }
}

// The same column step with every input of the launch loaded in one batch
// at its start -- the workgroup's own rows (at most RPW per wavefront) into
// registers, row j, column j-1's p and v and the partial sums -- so a launch
// waits on memory once, not once per dependent stage.  For n <= 1024
// (grid of up to 128 workgroups, 16 column chunks per lane).
template <int RPW, int CH>
__global__ __launch_bounds__(kStepThreads) void sytrd_step_prefetch_kernel(
    int n, double* __restrict__ W, int ldw, int j, double* __restrict__ vbuf,
    double* __restrict__ pbuf, double* __restrict__ sbuf, double* __restrict__ VhT,
    double* __restrict__ d, double* __restrict__ e, double* __restrict__ tau) {
    extern __shared__ double sm[];
    const int G = gridDim.x, g = blockIdx.x;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    constexpr int NW = kStepThreads / 64;
    constexpr int PT = (64 * CH + kStepThreads - 1) / kStepThreads;  // row-j entries per thread
    const int m0 = n - j;
    const int m = n - j - 1;
    double* wprev = sm;
    double* vprev = wprev + n;
    double* row = vprev + n;
    double* v = row + n;
    double* red = v + n;
    const int cur = j & 1, prv = cur ^ 1;
    // ---- every load of the launch, issued together
    const int rfirst = j + 1 + ((g - (j + 1) % G) % G + G) % G;
    double a[RPW][CH];
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        const int r = rfirst + (wv + NW * k) * G;
        const double* Wr = W + (size_t)(r < n ? r : 0) * ldw + (j + 1);
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int c = lane + 64 * q;
            a[k][q] = (r < n && c < m) ? Wr[c] : 0.0;
        }
    }
    double rj[PT], pp[PT], vp[PT];
#pragma unroll
    for (int k = 0; k < PT; ++k) {
        const int c = t + kStepThreads * k;
        const bool ok = c < m0;
        rj[k] = ok ? W[(size_t)j * ldw + j + c] : 0.0;
        pp[k] = (ok && j > 0) ? pbuf[(size_t)prv * n + c] : 0.0;
        vp[k] = (ok && j > 0) ? vbuf[(size_t)prv * n + c] : 0.0;
    }
    const double sp = (j > 0 && t < G) ? sbuf[(size_t)prv * G + t] : 0.0;
    const double tprev = j > 0 ? tau[j - 1] : 0.0;
    // ---- column j-1's w
    double kk = wave_sum(sp);
    if (lane == 0) red[wv] = kk;
    __syncthreads();
    kk = 0.0;
    for (int q = 0; q < NW; ++q) kk += red[q];
    kk *= 0.5 * tprev;
#pragma unroll
    for (int k = 0; k < PT; ++k) {
        const int c = t + kStepThreads * k;
        if (c < m0) {
            vprev[c] = vp[k];
            wprev[c] = pp[k] - kk * vp[k];
        }
    }
    __syncthreads();
    // ---- row j after column j-1's update, and its reflector
    double s2 = 0.0;
#pragma unroll
    for (int k = 0; k < PT; ++k) {
        const int c = t + kStepThreads * k;
        if (c < m0) {
            double x = rj[k];
            if (j > 0) x -= vprev[0] * wprev[c] + wprev[0] * vprev[c];
            row[c] = x;
            if (c >= 2) s2 += x * x;
        }
    }
    s2 = wave_sum(s2);
    if (lane == 0) red[NW + wv] = s2;
    __syncthreads();
    s2 = 0.0;
    for (int q = 0; q < NW; ++q) s2 += red[NW + q];
    const double alpha = row[1];
    double beta, tj, scal;
    if (s2 == 0.0) {
        tj = 0.0;
        beta = alpha;
        scal = 0.0;
    } else {
        beta = -copysign(sqrt(alpha * alpha + s2), alpha);
        tj = (beta - alpha) / beta;
        scal = 1.0 / (alpha - beta);
    }
    for (int i = t; i < m; i += kStepThreads) v[i] = i == 0 ? 1.0 : row[1 + i] * scal;
    __syncthreads();
    if (g == 0) {
        for (int i = t; i < m; i += kStepThreads) {
            vbuf[(size_t)cur * n + i] = v[i];
            VhT[(size_t)j * n + j + 1 + i] = v[i];
        }
        if (t == 0) {
            d[j] = row[0];
            e[j] = beta;
            tau[j] = tj;
        }
    }
    // ---- own rows: column j-1's update, then p_r
    double dot = 0.0;
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
        const int r = rfirst + (wv + NW * k) * G;
        if (r < n) {
            const int ir = r - j;
            const double vr = j > 0 ? vprev[ir] : 0.0, wr = j > 0 ? wprev[ir] : 0.0;
            double* Wr = W + (size_t)r * ldw + (j + 1);
            double acc = 0.0;
#pragma unroll
            for (int q = 0; q < CH; ++q) {
                const int c = lane + 64 * q;
                if (c < m) {
                    double x = a[k][q];
                    if (j > 0) {
                        x -= vr * wprev[c + 1] + wr * vprev[c + 1];
                        Wr[c] = x;
                    }
                    acc += x * v[c];
                }
            }
            acc = wave_sum(acc);
            if (lane == 0) {
                const double pr = tj * acc;
                pbuf[(size_t)cur * n + (r - j - 1)] = pr;
                dot += v[r - j - 1] * pr;
                if (j == n - 2 && r == n - 1) d[n - 1] = Wr[0];
            }
        }
    }
    dot = wave_sum(dot);
    __syncthreads();
    if (lane == 0) red[wv] = dot;
    __syncthreads();
    if (t == 0) {
        double sacc = 0.0;
        for (int q = 0; q < NW; ++q) sacc += red[q];
        sbuf[(size_t)cur * G + g] = sacc;
    }
}

// W <- A (full symmetric storage), row stride ldw.
__global__ __launch_bounds__(256) void copy_full_kernel(int n, const double* __restrict__ A, int lda,
                                                        double* __restrict__ W, int ldw) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * n) return;
    const int r = (int)(e / n), c = (int)(e - (size_t)r * n);
    W[(size_t)r * ldw + c] = A[(size_t)r * lda + c];
}

__global__ void diag_one_kernel(const double* __restrict__ W, double* __restrict__ d) {
    if (threadIdx.x == 0) d[0] = W[0];
}

__global__ void fill_kernel(size_t cnt, double* __restrict__ p, double val) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) p[i] = val;
}

// Number of eigenvalues of T below x (Sturm sequence of the LDL^T pivots,
// with LAPACK's pivmin safeguard against zero pivots).
__device__ __forceinline__ int sturm_count(int n, const double* dd, const double* e2, double x,
                                           double pivmin) {
    double q = dd[0] - x;
    if (fabs(q) <= pivmin) q = -pivmin;
    int c = q < 0.0;
    for (int i = 1; i < n; ++i) {
        q = dd[i] - x - e2[i - 1] / q;
        if (fabs(q) <= pivmin) q = -pivmin;
        c += q < 0.0;
    }
    return c;
}

// Eigenvalue i of T (ascending) per wavefront: 64-point multisection of the
// Gershgorin interval.  bounds = (lower, upper, pivmin) from tri_bounds_kernel.
__global__ __launch_bounds__(64) void stebz_kernel(int n, const double* __restrict__ d,
                                                   const double* __restrict__ e,
                                                   const double* __restrict__ bounds,
                                                   double* __restrict__ w) {
    extern __shared__ double sm[];
    double* dd = sm;
    double* e2 = sm + n;
    const int lane = threadIdx.x;
    const int i = blockIdx.x;
    for (int q = lane; q < n; q += 64) {
        dd[q] = d[q];
        if (q < n - 1) e2[q] = e[q] * e[q];
    }
    __syncthreads();
    const double pivmin = bounds[2];
    double lo = bounds[0], hi = bounds[1];
    for (int round = 0; round < 24; ++round) {
        const double tol = 2.0 * 2.2204460492503131e-16 * fmax(fabs(lo), fabs(hi)) + 2.0 * pivmin;
        if (hi - lo <= tol) break;
        const double x = lo + (hi - lo) * (double)(lane + 1) / 65.0;
        const int c = sturm_count(n, dd, e2, x, pivmin);
        // lanes with count(x) <= i lie below eigenvalue i; counts grow with
        // the lane, and only the leading run of such lanes is trusted
        const unsigned long long below = __ballot(c <= i);
        const int nb = below == ~0ull ? 64 : __builtin_ctzll(~below);
        const double nlo = nb > 0 ? lo + (hi - lo) * (double)nb / 65.0 : lo;
        const double nhi = nb < 64 ? lo + (hi - lo) * (double)(nb + 1) / 65.0 : hi;
        lo = nlo;
        hi = nhi;
    }
    if (lane == 0) w[i] = 0.5 * (lo + hi);
}

__global__ __launch_bounds__(256) void tri_bounds_kernel(int n, const double* __restrict__ d,
                                                         const double* __restrict__ e,
                                                         double* __restrict__ bounds) {
    __shared__ double slo[256], shi[256], se[256];
    double lo = 1e300, hi = -1e300, em = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        const double a = i > 0 ? fabs(e[i - 1]) : 0.0;
        const double b = i < n - 1 ? fabs(e[i]) : 0.0;
        lo = fmin(lo, d[i] - a - b);
        hi = fmax(hi, d[i] + a + b);
        if (i < n - 1) em = fmax(em, e[i] * e[i]);
    }
    slo[threadIdx.x] = lo;
    shi[threadIdx.x] = hi;
    se[threadIdx.x] = em;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            slo[threadIdx.x] = fmin(slo[threadIdx.x], slo[threadIdx.x + s]);
            shi[threadIdx.x] = fmax(shi[threadIdx.x], shi[threadIdx.x + s]);
            se[threadIdx.x] = fmax(se[threadIdx.x], se[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double tn = fmax(fabs(slo[0]), fabs(shi[0]));
        const double widen = 2.2204460492503131e-16 * tn * (2.0 * n + 2.0) + 1e-300;
        bounds[0] = slo[0] - widen;
        bounds[1] = shi[0] + widen;
        bounds[2] = 2.2250738585072014e-308 * fmax(1.0, se[0]) / 2.2204460492503131e-16;
        bounds[3] = tn;
    }
}

// Eigenvectors of T by inverse iteration (LAPACK stein): one wavefront per
// cluster leader.  A cluster is a run of eigenvalues with consecutive gaps
// below 1e-7 |T|; its members are computed in order, each iterate
// orthogonalised (modified Gram-Schmidt) against the members before it.
// Eigenvector i is written to row i of Zt (n x n).  The tridiagonal
// factorisation with partial pivoting (dlagtf) and the solve (dlagts) run on
// lane 0; the Gram-Schmidt and normalisation use the whole wavefront.
__global__ __launch_bounds__(64) void stein_kernel(int n, const double* __restrict__ d,
                                                   const double* __restrict__ e,
                                                   const double* __restrict__ w,
                                                   const double* __restrict__ bounds,
                                                   double* __restrict__ Zt) {
    extern __shared__ double sm[];
    double* a = sm;          // U diagonal
    double* b = a + n;       // U first super-diagonal
    double* c = b + n;       // multipliers
    double* dd = c + n;      // U second super-diagonal
    double* z = dd + n;      // iterate
    int* piv = reinterpret_cast<int*>(z + n);
    const int lane = threadIdx.x;
    const int i0 = blockIdx.x;
    const double tn = bounds[3];
    const double ctol = 1e-7 * tn;
    if (i0 > 0 && w[i0] - w[i0 - 1] <= ctol) return;  // not a cluster leader
    int i1 = i0 + 1;
    while (i1 < n && w[i1] - w[i1 - 1] <= ctol) ++i1;
    const double eps = 2.2204460492503131e-16;
    // pivot floor: eps |T|, and a normal number even for T = 0
    const double delta = eps * fmax(tn, 1e-200);
    double lam_prev = -1e300;
    for (int i = i0; i < i1; ++i) {
        // equal eigenvalues of a cluster are separated so that their shifted
        // systems differ (LAPACK stein)
        double lam = w[i];
        if (i > i0 && lam - lam_prev < 10.0 * eps * fabs(lam) + delta) lam = lam_prev + 10.0 * eps * fabs(lam) + delta;
        lam_prev = lam;
        if (lane == 0) {
            for (int k = 0; k < n; ++k) {
                a[k] = d[k] - lam;
                b[k] = k < n - 1 ? e[k] : 0.0;
                c[k] = k < n - 1 ? e[k] : 0.0;
                dd[k] = 0.0;
                piv[k] = 0;
            }
            for (int k = 0; k < n - 1; ++k) {
                if (fabs(a[k]) >= fabs(c[k])) {
                    const double ak = fabs(a[k]) < delta ? copysign(delta, a[k] == 0.0 ? 1.0 : a[k]) : a[k];
                    a[k] = ak;
                    const double mult = c[k] / ak;
                    c[k] = mult;
                    a[k + 1] -= mult * b[k];
                    dd[k] = 0.0;
                    piv[k] = 0;
                } else {
                    const double mult = a[k] / c[k];
                    a[k] = c[k];
                    const double temp = a[k + 1];
                    a[k + 1] = b[k] - mult * temp;
                    if (k < n - 2) {
                        dd[k] = b[k + 1];
                        b[k + 1] = -mult * dd[k];
                    }
                    b[k] = temp;
                    c[k] = mult;
                    piv[k] = 1;
                }
            }
            if (fabs(a[n - 1]) < delta) a[n - 1] = copysign(delta, a[n - 1] == 0.0 ? 1.0 : a[n - 1]);
        }
        // deterministic, non-degenerate starting vector
        for (int k = lane; k < n; k += 64) {
            const unsigned h = (unsigned)(k * 2654435761u) ^ (unsigned)(i * 40503u + 17u);
            z[k] = 0.5 + (double)(h & 0xffff) / 65536.0;
        }
        __syncthreads();
        for (int it = 0; it < 3; ++it) {
            if (lane == 0) {
                // forward: apply the row interchanges and multipliers
                for (int k = 0; k < n - 1; ++k) {
                    if (piv[k] == 0) {
                        z[k + 1] -= c[k] * z[k];
                    } else {
                        const double tmp = z[k];
                        z[k] = z[k + 1];
                        z[k + 1] = tmp - c[k] * z[k];
                    }
                }
                // back substitution with U (diag a, supers b, dd)
                z[n - 1] /= a[n - 1];
                if (n > 1) z[n - 2] = (z[n - 2] - b[n - 2] * z[n - 1]) / a[n - 2];
                for (int k = n - 3; k >= 0; --k)
                    z[k] = (z[k] - b[k] * z[k + 1] - dd[k] * z[k + 2]) / a[k];
            }
            __syncthreads();
            // orthogonalise against the cluster's earlier members, normalise
            for (int q = i0; q < i; ++q) {
                const double* zq = Zt + (size_t)q * n;
                double s = 0.0;
                for (int k = lane; k < n; k += 64) s += zq[k] * z[k];
                s = wave_sum(s);
                for (int k = lane; k < n; k += 64) z[k] -= s * zq[k];
                __syncthreads();
            }
            // normalise: scale by the largest component first (a solve
            // against a nearly singular system can exceed sqrt(DBL_MAX))
            double amax = 0.0;
            for (int k = lane; k < n; k += 64) amax = fmax(amax, fabs(z[k]));
            amax = wave_max(amax);
            const double sc = amax > 0.0 ? 1.0 / amax : 1.0;
            double nrm = 0.0;
            for (int k = lane; k < n; k += 64) {
                const double zk = z[k] * sc;
                nrm += zk * zk;
            }
            nrm = wave_sum(nrm);
            const double inv = nrm > 0.0 ? sc / sqrt(nrm) : 1.0;
            __syncthreads();
            for (int k = lane; k < n; k += 64) z[k] *= inv;
            __syncthreads();
        }
        // a sign convention: the largest component positive
        double big = 0.0;
        int arg = 0;
        for (int k = lane; k < n; k += 64)
            if (fabs(z[k]) > big) { big = fabs(z[k]); arg = k; }
        for (int o = 32; o > 0; o >>= 1) {
            const double ob = __shfl_xor(big, o, 64);
            const int oa = __shfl_xor(arg, o, 64);
            if (ob > big || (ob == big && oa < arg)) { big = ob; arg = oa; }
        }
        const double sg = z[arg] < 0.0 ? -1.0 : 1.0;
        for (int k = lane; k < n; k += 64) Zt[(size_t)i * n + k] = sg * z[k];
        __syncthreads();
    }
}

// Triangular factor of each block of NB reflectors (LAPACK larft, forward,
// columnwise): I - Vb Tb Vb^T = H_j0 ... H_{j0+nb-1}, Vb = VhT[j0:j0+nb]^T.
// One workgroup per block; the Gram G = Vb^T Vb one wavefront per entry.
__global__ __launch_bounds__(256) void larft_kernel(int n, int nref, const double* __restrict__ VhT,
                                                    const double* __restrict__ tau,
                                                    double* __restrict__ Tb) {
    constexpr int NB = kSyevNB;
    __shared__ double G[NB][NB + 1], T[NB][NB + 1];
    const int blk = blockIdx.x;
    const int j0 = blk * NB;
    const int nb = min(NB, nref - j0);
    __shared__ double Vs[NB][129];
    for (int e = threadIdx.x; e < NB * (NB + 1); e += 256) {
        (&G[0][0])[e] = 0.0;
        (&T[0][0])[e] = 0.0;
    }
    // G = Vb^T Vb over row chunks of 128 staged in LDS; thread -> entries
    // (p, q) with p <= q, at most 3 per thread (528 for NB = 32)
    double gacc[3] = {0.0, 0.0, 0.0};
    for (int r0 = j0 + 1; r0 < n; r0 += 128) {
        __syncthreads();
        for (int e = threadIdx.x; e < NB * 128; e += 256) {
            const int p = e >> 7, c = e & 127;
            const int r = r0 + c;
            Vs[p][c] = (p < nb && r < n) ? VhT[(size_t)(j0 + p) * n + r] : 0.0;
        }
        __syncthreads();
        for (int k = 0; k < 3; ++k) {
            const int idx = threadIdx.x + 256 * k;
            if (idx >= NB * (NB + 1) / 2) break;
            int p = 0, rem = idx;
            while (rem >= NB - p) { rem -= NB - p; ++p; }
            const int q = p + rem;
            double sacc = 0.0;
            for (int c = 0; c < 128; ++c) sacc += Vs[p][c] * Vs[q][c];
            gacc[k] += sacc;
        }
    }
    for (int k = 0; k < 3; ++k) {
        const int idx = threadIdx.x + 256 * k;
        if (idx >= NB * (NB + 1) / 2) break;
        int p = 0, rem = idx;
        while (rem >= NB - p) { rem -= NB - p; ++p; }
        G[p][p + rem] = gacc[k];
    }
    __syncthreads();
    for (int q = 0; q < nb; ++q) {
        // T[0:q, q] = -tau_q T[0:q, 0:q] G[0:q, q]
        const double tq = tau[j0 + q];
        if ((int)threadIdx.x < q) {
            const int p = threadIdx.x;
            double sacc = 0.0;
            for (int l = p; l < q; ++l) sacc += T[p][l] * G[l][q];
            T[p][q] = -tq * sacc;
        }
        if (threadIdx.x == 0) T[q][q] = tq;
        __syncthreads();
    }
    for (int e = threadIdx.x; e < NB * NB; e += 256) {
        const int p = e / NB, q = e - p * NB;
        Tb[(size_t)blk * NB * NB + e] = (p < nb && q < nb) ? T[p][q] : 0.0;
    }
}

// V = H_0 H_1 ... H_{n-2} Z, Z = Zt^T (Zt row i = eigenvector i of T).
// Columns are independent under a left product, so workgroup c0 takes NC
// columns of Z into LDS and applies the reflector blocks last to first in
// compact WY form: Y1 = Vb^T Z, Y2 = Tb Y1, Z -= Vb Y2, streaming each
// block's reflectors (rows j0.. of VhT) through LDS in chunks of 128 rows.
template <int NC>
__global__ __launch_bounds__(256) void apply_q_kernel(int n, int nref, const double* __restrict__ VhT,
                                                      const double* __restrict__ Tb,
                                                      const double* __restrict__ Zt,
                                                      double* __restrict__ V, int ldv) {
    constexpr int NB = kSyevNB, RC = 128;
    extern __shared__ double sm[];
    double* Z = sm;                                // n x NC
    double* Vs = Z + (size_t)n * NC;               // NB x (RC + 1)
    double* Y = Vs + NB * (RC + 1);                // NB x NC
    double* Ts = Y + NB * NC;                      // NB x (NB + 1)
    const int t = threadIdx.x;
    const int c0 = blockIdx.x * NC;
    const int ncol = min(NC, n - c0);
    for (int e = t; e < n * NC; e += 256) {
        const int c = e / n, r = e - c * n;         // coalesced over r
        Z[r * NC + c] = c < ncol ? Zt[(size_t)(c0 + c) * n + r] : 0.0;
    }
    const int nblk = (nref + NB - 1) / NB;
    for (int b = nblk - 1; b >= 0; --b) {
        const int j0 = b * NB;
        const int nb = min(NB, nref - j0);
        __syncthreads();
        for (int e = t; e < NB * NB; e += 256) Ts[(e / NB) * (NB + 1) + e % NB] = Tb[(size_t)b * NB * NB + e];
        // Y1 = Vb^T Z: thread -> (p, c) entries
        double y1[(NB * NC + 255) / 256];
#pragma unroll
        for (int k = 0; k < (NB * NC + 255) / 256; ++k) y1[k] = 0.0;
        for (int r0 = j0 + 1; r0 < n; r0 += RC) {
            __syncthreads();
            for (int e = t; e < NB * RC; e += 256) {
                const int p = e / RC, cc = e - p * RC;
                const int r = r0 + cc;
                Vs[p * (RC + 1) + cc] = (p < nb && r < n) ? VhT[(size_t)(j0 + p) * n + r] : 0.0;
            }
            __syncthreads();
            const int rl = min(RC, n - r0);
#pragma unroll
            for (int k = 0; k < (NB * NC + 255) / 256; ++k) {
                const int idx = t + 256 * k;
                if (idx < NB * NC) {
                    const int p = idx / NC, c = idx - p * NC;
                    double sacc = 0.0;
                    for (int cc = 0; cc < rl; ++cc) sacc += Vs[p * (RC + 1) + cc] * Z[(r0 + cc) * NC + c];
                    y1[k] += sacc;
                }
            }
        }
        __syncthreads();
        // Y2 = Tb Y1 (Tb upper triangular)
#pragma unroll
        for (int k = 0; k < (NB * NC + 255) / 256; ++k) {
            const int idx = t + 256 * k;
            if (idx < NB * NC) Y[idx] = y1[k];
        }
        __syncthreads();
        double y2[(NB * NC + 255) / 256];
#pragma unroll
        for (int k = 0; k < (NB * NC + 255) / 256; ++k) {
            const int idx = t + 256 * k;
            y2[k] = 0.0;
            if (idx < NB * NC) {
                const int p = idx / NC, c = idx - p * NC;
                for (int l = p; l < NB; ++l) y2[k] += Ts[p * (NB + 1) + l] * Y[l * NC + c];
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < (NB * NC + 255) / 256; ++k) {
            const int idx = t + 256 * k;
            if (idx < NB * NC) Y[idx] = y2[k];
        }
        // Z -= Vb Y2, chunk by chunk
        for (int r0 = j0 + 1; r0 < n; r0 += RC) {
            __syncthreads();
            for (int e = t; e < NB * RC; e += 256) {
                const int p = e / RC, cc = e - p * RC;
                const int r = r0 + cc;
                Vs[p * (RC + 1) + cc] = (p < nb && r < n) ? VhT[(size_t)(j0 + p) * n + r] : 0.0;
            }
            __syncthreads();
            const int rl = min(RC, n - r0);
            for (int e = t; e < rl * NC; e += 256) {
                const int cc = e / NC, c = e - cc * NC;
                double sacc = 0.0;
#pragma unroll 8
                for (int p = 0; p < NB; ++p) sacc += Vs[p * (RC + 1) + cc] * Y[p * NC + c];
                Z[(r0 + cc) * NC + c] -= sacc;
            }
        }
    }
    __syncthreads();
    for (int e = t; e < n * NC; e += 256) {
        const int r = e / NC, c = e - r * NC;
        if (c < ncol) V[(size_t)r * ldv + c0 + c] = Z[e];
    }
}

struct SyevWs {
    size_t d, e, tau, bounds, W, Vh, vb, pb, sb, Zt, Tb, total;
    int ldw;
};

static SyevWs syev_layout(int n) {
    SyevWs L{};
    size_t off = 0;
    auto take = [&](size_t doubles) {
        const size_t o = off;
        off += align_up(doubles * sizeof(double), 256);
        return o;
    };
    const int nblk = (n + kSyevNB - 1) / kSyevNB;
    L.d = take(n);
    L.e = take(n);
    L.tau = take(n);
    L.bounds = take(4);
    L.ldw = n + (n & 1);
    L.W = take((size_t)n * L.ldw + 2);  // + the pair a 16-byte load may touch past the last row
    L.Vh = take((size_t)n * n);
    L.vb = take(2 * (size_t)n);
    L.pb = take(2 * (size_t)n);
    L.sb = take(2 * (size_t)kStepGroupsMax);
    L.Zt = take((size_t)n * n);
    L.Tb = take((size_t)nblk * kSyevNB * kSyevNB);
    L.total = off;
    return L;
}

}  // namespace rvgp

using namespace rvgp;

extern "C" size_t rvgp_syev_ws_bytes(int n) {
    if (n < 1 || n > kSyevMaxN) return 0;
    return syev_layout(n).total;
}

extern "C" int rvgp_syev(int n, double* A, int lda, double* w, double* V, int ldv, void* ws,
                         size_t ws_bytes, void* stream) {
    RVGP_CHECK_ARG(n >= 1 && n <= kSyevMaxN, "rvgp_syev: need 1 <= n <= %d", kSyevMaxN);
    RVGP_CHECK_ARG(lda >= n && ldv >= n, "rvgp_syev: leading dimension < n");
    const SyevWs L = syev_layout(n);
    RVGP_CHECK_ARG(ws != nullptr && ws_bytes >= L.total, "rvgp_syev: needs %zu bytes of workspace",
                   L.total);
    hipStream_t s = as_stream(stream);
    char* base = static_cast<char*>(ws);
    double* d = reinterpret_cast<double*>(base + L.d);
    double* e = reinterpret_cast<double*>(base + L.e);
    double* tau = reinterpret_cast<double*>(base + L.tau);
    double* bounds = reinterpret_cast<double*>(base + L.bounds);
    double* W = reinterpret_cast<double*>(base + L.W);
    double* VhT = reinterpret_cast<double*>(base + L.Vh);
    double* Vp = reinterpret_cast<double*>(base + L.vb);
    double* Wp = reinterpret_cast<double*>(base + L.pb);
    double* sb = reinterpret_cast<double*>(base + L.sb);
    double* Zt = reinterpret_cast<double*>(base + L.Zt);
    double* Tb = reinterpret_cast<double*>(base + L.Tb);
    const int ldw = L.ldw;

    // stage 1: tridiagonalisation, one launch per column (reflectors
    // j = 0 .. n-2; the last has length 1 and tau = 0)
    const size_t nn = (size_t)n * n;
    hipLaunchKernelGGL(copy_full_kernel, dim3(cdiv(nn, 256)), dim3(256), 0, s, n, A, lda, W, ldw);
    RVGP_LAUNCH_CHECK();
    hipLaunchKernelGGL(fill_kernel, dim3(cdiv(nn, 256)), dim3(256), 0, s, nn, VhT, 0.0);
    RVGP_LAUNCH_CHECK();
    if (n == 1) {
        hipLaunchKernelGGL(diag_one_kernel, dim3(1), dim3(64), 0, s, W, d);
        RVGP_LAUNCH_CHECK();
    }
    const size_t stsm = sizeof(double) * (4 * (size_t)n + 2 * (kStepThreads / 64));
    if (n <= 1024) {
        // 128 workgroups x 4 wavefronts x 2 rows cover the <= 1023 trailing rows
        const int G = n < kStepGroupsMax ? n : kStepGroupsMax;
        auto kern = sytrd_step_prefetch_kernel<2, 16>;
        RVGP_CHECK_HIP(hipFuncSetAttribute((const void*)kern,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)stsm));
        for (int j = 0; j + 1 < n; ++j) {
            hipLaunchKernelGGL(kern, dim3(G), dim3(kStepThreads), stsm, s, n, W, ldw, j, Vp, Wp,
                               sb, VhT, d, e, tau);
            RVGP_LAUNCH_CHECK();
        }
    } else {
        const int G = kStepGroups;
        RVGP_CHECK_HIP(hipFuncSetAttribute((const void*)sytrd_step_kernel,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)stsm));
        for (int j = 0; j + 1 < n; ++j) {
            hipLaunchKernelGGL(sytrd_step_kernel, dim3(G), dim3(kStepThreads), stsm, s, n, W, ldw,
                               j, Vp, Wp, sb, VhT, d, e, tau);
            RVGP_LAUNCH_CHECK();
        }
    }
    const int nref = n - 1;
    // stage 2: eigenvalues and eigenvectors of T
    hipLaunchKernelGGL(tri_bounds_kernel, dim3(1), dim3(256), 0, s, n, d, e, bounds);
    RVGP_LAUNCH_CHECK();
    hipLaunchKernelGGL(stebz_kernel, dim3(n), dim3(64), sizeof(double) * 2 * n, s, n, d, e, bounds,
                       w);
    RVGP_LAUNCH_CHECK();
    const size_t ssm = sizeof(double) * 5 * n + sizeof(int) * n;
    RVGP_CHECK_HIP(hipFuncSetAttribute((const void*)stein_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)ssm));
    hipLaunchKernelGGL(stein_kernel, dim3(n), dim3(64), ssm, s, n, d, e, w, bounds, Zt);
    RVGP_LAUNCH_CHECK();
    // stage 3: V = H_0 H_1 ... H_{n-2} Z with Z = Zt^T (one fused kernel,
    // a few columns of Z per workgroup)
    const int nblk = (nref + kSyevNB - 1) / kSyevNB;
    if (nblk > 0) {
        hipLaunchKernelGGL(larft_kernel, dim3(nblk), dim3(256), 0, s, n, nref, VhT, tau, Tb);
        RVGP_LAUNCH_CHECK();
    }
    const int nc = n <= 1024 ? 8 : 4;
    const size_t qsm = sizeof(double) * ((size_t)n * nc + kSyevNB * 129 + kSyevNB * nc +
                                         kSyevNB * (kSyevNB + 1));
    if (nc == 8) {
        RVGP_CHECK_HIP(hipFuncSetAttribute((const void*)apply_q_kernel<8>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)qsm));
        hipLaunchKernelGGL(apply_q_kernel<8>, dim3(cdiv(n, 8)), dim3(256), qsm, s, n, nref, VhT, Tb,
                           Zt, V, ldv);
    } else {
        RVGP_CHECK_HIP(hipFuncSetAttribute((const void*)apply_q_kernel<4>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)qsm));
        hipLaunchKernelGGL(apply_q_kernel<4>, dim3(cdiv(n, 4)), dim3(256), qsm, s, n, nref, VhT, Tb,
                           Zt, V, ldv);
    }
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}
