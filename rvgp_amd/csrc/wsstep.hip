// a14/a15  One optimiser step of a single weight-space fit in ONE launch:
// spectral filter (kernels.py:41-53), A = G + s2 diag(S)^-1, the Cholesky
// factor A = L L^T and X = L^-1, diag(A^-1) and w = A^-1 b, the NLML and its
// adjoint (the weight-space formulas of gp.hip) and -- when asked -- the Adam
// update of the four unconstrained hyperparameters.  It replaces the ~40
// launches per step of the multi-kernel path (potrf, blocked trsm, GEMVs,
// reductions) behind GPflow's optimiser loop (RVGP/main.py:87-95).
//
// The k x k system is cut into 32 x 32 FP64 tiles (T = ceil(k/32) tile rows,
// the padding is the identity).  Each workgroup owns ONE task of a tile
// dataflow and waits for the tiles it reads through per-task flags:
//   L(j,j)  C = A(j,j) - sum_{p<j} L(j,p) L(j,p)^T; L = chol(C), D(j) =
//           L(j,j)^-1, and the tile below, L(j+1,j) = C(j+1,j) D(j)^T, so
//           that the next diagonal task is one hand-off away (one wave
//           factors; the other three apply C(j+1,j)'s last update meanwhile)
//   L(i,j)  i > j+1: C = A(i,j) - sum_{p<j} L(i,p) L(j,p)^T; L = C D(j)^T
//   X(i,j)  i > j: X = -D(i) sum_{p=j}^{i-1} L(i,p) X(p,j),  X(p,p) = D(p)
//   filter  S, dS and sum log S (no dependencies)
//   row i   t_i = sum_{j<=i} X(i,j) b_j                      (t = X b)
//   col j   w_j = sum_{i>=j} X(i,j)^T t_i,  diag(A^-1)_j = column sums of
//           squares of X(i,j), i >= j                        (w = X^T X b),
//           and tile j's share of every sum the NLML and adjoint need
//   final   NLML, adjoint, Adam; advances the epoch.
// Tasks are numbered (L tasks column by column, then X tasks column by
// column, filter, rows, columns, final) so that every task depends only on tasks
// with smaller block indices: the dispatcher starts workgroups in index
// order, so a waiting workgroup never holds a slot its producer needs, and
// the grid drains for any number of resident workgroups.
//
// Synchronisation (HIP/HSA memory model, agent scope): a producer's every
// wave waits until its stores have reached the L2, the workgroup meets at a
// barrier, then one lane issues an agent-scope release fence (the L2 is
// written back) and stores the flag; a consumer's lane 0 polls the flag with relaxed agent-scope
// loads, issues an agent-scope acquire fence (L1/L2 invalidation) and the
// workgroup meets at a barrier before reading.  Flags hold the step's epoch
// (sync[ctl]), so nothing is cleared between steps.  Every wait is bounded:
// past kSpinLimit polls -- or once another workgroup has given up -- it sets
// an error flag and continues, the step completes with garbage and status
// RVGP_ST_TIMEOUT; no launch can hang.
//
// Summation orders are fixed, so results are identical from run to run.
#include "common.h"

#pragma clang fp contract(fast)

namespace rvgp {
namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

constexpr int TB = 32;          // tile edge
constexpr int TT = TB * TB;     // doubles per tile
constexpr int NTH = 256;        // threads per task workgroup (4 waves)
constexpr int LDT = TB + 1;     // LDS tile row stride
constexpr int kSpinLimit = 1 << 21;

// task numbering.  L column j: the diagonal task (which also produces
// L(j+1,j)), then L(i,j) for i >= j+2; X column j: X(i,j) for i > j.
__host__ __device__ inline int lbase(int T, int j) { return j * (T - 1) - j * (j - 1) / 2; }
__host__ __device__ inline int xbase(int T, int j) { return j * (T - 1) - j * (j - 1) / 2; }
__host__ __device__ inline int lid(int T, int i, int j) { return lbase(T, j) + (i - j - 1); }  // i >= j+2
__host__ __device__ inline int xid(int T, int i, int j) { return xbase(T, j) + (i - j - 1); }
__host__ __device__ inline int tix(int i, int j) { return i * (i + 1) / 2 + j; }  // packed tile

struct Layout {
    int T, nL, nX;
    int fL, fD, fX, fF, fR, fC, ctl, nsync;
    int nblocks;
    size_t oL, oX, oT, oW, oA, oS, oDS, oP, oSc, nws;  // workspace offsets (doubles)
};

__host__ __device__ inline Layout make_layout(int k) {
    Layout g;
    g.T = (k + TB - 1) / TB;
    const int T = g.T;
    g.nL = lbase(T, T - 1) + 1;  // tasks; T(T+1)/2 tiles are stored
    g.nX = T * (T - 1) / 2;
    g.fL = 0;
    g.fD = g.nL;
    g.fX = g.fD + T;
    g.fF = g.fX + g.nX;
    g.fR = g.fF + 1;
    g.fC = g.fR + T;
    g.ctl = g.fC + T;
    g.nsync = g.ctl + 4;
    g.nblocks = g.nL + g.nX + 2 * T + 2;
    const size_t kp = (size_t)T * TB;
    const size_t ntiles = (size_t)T * (T + 1) / 2;
    g.oL = 0;
    g.oX = g.oL + ntiles * TT;
    g.oT = g.oX + ntiles * TT;
    g.oW = g.oT + kp;
    g.oA = g.oW + kp;
    g.oS = g.oA + kp;
    g.oDS = g.oS + kp;
    g.oP = g.oDS + 3 * kp;  // per-tile partial sums, 8 per tile row
    g.oSc = g.oP + 8 * (size_t)T;
    g.nws = g.oSc + 8;
    return g;
}

// flag that announces L(i,j), i > j: L(j+1,j) comes with the diagonal task's D(j)
__host__ __device__ inline int lflag(const Layout& g, int i, int j) {
    return i == j + 1 ? g.fD + j : g.fL + lid(g.T, i, j);
}

struct StepArgs {
    int k, typ, ldg, do_adam;
    double n_rows, M, lr, b1, b2, eps;
    const double* lam;
    const double* G;
    const double* b;
    const double* yy;
    const double* lower;
    double* u;
    const int* trainable;
    double* m1;
    double* m2;
    double* tstep;
    double* ws;
    int* sync;
    double* out;
    int* status;
};

__device__ __forceinline__ int ld_flag(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_flag(int* p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bounded wait until *flag reaches epoch; afterwards the producer's stores
// are visible to every thread of the workgroup.
__device__ void wait_for(const int* flag, int epoch, int* ctl) {
    if (threadIdx.x == 0) {
        int n = 0;
        while ((int)(ld_flag(flag) - epoch) < 0) {
            ++n;
            if (n > kSpinLimit || ((n & 1023) == 0 && ld_flag(ctl + 1) != 0)) {
                st_flag(ctl + 1, 1);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// wait_for for one wave (the others of its workgroup keep working): its lane
// 0 polls, then the whole wave acquires.
__device__ void wait_for_wave(const int* flag, int epoch, int* ctl) {
    if ((threadIdx.x & 63) == 0) {
        int n = 0;
        while ((int)(ld_flag(flag) - epoch) < 0) {
            ++n;
            if (n > kSpinLimit || ((n & 1023) == 0 && ld_flag(ctl + 1) != 0)) {
                st_flag(ctl + 1, 1);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// wait_for over a set of n flags that are published at about the same time:
// the lanes of wave 0 poll flag_of(0 .. n-1) at once, then one acquire --
// one polling round trip for the set instead of one per flag.
template <typename FlagOf>
__device__ void wait_all(FlagOf flag_of, int n, int epoch, int* ctl) {
    if (threadIdx.x < 64) {
        for (int l = threadIdx.x; l < n; l += 64) {
            const int* f = flag_of(l);
            int cnt = 0;
            while ((int)(ld_flag(f) - epoch) < 0) {
                ++cnt;
                if (cnt > kSpinLimit || ((cnt & 1023) == 0 && ld_flag(ctl + 1) != 0)) {
                    st_flag(ctl + 1, 1);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// Every wave waits for its own stores to reach the L2 (vmcnt 0); after the
// barrier one lane's agent-scope release fence writes the L2 back -- the
// workgroup's waves share one CU and so one L2 -- and raises the flag.
__device__ void publish(int* flag, int epoch) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        st_flag(flag, epoch);
    }
}

// acc += sgn * P Q^T (TRANS) or sgn * P Q: the 16 x 16 quadrant (qi, qj) of
// a 32 x 32 tile product on the matrix cores.  Quadrant element r of a lane
// is row qi*16 + lane/16 + 4r, column qj*16 + lane%16.
template <bool TRANS>
__device__ __forceinline__ void tile_mma(dbl4& acc, const double* P, int ldp, const double* Q,
                                         int ldq, double sgn, int qi, int qj, int lane) {
    const int m = qi * 16 + (lane & 15), n = qj * 16 + (lane & 15), kq = lane >> 4;
    double a[8], b[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        a[s] = sgn * P[m * ldp + 4 * s + kq];
        b[s] = TRANS ? Q[n * ldq + 4 * s + kq] : Q[(4 * s + kq) * ldq + n];
    }
#pragma unroll
    for (int s = 0; s < 8; ++s) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b[s], acc, 0, 0, 0);
}

__device__ __forceinline__ void quad_store(const dbl4& acc, double* C, int ldc, int qi, int qj,
                                           int lane) {
#pragma unroll
    for (int r = 0; r < 4; ++r)
        C[(qi * 16 + (lane >> 4) + 4 * r) * ldc + qj * 16 + (lane & 15)] = acc[r];
}

__device__ __forceinline__ dbl4 quad_load(const double* C, int ldc, int qi, int qj, int lane) {
    dbl4 v;
#pragma unroll
    for (int r = 0; r < 4; ++r)
        v[r] = C[(qi * 16 + (lane >> 4) + 4 * r) * ldc + qj * 16 + (lane & 15)];
    return v;
}

// 16 x 16 products on one wave: acc += sgn * P Q^T (TRANS) or sgn * P Q,
// P and Q in LDS; element r of a lane is row lane/16 + 4r, column lane%16
template <bool TRANS>
__device__ __forceinline__ dbl4 mma16(dbl4 acc, const double* P, int ldp, const double* Q, int ldq,
                                      double sgn, int lane) {
    const int m = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const double a = sgn * P[m * ldp + 4 * s + kq];
        const double b = TRANS ? Q[m * ldq + 4 * s + kq] : Q[(4 * s + kq) * ldq + m];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    return acc;
}

__device__ __forceinline__ void store16(const dbl4& acc, double* C, int ldc, int lane) {
#pragma unroll
    for (int r = 0; r < 4; ++r) C[((lane >> 4) + 4 * r) * ldc + (lane & 15)] = acc[r];
}

__device__ __forceinline__ dbl4 load16(const double* C, int ldc, int lane) {
    dbl4 v;
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = C[((lane >> 4) + 4 * r) * ldc + (lane & 15)];
    return v;
}

// LDS ordering between the lanes of one wave
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// 1/sqrt(a): the hardware estimate refined by two Newton steps for the
// inverse square root, y <- y (3 - a y^2) / 2 (each squares the relative
// error; two take v_rsq_f64's estimate to a few ulps).  No division and no
// IEEE square-root sequence on the factorisation's serial path.
__device__ __forceinline__ double rsqrt_nr(double a) {
    double y = __builtin_amdgcn_rsq(a);
    const double h = 0.5 * a;
    y = y * fma(-h * y, y, 1.5);
    y = y * fma(-h * y, y, 1.5);
    return y;
}

// Cholesky factor (lower, zeros above) and its inverse of the 16 x 16 SPD
// block C (LDS), one wave: lane r < 16 holds row r of the factor in x[],
// lane c < 16 column c of the inverse in y[]; entries of other lanes come
// through v_readlane (SGPRs).  Both loops are right-looking, so the updates
// of one step are independent of each other and only the pivots are serial.
__device__ __forceinline__ void chol_inv16(const double* C, int ldc, double* L, double* D, int ldo,
                                           int lane, int& bad) {
    double x[16], y[16], rinv[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) x[c] = C[(lane & 15) * ldc + c];
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        double piv = bcast(x[c], c);
        if (!(piv > 0.0)) {  // not positive definite (or NaN)
            bad = 1;
            piv = 1.0;
        }
        const double r = rsqrt_nr(piv);
        rinv[c] = r;  // 1 / L[c][c]
        const double lrc = lane > c ? x[c] * r : (lane == c ? piv * r : 0.0);
        x[c] = lrc;
#pragma unroll
        for (int l = c + 1; l < 16; ++l) {
            x[l] -= lrc * bcast(lrc, l);
            // the update happens now: sinking it to x[l]'s next use would keep
            // every broadcast of this column live in SGPRs until then
            asm volatile("" : "+v"(x[l]));
        }
    }
    // the broadcasts below re-read the factor: an opaque copy keeps the
    // compiler from holding the ones above live (SGPR spills) to reuse them
#pragma unroll
    for (int l = 0; l < 16; ++l) asm volatile("" : "+v"(x[l]));
    // inverse, column c = lane: L y = e_c by right-looking substitution
#pragma unroll
    for (int r = 0; r < 16; ++r) y[r] = lane == r ? 1.0 : 0.0;
#pragma unroll
    for (int l = 0; l < 16; ++l) {
        y[l] *= rinv[l];
#pragma unroll
        for (int r = l + 1; r < 16; ++r) {
            y[r] -= bcast(x[l], r) * y[l];
            asm volatile("" : "+v"(y[r]));
        }
    }
    if (lane < 16) {
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            L[lane * ldo + c] = c <= lane ? x[c] : 0.0;
            D[c * ldo + lane] = c >= lane ? y[c] : 0.0;
        }
    }
}

__device__ double block_sum(double v, double* sh) {
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < NTH / 64; ++i) s += sh[i];
    return s;
}

// filter terms r_j, dlog r_j / d(nu, kappa) (kernels.py:41-53)
__device__ __forceinline__ void filter_terms(int typ, double lam, double nu, double kappa,
                                             double& r, double& dn, double& dk) {
    if (typ == 0) {
        const double base = lam + 2.0 * nu / (kappa * kappa);
        const double lb = log(base);
        r = exp(-nu * lb);  // base^-nu
        dn = -lb - nu * (2.0 / (kappa * kappa)) / base;
        dk = 4.0 * nu * nu / (kappa * kappa * kappa * base);
    } else {
        r = exp(-0.5 * lam * kappa * kappa);
        dn = 0.0;
        dk = -lam * kappa;
    }
}

__device__ __forceinline__ double softplus_d(double x) {
    return x > 0 ? x + log1p(exp(-x)) : log1p(exp(x));
}

// ---------------------------------------------------------------- tasks
// A(i,j) quadrant: G padded with the identity, plus s2 / S on the diagonal
__device__ dbl4 a_quadrant(const StepArgs& a, int i, int j, int qi, int qj, int lane,
                           const double* dadd /* LDS, 32 entries, diagonal tiles only */) {
    dbl4 v;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int lr = qi * 16 + (lane >> 4) + 4 * r, lc = qj * 16 + (lane & 15);
        const int row = i * TB + lr, col = j * TB + lc;
        double x;
        if (row < a.k && col < a.k) x = a.G[(size_t)row * a.ldg + col];
        else x = row == col ? 1.0 : 0.0;
        if (dadd != nullptr && lr == lc) x += dadd[lr];
        v[r] = x;
    }
    return v;
}

__device__ void task_diag(const StepArgs& a, const Layout& g, int j, int epoch, double* lds) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int qi = wv >> 1, qj = wv & 1, T = g.T;
    int* sync = a.sync;
    int* ctl = sync + g.ctl;
    double* Lt = a.ws + g.oL;
    double* Xt = a.ws + g.oX;
    double* Cs = lds;                            // 32 x 33 (then L, D, scratch)
    double* dadd = lds + 4 * TB * LDT + 16 * 17;  // 32
    double* sh = dadd + TB;                       // 4
    // theta and the filter normalisation sum over all k eigenvalues
    const double nu = softplus_d(a.u[0]) + a.lower[0];
    const double kappa = softplus_d(a.u[1]) + a.lower[1];
    const double sf = softplus_d(a.u[2]) + a.lower[2];
    const double s2 = softplus_d(a.u[3]) + a.lower[3];
    double sr = 0.0;
    for (int q = tid; q < a.k; q += NTH) {
        double r, dn, dk;
        filter_terms(a.typ, a.lam[q], nu, kappa, r, dn, dk);
        sr += r;
    }
    sr = block_sum(sr, sh);
    if (tid < TB) {
        const int row = j * TB + tid;
        double add = 0.0;
        if (row < a.k) {
            double r, dn, dk;
            filter_terms(a.typ, a.lam[row], nu, kappa, r, dn, dk);
            add = s2 / (sf * a.n_rows * r / sr);
        }
        dadd[tid] = add;
    }
    __syncthreads();
    // C(j,j) and, below it, C(j+1,j): every update but C(j+1,j)'s last one
    // (from L(j+1,j-1), a task of the previous column) before the factor
    const bool has1 = j + 1 < T;
    dbl4 acc = a_quadrant(a, j, j, qi, qj, lane, dadd);
    dbl4 acc1 = has1 ? a_quadrant(a, j + 1, j, qi, qj, lane, nullptr) : dbl4{0, 0, 0, 0};
    for (int p = 0; p < j; ++p) {
        wait_for(sync + lflag(g, j, p), epoch, ctl);
        const double* Ljp = Lt + (size_t)tix(j, p) * TT;
        tile_mma<true>(acc, Ljp, TB, Ljp, TB, -1.0, qi, qj, lane);
        if (has1 && p + 1 < j) {
            wait_for(sync + lflag(g, j + 1, p), epoch, ctl);
            tile_mma<true>(acc1, Lt + (size_t)tix(j + 1, p) * TT, TB, Ljp, TB, -1.0, qi, qj, lane);
        }
    }
    quad_store(acc, Cs, LDT, qi, qj, lane);
    double* C1 = Cs + 3 * TB * LDT + 16 * 17;  // 32 x 33: C(j+1,j), past wave 0's scratch
    if (has1) quad_store(acc1, C1, LDT, qi, qj, lane);
    __syncthreads();
    // Blocked factor and inverse of the tile on wave 0 (2 x 2 blocks of 16):
    //   L11 = chol(C11), D11 = L11^-1;  L21 = C21 D11^T;
    //   L22 = chol(C22 - L21 L21^T), D22 = L22^-1;  D21 = -D22 L21 D11
    int bad = 0;
    double* Ls = Cs + TB * LDT;  // 32 x 33: L(j,j)
    double* Ds = Ls + TB * LDT;  // 32 x 33: D(j) = L(j,j)^-1
    double* Ws = Ds + TB * LDT;  // 16 x 17 scratch
    if (wv == 0) {
        // upper-right 16 x 16 blocks of L and D are zero; the rest is written
        for (int e = lane; e < 256; e += 64) {
            Ls[(e >> 4) * LDT + 16 + (e & 15)] = 0.0;
            Ds[(e >> 4) * LDT + 16 + (e & 15)] = 0.0;
        }
        wave_lds_sync();
        chol_inv16(Cs, LDT, Ls, Ds, LDT, lane, bad);
        wave_lds_sync();
        dbl4 t = mma16<true>(dbl4{0, 0, 0, 0}, Cs + 16 * LDT, LDT, Ds, LDT, 1.0, lane);  // C21 D11^T
        store16(t, Ls + 16 * LDT, LDT, lane);                                             // L21
        wave_lds_sync();
        dbl4 c22 = load16(Cs + 16 * LDT + 16, LDT, lane);
        c22 = mma16<true>(c22, Ls + 16 * LDT, LDT, Ls + 16 * LDT, LDT, -1.0, lane);
        store16(c22, Ws, 17, lane);
        wave_lds_sync();
        chol_inv16(Ws, 17, Ls + 16 * LDT + 16, Ds + 16 * LDT + 16, LDT, lane, bad);
        wave_lds_sync();
        t = mma16<false>(dbl4{0, 0, 0, 0}, Ls + 16 * LDT, LDT, Ds, LDT, 1.0, lane);  // L21 D11
        store16(t, Ws, 17, lane);
        wave_lds_sync();
        t = mma16<false>(dbl4{0, 0, 0, 0}, Ds + 16 * LDT + 16, LDT, Ws, 17, -1.0, lane);
        store16(t, Ds + 16 * LDT, LDT, lane);  // D21
        wave_lds_sync();
        bad = (int)wave_max((double)bad);
        if (bad && lane == 0) atomicOr(ctl + 2, 1);
    } else if (has1 && j > 0) {
        // meanwhile waves 1-3 apply C(j+1,j)'s last update, from L(j+1,j-1),
        // quadrant by quadrant in LDS (wave 1: (0,0) and (1,1); wave 2: (0,1);
        // wave 3: (1,0)) -- the accumulation the owner waves would do in
        // registers, so the result is the same
        wait_for_wave(sync + lflag(g, j + 1, j - 1), epoch, ctl);
        const double* P = Lt + (size_t)tix(j + 1, j - 1) * TT;
        const double* Q = Lt + (size_t)tix(j, j - 1) * TT;
        for (int t = 0; t < (wv == 1 ? 2 : 1); ++t) {
            const int pi = wv == 1 ? t : wv - 2, pj = wv == 1 ? t : 3 - wv;
            dbl4 c = quad_load(C1, LDT, pi, pj, lane);
            tile_mma<true>(c, P, TB, Q, TB, -1.0, pi, pj, lane);
            quad_store(c, C1, LDT, pi, pj, lane);
        }
    }
    __syncthreads();
    if (has1) {
        // L(j+1,j) = C(j+1,j) D(j)^T
        dbl4 l1 = dbl4{0, 0, 0, 0};
        tile_mma<true>(l1, C1, LDT, Ds, LDT, 1.0, qi, qj, lane);
        quad_store(l1, Lt + (size_t)tix(j + 1, j) * TT, TB, qi, qj, lane);
    }
    {
        double* Lg = Lt + (size_t)tix(j, j) * TT;
        double* Dg = Xt + (size_t)tix(j, j) * TT;
        for (int e = tid; e < TT; e += NTH) {
            Lg[e] = Ls[(e / TB) * LDT + e % TB];
            Dg[e] = Ds[(e / TB) * LDT + e % TB];
        }
    }
    // L(j,j), D(j) and L(j+1,j) go out together on the D flag
    publish(sync + g.fD + j, epoch);
}

__device__ void task_offdiag(const StepArgs& a, const Layout& g, int i, int j, int epoch,
                             double* lds) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int qi = wv >> 1, qj = wv & 1, T = g.T;
    int* sync = a.sync;
    int* ctl = sync + g.ctl;
    double* Lt = a.ws + g.oL;
    const double* Xt = a.ws + g.oX;
    double* Cs = lds;
    dbl4 acc = a_quadrant(a, i, j, qi, qj, lane, nullptr);
    for (int p = 0; p < j; ++p) {
        wait_for(sync + lflag(g, i, p), epoch, ctl);
        wait_for(sync + lflag(g, j, p), epoch, ctl);
        tile_mma<true>(acc, Lt + (size_t)tix(i, p) * TT, TB, Lt + (size_t)tix(j, p) * TT, TB, -1.0,
                       qi, qj, lane);
    }
    quad_store(acc, Cs, LDT, qi, qj, lane);
    wait_for(sync + g.fD + j, epoch, ctl);  // also orders the LDS stores above
    dbl4 l = dbl4{0, 0, 0, 0};
    tile_mma<true>(l, Cs, LDT, Xt + (size_t)tix(j, j) * TT, TB, 1.0, qi, qj, lane);  // C D(j)^T
    quad_store(l, Lt + (size_t)tix(i, j) * TT, TB, qi, qj, lane);
    publish(sync + lflag(g, i, j), epoch);
}

__device__ void task_inverse(const StepArgs& a, const Layout& g, int i, int j, int epoch,
                             double* lds) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int qi = wv >> 1, qj = wv & 1, T = g.T;
    int* sync = a.sync;
    int* ctl = sync + g.ctl;
    const double* Lt = a.ws + g.oL;
    double* Xt = a.ws + g.oX;
    double* Ys = lds;
    dbl4 acc = dbl4{0, 0, 0, 0};
    for (int p = j; p < i; ++p) {
        wait_for(sync + lflag(g, i, p), epoch, ctl);
        wait_for(p == j ? sync + g.fD + j : sync + g.fX + xid(T, p, j), epoch, ctl);
        tile_mma<false>(acc, Lt + (size_t)tix(i, p) * TT, TB, Xt + (size_t)tix(p, j) * TT, TB, 1.0,
                        qi, qj, lane);
    }
    quad_store(acc, Ys, LDT, qi, qj, lane);
    wait_for(sync + g.fD + i, epoch, ctl);
    dbl4 x = dbl4{0, 0, 0, 0};
    tile_mma<false>(x, Xt + (size_t)tix(i, i) * TT, TB, Ys, LDT, -1.0, qi, qj, lane);  // -D(i) Y
    quad_store(x, Xt + (size_t)tix(i, j) * TT, TB, qi, qj, lane);
    publish(sync + g.fX + xid(T, i, j), epoch);
}

// S, dS/d(nu, kappa, sigma_f) of every eigenvalue and sum_j log S_j; no
// dependencies, so it runs while the factorisation does
__device__ void task_filter(const StepArgs& a, const Layout& g, int epoch, double* lds) {
    const int tid = threadIdx.x, k = a.k;
    double* sh = lds;
    const double nu = softplus_d(a.u[0]) + a.lower[0];
    const double kappa = softplus_d(a.u[1]) + a.lower[1];
    const double sf = softplus_d(a.u[2]) + a.lower[2];
    double* S = a.ws + g.oS;
    double* dS = a.ws + g.oDS;
    double sr = 0.0, snu = 0.0, ska = 0.0;
    for (int q = tid; q < k; q += NTH) {
        double r, dn, dk;
        filter_terms(a.typ, a.lam[q], nu, kappa, r, dn, dk);
        sr += r;
        snu += r * dn;
        ska += r * dk;
    }
    sr = block_sum(sr, sh);
    snu = block_sum(snu, sh) / sr;
    ska = block_sum(ska, sh) / sr;
    double lS = 0.0;
    for (int q = tid; q < k; q += NTH) {
        double r, dn, dk;
        filter_terms(a.typ, a.lam[q], nu, kappa, r, dn, dk);
        const double v = sf * a.n_rows * r / sr;
        S[q] = v;
        dS[q] = v * (dn - snu);
        dS[k + q] = v * (dk - ska);
        dS[2 * k + q] = v / sf;
        lS += log(v);
    }
    lS = block_sum(lS, sh);
    if (tid == 0) a.ws[g.oSc] = lS;
    publish(a.sync + g.fF, epoch);
}

// t_i = sum_{j<=i} X(i,j) b_j, tile by tile as the X tiles arrive: thread
// (r = tid/8, q = tid%8) takes row r, columns 4q..4q+3 of every tile, then
// the 8 partial sums meet by shuffles
__device__ void task_row(const StepArgs& a, const Layout& g, int i, int epoch) {
    const int tid = threadIdx.x, T = g.T;
    int* sync = a.sync;
    int* ctl = sync + g.ctl;
    const double* Xt = a.ws + g.oX;
    const int r = tid >> 3, q = tid & 7;
    double s = 0.0;
    // every X(i, .) tile follows D(i) within a hand-off or two: wait for the
    // whole row at once
    wait_all([&](int j) { return j == i ? sync + g.fD + i : sync + g.fX + xid(T, i, j); }, i + 1,
             epoch, ctl);
    for (int j = 0; j <= i; ++j) {
        const double* X = Xt + (size_t)tix(i, j) * TT + r * TB + 4 * q;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int col = j * TB + 4 * q + c;
            s += X[c] * (col < a.k ? a.b[col] : 0.0);
        }
    }
    s += __shfl_xor(s, 1, 64);
    s += __shfl_xor(s, 2, 64);
    s += __shfl_xor(s, 4, 64);
    if (q == 0) a.ws[g.oT + (size_t)i * TB + r] = s;
    publish(sync + g.fR + i, epoch);
}

// Tile column j of w = X^T t and diag(A^-1) (thread (c = tid%32, q = tid/32)
// takes column c, rows 4q..4q+3 of every tile; the 8 partials are summed in
// q order), then this column's share of every sum the NLML and adjoint need:
//   P[0] = sum b_q w_q       P[1] = sum w_q^2 / S_q    P[2] = sum A^-1_qq / S_q
//   P[3] = sum log L_qq      P[4..6] = sum gS_q dS_q/d(nu, kappa, sigma_f)
// with gS_q = (1/S_q - s2 A^-1_qq / S_q^2 - w_q^2 / S_q^2) / 2.
__device__ void task_col(const StepArgs& a, const Layout& g, int j, int epoch, double* lds) {
    const int tid = threadIdx.x, T = g.T, k = a.k;
    int* sync = a.sync;
    int* ctl = sync + g.ctl;
    const double* Xt = a.ws + g.oX;
    const double* tv = a.ws + g.oT;
    const int c = tid & 31, q = tid >> 5;
    double sw = 0.0, sa = 0.0;
    for (int i = j; i < T; ++i) {
        wait_for(sync + g.fR + i, epoch, ctl);  // row i has seen every X(i, .)
        const double* X = Xt + (size_t)tix(i, j) * TT;
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
            const int r = 4 * q + rr;
            const double x = X[r * TB + c];
            sw += x * tv[i * TB + r];
            sa += x * x;
        }
    }
    double* pw = lds;            // 8 x 32
    double* pa = lds + 8 * TB;   // 8 x 32
    pw[q * TB + c] = sw;
    pa[q * TB + c] = sa;
    wait_for(sync + g.fF, epoch, ctl);  // S, dS (also orders the LDS stores)
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    if (tid < TB) {
        const int qq = j * TB + tid;
        double w = 0.0, av = 0.0;
#pragma unroll
        for (int h = 0; h < 8; ++h) {
            w += pw[h * TB + tid];
            av += pa[h * TB + tid];
        }
        a.ws[g.oW + qq] = w;
        a.ws[g.oA + qq] = av;
        if (qq < k) {
            const double s2 = softplus_d(a.u[3]) + a.lower[3];
            const double S = a.ws[g.oS + qq];
            const double* dS = a.ws + g.oDS;
            const double gS = 0.5 * (1.0 / S - s2 * av / (S * S) - w * w / (S * S));
            v[0] = a.b[qq] * w;
            v[1] = w * w / S;
            v[2] = av / S;
            v[3] = log(a.ws[g.oL + (size_t)tix(j, j) * TT + tid * TB + tid]);
            v[4] = gS * dS[qq];
            v[5] = gS * dS[k + qq];
            v[6] = gS * dS[2 * k + qq];
        }
    }
    if (tid < 64) {  // wave 0: lanes 32..63 carry zeros
#pragma unroll
        for (int m = 0; m < 7; ++m) v[m] = wave_sum(v[m]);
        if (tid == 0)
#pragma unroll
            for (int m = 0; m < 7; ++m) a.ws[g.oP + 8 * (size_t)j + m] = v[m];
    }
    publish(sync + g.fC + j, epoch);
}

// NLML and adjoint (gp.hip's weight-space finish, one right-hand side, with
// w^T G w = b^T w - s2 sum_j w_j^2 / S_j since A w = b) from the column
// partials, summed in tile order; then Adam.
__device__ void task_final(const StepArgs& a, const Layout& g, int epoch) {
    const int tid = threadIdx.x, T = g.T, k = a.k;
    int* sync = a.sync;
    int* ctl = sync + g.ctl;
    // the column tasks all finish after the last row: wait for them (and the
    // filter) at once
    wait_all([&](int j) { return j < T ? sync + g.fC + j : sync + g.fF; }, T + 1, epoch, ctl);
    if (tid < 4) {
        const int p = tid;
        double P[7] = {0, 0, 0, 0, 0, 0, 0};
        for (int j = 0; j < T; ++j)
#pragma unroll
            for (int m = 0; m < 7; ++m) P[m] += a.ws[g.oP + 8 * (size_t)j + m];
        const double bw = P[0], ww = P[1], tr = P[2], ld = P[3];
        const double lS = a.ws[g.oSc];
        const double s2 = softplus_d(a.u[3]) + a.lower[3];
        const double y2 = a.yy[0], M = a.M;
        const double two_pi = 6.283185307179586;
        const double nlml = 0.5 * (y2 - bw) / s2 + 0.5 * ((M - k) * log(s2) + lS + 2.0 * ld) +
                            0.5 * M * log(two_pi);
        const double wGw = bw - s2 * ww;
        const double res2 = y2 - 2.0 * bw + wGw;
        const double gs2 = 0.5 * ((M - k) / s2 + tr - res2 / (s2 * s2));
        const double gr[4] = {P[4], P[5], P[6], gs2};
        const double up = a.u[p];
        const double sig = 1.0 / (1.0 + exp(-up));
        const double gu = gr[p] * sig;
        if (p == 0) a.out[0] = nlml;
        a.out[1 + p] = gr[p];
        a.out[5 + p] = gu;
        if (a.do_adam) {
            const double step = a.tstep[0] + 1.0;
            if (a.trainable[p]) {
                const double m = a.b1 * a.m1[p] + (1.0 - a.b1) * gu;
                const double v = a.b2 * a.m2[p] + (1.0 - a.b2) * gu * gu;
                a.m1[p] = m;
                a.m2[p] = v;
                const double mh = m / (1.0 - pow(a.b1, step));
                const double vh = v / (1.0 - pow(a.b2, step));
                a.u[p] = up - a.lr * mh / (sqrt(vh) + a.eps);
            }
        }
    }
    __syncthreads();
    if (tid == 0) {
        if (a.do_adam) a.tstep[0] += 1.0;
        const int err = ld_flag(ctl + 1), nspd = ld_flag(ctl + 2);
        a.status[0] = err ? RVGP_ST_TIMEOUT : (nspd ? RVGP_ST_NOT_SPD : RVGP_ST_OK);
        st_flag(ctl + 1, 0);
        st_flag(ctl + 2, 0);
        st_flag(ctl, epoch);  // the next launch runs epoch + 1
    }
}

__global__ __launch_bounds__(NTH) void ws_step_kernel(StepArgs a) {
    __shared__ double lds[4 * TB * LDT + 16 * 17 + 64];
    __shared__ int s_epoch;
    const Layout g = make_layout(a.k);
    // the previous launch's final task stored the epoch; every task reads it
    // before it publishes anything, so the final task of THIS launch (which
    // depends on all of them) cannot have advanced it yet
    if (threadIdx.x == 0) s_epoch = ld_flag(a.sync + g.ctl) + 1;
    __syncthreads();
    const int epoch = s_epoch;
    const int T = g.T;
    int b = blockIdx.x;
    if (b < g.nL) {
        int j = 0;
        while (j + 1 < T && lbase(T, j + 1) <= b) ++j;
        const int off = b - lbase(T, j);
        if (off == 0) task_diag(a, g, j, epoch, lds);
        else task_offdiag(a, g, j + 1 + off, j, epoch, lds);
        return;
    }
    b -= g.nL;
    if (b < g.nX) {
        int j = 0;
        while (j + 1 < T - 1 && xbase(T, j + 1) <= b) ++j;
        const int i = j + 1 + (b - xbase(T, j));
        task_inverse(a, g, i, j, epoch, lds);
        return;
    }
    b -= g.nX;
    if (b == 0) { task_filter(a, g, epoch, lds); return; }
    b -= 1;
    if (b < T) { task_row(a, g, b, epoch); return; }
    b -= T;
    if (b < T) { task_col(a, g, b, epoch, lds); return; }
    task_final(a, g, epoch);
}

}  // namespace
}  // namespace rvgp

using namespace rvgp;

extern "C" size_t rvgp_ws_step_ws_bytes(int k) {
    if (k < 1 || k > RVGP_WS_STEP_MAX_K) return 0;
    return sizeof(double) * make_layout(k).nws;
}

extern "C" size_t rvgp_ws_step_sync_ints(int k) {
    if (k < 1 || k > RVGP_WS_STEP_MAX_K) return 0;
    return (size_t)make_layout(k).nsync;
}

extern "C" int rvgp_ws_step(int k, int typ, const double* evals, double n_rows, const double* G,
                            int ldg, const double* b, const double* yy, double M,
                            const double* lower, double* u, const int* trainable, double* m1,
                            double* m2, double* t, int do_adam, double lr, double b1, double b2,
                            double eps, void* ws, size_t ws_bytes, int* sync, size_t sync_ints,
                            double* out, int* status, void* stream) {
    RVGP_CHECK_ARG(k >= 1 && k <= RVGP_WS_STEP_MAX_K, "rvgp_ws_step: need 1 <= k <= %d",
                   RVGP_WS_STEP_MAX_K);
    RVGP_CHECK_ARG(ldg >= k, "rvgp_ws_step: ldg < k");
    RVGP_CHECK_ARG(typ == 0 || typ == 1, "rvgp_ws_step: typ must be 0 (matern) or 1 (se)");
    const Layout g = make_layout(k);
    RVGP_CHECK_ARG(ws_bytes >= sizeof(double) * g.nws, "rvgp_ws_step: workspace too small");
    RVGP_CHECK_ARG(sync_ints >= (size_t)g.nsync, "rvgp_ws_step: sync array too small");
    RVGP_CHECK_ARG(do_adam == 0 || (m1 && m2 && t && trainable),
                   "rvgp_ws_step: Adam state missing");
    StepArgs a;
    a.k = k;
    a.typ = typ;
    a.ldg = ldg;
    a.do_adam = do_adam;
    a.n_rows = n_rows;
    a.M = M;
    a.lr = lr;
    a.b1 = b1;
    a.b2 = b2;
    a.eps = eps;
    a.lam = evals;
    a.G = G;
    a.b = b;
    a.yy = yy;
    a.lower = lower;
    a.u = u;
    a.trainable = trainable;
    a.m1 = m1;
    a.m2 = m2;
    a.tstep = t;
    a.ws = static_cast<double*>(ws);
    a.sync = sync;
    a.out = out;
    a.status = status;
    hipLaunchKernelGGL(ws_step_kernel, dim3(g.nblocks), dim3(NTH), 0, as_stream(stream), a);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}
