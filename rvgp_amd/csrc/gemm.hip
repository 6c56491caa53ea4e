// Dense GEMM on the CDNA4 matrix cores:
//     C = alpha * op(A) * diag(s) * op(B) + beta * C
// with op = N or T on row-major operands, an optional per-k scale vector s
// (the spectral filter of the GP Gram Phi diag(S) Phi^T, kernels.py:55-61),
// and split-K for the tall-skinny Grams of the eigensolver (K = n*d rows,
// M, N = block width) whose output tile count alone cannot fill 256 CUs.
//
// Tiles: 64 x 64 per 256-thread workgroup, four waves in a 2 x 2 grid, each
// wave a 32 x 32 block of 2 x 2 MFMA 16x16 tiles.  A and B k-slices are
// staged through LDS k-major (a lane reading its fragment walks consecutive
// m / n -> conflict-free), loaded from global memory along whichever index
// is contiguous for the given op.
//
// Precision variants (TIn -> compute -> TOut):
//   f64 -> v_mfma_f64_16x16x4f64 -> f64
//   f32 -> v_mfma_f32_16x16x4f32 -> f32
//   f32 -> widened to f64 in LDS, v_mfma_f64_16x16x4f64 -> f64
//          (Rayleigh-Ritz Grams of f32 blocks accumulated in FP64)
//   f32 -> rounded to bf16 in LDS, v_mfma_f32_16x16x32_bf16 -> f32
//          (BF16 tiles with FP32 accumulation)
#include "common.h"

// The library builds with -ffp-contract=off (the distance kernels need a
// fixed operation order); the epilogues and the matrix-vector kernels use FMAs; results are compared
// with tolerances, never bit for bit.
#pragma clang fp contract(fast)

namespace rvgp {

constexpr int BM = 64, BN = 64, GT = 256;

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

enum class Mode { F64, F32, F32ACC64, BF16 };

template <Mode M>
struct Tr;
template <>
struct Tr<Mode::F64> {
    using In = double; using Lds = double; using Acc = double; using Vec = dbl4;
    static constexpr int BK = 16, KSTEP = 4;
};
template <>
struct Tr<Mode::F32> {
    using In = float; using Lds = float; using Acc = float; using Vec = flt4;
    static constexpr int BK = 16, KSTEP = 4;
};
template <>
struct Tr<Mode::F32ACC64> {
    using In = float; using Lds = double; using Acc = double; using Vec = dbl4;
    static constexpr int BK = 16, KSTEP = 4;
};
template <>
struct Tr<Mode::BF16> {
    using In = float; using Lds = __bf16; using Acc = float; using Vec = flt4;
    static constexpr int BK = 32, KSTEP = 32;
};

template <Mode M>
__device__ __forceinline__ typename Tr<M>::Lds to_lds(float v) {
    return static_cast<typename Tr<M>::Lds>(v);  // bf16: round to nearest even
}
template <Mode M>
__device__ __forceinline__ typename Tr<M>::Lds to_lds(double v) {
    return (typename Tr<M>::Lds)v;
}

// Load element (m, k) of op(X) where X is row-major with leading dim ld.
template <typename T>
__device__ __forceinline__ T ld_op(const T* X, int ld, bool trans, int m, int k) {
    return trans ? X[(size_t)k * ld + m] : X[(size_t)m * ld + k];
}

template <Mode MD>
__global__ __launch_bounds__(GT) void gemm_kernel(int transA, int transB, int M, int N, int K,
                                                  int kchunk, typename Tr<MD>::Acc alpha,
                                                  const typename Tr<MD>::In* __restrict__ A,
                                                  int lda, const typename Tr<MD>::In* __restrict__ B,
                                                  int ldb, const double* __restrict__ kscale,
                                                  typename Tr<MD>::Acc beta,
                                                  typename Tr<MD>::Acc* __restrict__ C, int ldc,
                                                  typename Tr<MD>::Acc* __restrict__ part,
                                                  int upper) {
    using In = typename Tr<MD>::In;
    using Lds = typename Tr<MD>::Lds;
    using Acc = typename Tr<MD>::Acc;
    using Vec = typename Tr<MD>::Vec;
    constexpr int BK = Tr<MD>::BK;
    constexpr int PAD = (sizeof(Lds) == 8) ? 1 : 2;
    __shared__ Lds As[BK][BM + PAD];
    __shared__ Lds Bs[BK][BN + PAD];

    if (upper && blockIdx.y < blockIdx.x) return;  // strictly-lower tile of a symmetric C
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    const int kbeg = blockIdx.z * kchunk;
    const int kend = min(K, kbeg + kchunk);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    const bool tA = transA != 0, tB = transB != 0;

    Vec acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = Vec{0, 0, 0, 0};

    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        // stage A: BM x BK elements, B: BK x BN elements
        for (int e = tid; e < BM * BK; e += GT) {
            int mm, kk;
            if (tA) { kk = e / BM; mm = e - kk * BM; }   // contiguous along m
            else { mm = e / BK; kk = e - mm * BK; }      // contiguous along k
            const int gm = m0 + mm, gk = k0 + kk;
            double v = 0.0;
            if (gm < M && gk < kend) {
                v = (double)ld_op(A, lda, tA, gm, gk);
                if (kscale) v *= kscale[gk];
            }
            if constexpr (MD == Mode::F32 || MD == Mode::BF16) As[kk][mm] = to_lds<MD>((float)v);
            else As[kk][mm] = (Lds)v;
        }
        for (int e = tid; e < BN * BK; e += GT) {
            int nn, kk;
            if (tB) { nn = e / BK; kk = e - nn * BK; }   // B^T: element (k, n) at B[n*ldb + k]
            else { kk = e / BN; nn = e - kk * BN; }      // B: element (k, n) at B[k*ldb + n]
            const int gn = n0 + nn, gk = k0 + kk;
            In v = 0;
            if (gn < N && gk < kend) v = tB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn];
            if constexpr (MD == Mode::F32 || MD == Mode::BF16) Bs[kk][nn] = to_lds<MD>((float)v);
            else Bs[kk][nn] = (Lds)v;
        }
        __syncthreads();
        if constexpr (MD == Mode::BF16) {
            // one 16x16x32 MFMA per sub-tile: lane holds 8 consecutive k
            bf16x8 a[2], b[2];
            const int kb = 8 * (lane >> 4);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    a[i][t] = As[kb + t][wm + i * 16 + (lane & 15)];
                    b[i][t] = Bs[kb + t][wn + i * 16 + (lane & 15)];
                }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        } else {
#pragma unroll
            for (int ks = 0; ks < BK; ks += 4) {
                const int kr = ks + (lane >> 4);
                Lds a[2], b[2];
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    a[i] = As[kr][wm + i * 16 + (lane & 15)];
                    b[i] = Bs[kr][wn + i * 16 + (lane & 15)];
                }
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        if constexpr (sizeof(Lds) == 8)
                            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
                        else
                            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
                    }
            }
        }
        __syncthreads();
    }
    // epilogue: lane holds column lane%16 of each 16x16 tile, in rows
    //   f32 / bf16 MFMA: 4*(lane/16) + r       (four consecutive rows)
    //   f64 MFMA:        lane/16 + 4*r         (rows interleaved by 4)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rr = (sizeof(Lds) == 8) ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r;
                const int gm = m0 + wm + i * 16 + rr;
                const int gn = n0 + wn + j * 16 + (lane & 15);
                if (gm < M && gn < N) {
                    const Acc v = acc[i][j][r];
                    if (part) {
                        part[(size_t)blockIdx.z * M * N + (size_t)gm * N + gn] = v;
                    } else {
                        Acc o = alpha * v;
                        if (beta != Acc(0)) o += beta * C[(size_t)gm * ldc + gn];
                        C[(size_t)gm * ldc + gn] = o;
                    }
                }
            }
}

// Large-tile variant: 128 x 128 per workgroup, each wave a 64 x 64 block of
// 4 x 4 MFMA tiles (half the LDS fragment traffic per MFMA of the 64 x 64
// kernel), with the next k-slice prefetched into registers while the current
// one is multiplied out of LDS.
constexpr int BBM = 128, BBN = 128;

template <Mode MD>
__global__ __launch_bounds__(GT) __attribute__((amdgpu_waves_per_eu(2))) void gemm_big_kernel(int transA, int transB, int M, int N, int K,
                                                      int kchunk, typename Tr<MD>::Acc alpha,
                                                      const typename Tr<MD>::In* __restrict__ A,
                                                      int lda,
                                                      const typename Tr<MD>::In* __restrict__ B,
                                                      int ldb, const double* __restrict__ kscale,
                                                      typename Tr<MD>::Acc beta,
                                                      typename Tr<MD>::Acc* __restrict__ C, int ldc,
                                                      typename Tr<MD>::Acc* __restrict__ part,
                                                      int upper, int vec) {
    using In = typename Tr<MD>::In;
    using Lds = typename Tr<MD>::Lds;
    using Acc = typename Tr<MD>::Acc;
    using Vec = typename Tr<MD>::Vec;
    // The prefetch registers hold the raw input: widening (f32 -> f64), the
    // k-scale and the bf16 rounding happen at the LDS store, after the loads
    // have landed, so the wait for them stays one k-slice behind their issue.
    using Reg = In;
    constexpr int BK = Tr<MD>::BK;
    constexpr int PAD = (sizeof(Lds) == 8) ? 1 : 4;
    constexpr int EA = BBM * BK / GT;  // elements of each tile staged per thread
    // 16-byte loads of FP32 inputs along the contiguous index of each operand
    // (when `vec`: 16-byte aligned bases, leading dimensions a multiple of 4;
    // FP64 inputs load element by element, which measured faster)
    constexpr int VW = sizeof(In) == 4 ? 4 : 1;
    constexpr int EV = EA / VW;        // vectors of each tile per thread
    __shared__ Lds As[BK][BBM + PAD];
    __shared__ Lds Bs[BK][BBN + PAD];

    if (upper && blockIdx.y < blockIdx.x) return;  // strictly-lower tile of a symmetric C
    const int m0 = blockIdx.x * BBM, n0 = blockIdx.y * BBN;
    const int kbeg = blockIdx.z * kchunk;
    const int kend = min(K, kbeg + kchunk);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = (w >> 1) * 64, wn = (w & 1) * 64;
    const bool tA = transA != 0, tB = transB != 0;

    Vec acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = Vec{0, 0, 0, 0};

    Reg ra[EA], rb[EA];
    // Element q of this thread's share: vector v = tid + GT * (q / VW), lane
    // t = q % VW of it.  A (m, k): along k when row-major, along m when
    // transposed; B (k, n): along n when row-major, along k when transposed.
#define RVGP_A_AT(q, mm, kk)                                                        \
    const int mm = tA ? ((tid + GT * ((q) / VW)) % (BBM / VW)) * VW + (q) % VW     \
                      : (tid + GT * ((q) / VW)) / (BK / VW);                       \
    const int kk = tA ? (tid + GT * ((q) / VW)) / (BBM / VW)                       \
                      : ((tid + GT * ((q) / VW)) % (BK / VW)) * VW + (q) % VW;
#define RVGP_B_AT(q, nn, kk)                                                        \
    const int nn = tB ? (tid + GT * ((q) / VW)) / (BK / VW)                        \
                      : ((tid + GT * ((q) / VW)) % (BBN / VW)) * VW + (q) % VW;    \
    const int kk = tB ? ((tid + GT * ((q) / VW)) % (BK / VW)) * VW + (q) % VW      \
                      : (tid + GT * ((q) / VW)) / (BBN / VW);
    typedef In InV __attribute__((ext_vector_type(VW > 1 ? VW : 2)));
    auto load = [&](int k0) {
#pragma unroll
        for (int u = 0; u < EV; ++u) {
            {
                RVGP_A_AT(u * VW, mm, kk)
                const int gm = m0 + mm, gk = k0 + kk;
                const In* src = tA ? A + (size_t)gk * lda + gm : A + (size_t)gm * lda + gk;
                if (VW > 1 && vec && (tA ? (gk < kend && gm + VW <= M) : (gm < M && gk + VW <= kend))) {
                    const InV x = *reinterpret_cast<const InV*>(src);
#pragma unroll
                    for (int t = 0; t < VW; ++t) ra[u * VW + t] = x[t];
                } else {
#pragma unroll
                    for (int t = 0; t < VW; ++t) {
                        const bool in = tA ? (gm + t < M && gk < kend) : (gm < M && gk + t < kend);
                        ra[u * VW + t] = in ? src[t] : In(0);
                    }
                }
            }
            {
                RVGP_B_AT(u * VW, nn, kk)
                const int gn = n0 + nn, gk = k0 + kk;
                const In* src = tB ? B + (size_t)gn * ldb + gk : B + (size_t)gk * ldb + gn;
                if (VW > 1 && vec && (tB ? (gn < N && gk + VW <= kend) : (gk < kend && gn + VW <= N))) {
                    const InV x = *reinterpret_cast<const InV*>(src);
#pragma unroll
                    for (int t = 0; t < VW; ++t) rb[u * VW + t] = x[t];
                } else {
#pragma unroll
                    for (int t = 0; t < VW; ++t) {
                        const bool in = tB ? (gn < N && gk + t < kend) : (gn + t < N && gk < kend);
                        rb[u * VW + t] = in ? src[t] : In(0);
                    }
                }
            }
        }
    };
    auto store = [&](int k0) {
#pragma unroll
        for (int q = 0; q < EA; ++q) {
            {
                RVGP_A_AT(q, mm, kk)
                if (kscale != nullptr && k0 + kk < kend) {
                    const double v = (double)ra[q] * kscale[k0 + kk];
                    if constexpr (MD == Mode::F32 || MD == Mode::BF16) As[kk][mm] = to_lds<MD>((float)v);
                    else As[kk][mm] = (Lds)v;
                } else {
                    As[kk][mm] = to_lds<MD>(ra[q]);
                }
            }
            {
                RVGP_B_AT(q, nn, kk)
                Bs[kk][nn] = static_cast<Lds>(rb[q]);
            }
        }
    };
#undef RVGP_A_AT
#undef RVGP_B_AT

    if (kbeg < kend) load(kbeg);
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        store(k0);
        __syncthreads();
        if (k0 + BK < kend) load(k0 + BK);
        if constexpr (MD == Mode::BF16) {
            bf16x8 a[4], b[4];
            const int kb = 8 * (lane >> 4);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    a[i][t] = As[kb + t][wm + i * 16 + (lane & 15)];
                    b[i][t] = Bs[kb + t][wn + i * 16 + (lane & 15)];
                }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        } else {
#pragma unroll
            for (int ks = 0; ks < BK; ks += 4) {
                const int kr = ks + (lane >> 4);
                Lds a[4], b[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    a[i] = As[kr][wm + i * 16 + (lane & 15)];
                    b[i] = Bs[kr][wn + i * 16 + (lane & 15)];
                }
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        if constexpr (sizeof(Lds) == 8)
                            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
                        else
                            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
                    }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rr = (sizeof(Lds) == 8) ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r;
                const int gm = m0 + wm + i * 16 + rr;
                const int gn = n0 + wn + j * 16 + (lane & 15);
                if (gm < M && gn < N) {
                    const Acc v = acc[i][j][r];
                    if (part) {
                        part[(size_t)blockIdx.z * M * N + (size_t)gm * N + gn] = v;
                    } else {
                        Acc o = alpha * v;
                        if (beta != Acc(0)) o += beta * C[(size_t)gm * ldc + gn];
                        C[(size_t)gm * ldc + gn] = o;
                    }
                }
            }
}

// Fixed-order reduction of split-K partials: C = alpha * sum_z part[z] + beta * C.
template <typename Acc>
__global__ void splitk_reduce(int M, int N, int S, Acc alpha, const Acc* __restrict__ part,
                              Acc beta, Acc* __restrict__ C, int ldc) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)M * N) return;
    Acc s = 0;
    for (int z = 0; z < S; ++z) s += part[(size_t)z * M * N + e];
    const size_t gm = e / N, gn = e - gm * N;
    Acc o = alpha * s;
    if (beta != Acc(0)) o += beta * C[gm * ldc + gn];
    C[gm * ldc + gn] = o;
}

// Tile size and K split.  split_k <= 0 picks the split: enough K slices for
// ~1024 workgroups (the tall-skinny Grams have few output tiles), at least 8
// k-steps per slice, at most 256 slices; the split depends only on the shape,
// so results are reproducible.
// Small GEMM for the latency-bound regime (few output tiles, K <= 1024: the
// Cholesky / triangular-solve updates and matrix-vector products of the GP
// step).  One wavefront per 16 x 16 output tile; the MFMA operands are loaded
// straight from global memory into registers (no LDS, no barriers), 16
// k-steps per batch with the next batch's loads in flight during the current
// batch's MFMAs, and C (when beta != 0) fetched before the k loop.  `lower`
// restricts the output to the lower triangle (Cholesky trailing update).
constexpr int SU = 16;

template <Mode MD>
__global__ __launch_bounds__(256) void gemm_small_kernel(
    int transA, int transB, int M, int N, int K, typename Tr<MD>::Acc alpha,
    const typename Tr<MD>::In* __restrict__ A, int lda, const typename Tr<MD>::In* __restrict__ B,
    int ldb, const double* __restrict__ kscale, typename Tr<MD>::Acc beta,
    typename Tr<MD>::Acc* __restrict__ C, int ldc, int lower, int tiles_n) {
    using In = typename Tr<MD>::In;
    using Acc = typename Tr<MD>::Acc;
    using Vec = typename Tr<MD>::Vec;
    using Op = typename Tr<MD>::Lds;  // MFMA operand type: f64, or f32 for Mode::F32
    constexpr bool F64OUT = sizeof(Op) == 8;
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int ti = tile / tiles_n, tj = tile - ti * tiles_n;
    const int m0 = ti * 16, n0 = tj * 16;
    if (m0 >= M) return;
    if (lower && n0 > m0 + 15) return;  // tile strictly above the diagonal
    const int am = m0 + (lane & 15), bn = n0 + (lane & 15), kq = lane >> 4;
    const bool mok = am < M, nok = bn < N;

    Acc cpre[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = m0 + (F64OUT ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r);
        cpre[r] = (beta != Acc(0) && row < M && nok) ? C[(size_t)row * ldc + bn] : Acc(0);
    }
    In ra0[SU], rb0[SU], ra1[SU], rb1[SU];
    auto load = [&](In (&ra)[SU], In (&rb)[SU], int k0) {
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            const int k = k0 + 4 * u + kq;
            ra[u] = (mok && k < K) ? (transA ? A[(size_t)k * lda + am] : A[(size_t)am * lda + k]) : In(0);
            rb[u] = (nok && k < K) ? (transB ? B[(size_t)bn * ldb + k] : B[(size_t)k * ldb + bn]) : In(0);
        }
    };
    Vec acc = Vec{0, 0, 0, 0};
    auto compute = [&](In (&ra)[SU], In (&rb)[SU], int k0) {
#pragma unroll
        for (int u = 0; u < SU; ++u) {
            if (k0 + 4 * u < K) {  // wave-uniform
                Op a = (Op)ra[u];
                const int k = k0 + 4 * u + kq;
                if (kscale != nullptr && k < K) a = (Op)((double)ra[u] * kscale[k]);
                if constexpr (F64OUT)
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, (Op)rb[u], acc, 0, 0, 0);
                else
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, (Op)rb[u], acc, 0, 0, 0);
            }
        }
    };
    load(ra0, rb0, 0);
    for (int k0 = 0; k0 < K; k0 += 8 * SU) {
        if (k0 + 4 * SU < K) load(ra1, rb1, k0 + 4 * SU);
        compute(ra0, rb0, k0);
        if (k0 + 4 * SU >= K) break;
        if (k0 + 8 * SU < K) load(ra0, rb0, k0 + 8 * SU);
        compute(ra1, rb1, k0 + 4 * SU);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = m0 + (F64OUT ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r);
        if (row < M && nok && (!lower || bn <= row)) {
            Acc o = alpha * acc[r];
            if (beta != Acc(0)) o += beta * cpre[r];
            C[(size_t)row * ldc + bn] = o;
        }
    }
}

// Matrix-vector product (N = 1, K <= 8192, no k-scale):
//     y = alpha op(A) x + beta y.
// op = N: one wavefront per row, lanes striding K (coalesced), wave sum.
// op = T: a workgroup per 64 outputs, lane = output column (coalesced along
// the row), sixteen waves splitting K (eight loads in flight per lane) and
// reducing through LDS in a fixed order.  Accumulates in Acc (f64, or f32
// for Mode::F32).
template <Mode MD>
__global__ __launch_bounds__(1024) void gemv_kernel(int transA, int M, int K,
                                                   typename Tr<MD>::Acc alpha,
                                                   const typename Tr<MD>::In* __restrict__ A,
                                                   int lda, const typename Tr<MD>::In* __restrict__ x,
                                                   int incx, typename Tr<MD>::Acc beta,
                                                   typename Tr<MD>::Acc* __restrict__ y, int incy) {
    using Acc = typename Tr<MD>::Acc;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (!transA) {
        const int m = blockIdx.x * 4 + w;
        if (m >= M) return;
        const auto* row = A + (size_t)m * lda;
        Acc p[4] = {Acc(0), Acc(0), Acc(0), Acc(0)};
        int k = lane, t = 0;
        for (; k < K; k += 64, t = (t + 1) & 3) p[t] += (Acc)row[k] * (Acc)x[(size_t)k * incx];
        Acc s = wave_sum((p[0] + p[1]) + (p[2] + p[3]));
        if (lane == 0) {
            Acc o = alpha * s;
            if (beta != Acc(0)) o += beta * y[(size_t)m * incy];
            y[(size_t)m * incy] = o;
        }
    } else {
        constexpr int NW = 16;
        __shared__ Acc part[NW][64];
        const int m = blockIdx.x * 64 + lane;
        const int per = (K + NW - 1) / NW, k0 = w * per, k1 = min(K, k0 + per);
        Acc p[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) p[t] = Acc(0);
        if (m < M) {
            int k = k0;
            for (; k + 7 < k1; k += 8) {
#pragma unroll
                for (int t = 0; t < 8; ++t)
                    p[t] += (Acc)A[(size_t)(k + t) * lda + m] * (Acc)x[(size_t)(k + t) * incx];
            }
            for (int t = 0; k < k1; ++k, ++t)
                p[t] += (Acc)A[(size_t)k * lda + m] * (Acc)x[(size_t)k * incx];
        }
        part[w][lane] = ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]));
        __syncthreads();
        if (w == 0 && m < M) {
            Acc o = Acc(0);
#pragma unroll
            for (int q = 0; q < NW; ++q) o += part[q][lane];
            o *= alpha;
            if (beta != Acc(0)) o += beta * y[(size_t)m * incy];
            y[(size_t)m * incy] = o;
        }
    }
}

// C[i][j] = C[j][i] for i > j: completes a symmetric C whose upper tiles
// were computed.
template <typename Acc>
__global__ void mirror_lower(int n, Acc* __restrict__ C, int ldc) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * n) return;
    const int i = (int)(e / n), j = (int)(e - (size_t)i * n);
    if (i > j) C[(size_t)i * ldc + j] = C[(size_t)j * ldc + i];
}

// 16-byte vector loads in the 128 x 128 kernel: both operands 16-byte
// aligned with leading dimensions a multiple of the vector width.
static inline int vec_ok(const void* A, int lda, const void* B, int ldb, int elem) {
    const int vw = 16 / elem;
    return ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(B)) & 15) == 0 &&
           lda % vw == 0 && ldb % vw == 0;
}

struct GemmPlan {
    bool small, big;
    int split, kchunk;
};

// Workgroups of the 128 x 128 FP32 kernel resident at once on the device.
// Queried once and cached, the fallback included, so that the workspace size
// query and the launch that follows it always plan the same split.
static int f32_big_slots() {
    static int slots = 0;  // the same for every device of one process's type
    if (slots == 0) {
        int dev = 0, per = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, gemm_big_kernel<Mode::F32>, GT, 0) !=
                hipSuccess)
            slots = 768;
        else
            slots = (per > 0 ? per : 1) * prop.multiProcessorCount;
    }
    return slots;
}

// small: K <= 1024 and at most 64 tiles of the 128 x 128 kernel -- too few
// workgroups to hide the k-loop latency, so one wave per 16 x 16 tile wins.
static GemmPlan gemm_plan(int mode, int M, int N, int K, int BK, int split_k) {
    GemmPlan p;
    p.small = mode != RVGP_GEMM_BF16 && split_k <= 1 && K <= 1024 &&
              (long)cdiv(M, BBM) * cdiv(N, BBN) <= 64;
    p.big = M >= 96 && N >= 96;
    if (p.small) {
        p.split = 1;
        p.kchunk = K;
        return p;
    }
    const long tiles = p.big ? (long)cdiv(M, BBM) * cdiv(N, BBN) : (long)cdiv(M, BM) * cdiv(N, BN);
    int split = split_k;
    if (split <= 0) {
        split = 1;
        if (K >= 2048) {
            long s = (1024 + tiles - 1) / tiles;
            long smax = K / (8 * BK);
            if (s > smax) s = smax;
            if (s > 256) s = 256;
            split = s < 1 ? 1 : (int)s;
        }
    }
    int kchunk = (K + split - 1) / split;
    kchunk = ((kchunk + BK - 1) / BK) * BK;
    if (kchunk < BK) kchunk = BK;
    p.kchunk = kchunk;
    p.split = K > 0 ? (K + kchunk - 1) / kchunk : 1;
    return p;
}

template <Mode MD>
static int gemm_launch(int transA, int transB, int M, int N, int K, double alpha, const void* A,
                       int lda, const void* B, int ldb, const double* kscale, double beta,
                       void* C, int ldc, int split, void* ws, size_t ws_bytes, void* stream,
                       int lower = 0, int upper = 0) {
    using In = typename Tr<MD>::In;
    using Acc = typename Tr<MD>::Acc;
    RVGP_CHECK_ARG(M >= 0 && N >= 0 && K >= 0, "rvgp_gemm: negative size");
    RVGP_CHECK_ARG(ldc >= N, "rvgp_gemm: ldc < N");
    RVGP_CHECK_ARG(lda >= (transA ? M : K) && ldb >= (transB ? K : N), "rvgp_gemm: bad lda/ldb");
    if (M == 0 || N == 0) return RVGP_OK;
    hipStream_t s = as_stream(stream);
    constexpr int MODE = MD == Mode::F64 ? RVGP_GEMM_F64
                         : MD == Mode::F32 ? RVGP_GEMM_F32
                         : MD == Mode::F32ACC64 ? RVGP_GEMM_F32_ACC64 : RVGP_GEMM_BF16;
    const GemmPlan pl = gemm_plan(MODE, M, N, K, Tr<MD>::BK, split);
    if constexpr (MD != Mode::BF16) {
        if (N == 1 && K <= 8192 && kscale == nullptr && !lower && !upper && split <= 1) {
            const int incx = transB ? 1 : ldb;
            if (transA)
                hipLaunchKernelGGL(gemv_kernel<MD>, dim3(cdiv(M, 64)), dim3(1024), 0, s, 1, M, K,
                                   (Acc)alpha, static_cast<const In*>(A), lda,
                                   static_cast<const In*>(B), incx, (Acc)beta,
                                   static_cast<Acc*>(C), ldc);
            else
                hipLaunchKernelGGL(gemv_kernel<MD>, dim3(cdiv(M, 4)), dim3(256), 0, s, 0, M, K,
                                   (Acc)alpha, static_cast<const In*>(A), lda,
                                   static_cast<const In*>(B), incx, (Acc)beta,
                                   static_cast<Acc*>(C), ldc);
            RVGP_LAUNCH_CHECK();
            return RVGP_OK;
        }
    }
    if (pl.small && !upper) {
        if constexpr (MD != Mode::BF16) {
            const int tn = (int)cdiv(N, 16);
            const size_t tiles = (size_t)cdiv(M, 16) * tn;
            hipLaunchKernelGGL(gemm_small_kernel<MD>, dim3(cdiv(tiles, 4)), dim3(256), 0, s, transA,
                               transB, M, N, K, (Acc)alpha, static_cast<const In*>(A), lda,
                               static_cast<const In*>(B), ldb, kscale, (Acc)beta,
                               static_cast<Acc*>(C), ldc, lower, tn);
            RVGP_LAUNCH_CHECK();
            return RVGP_OK;
        }
    }
    Acc* part = nullptr;
    if (pl.split > 1) {
        const size_t need = sizeof(Acc) * (size_t)pl.split * M * N;
        RVGP_CHECK_ARG(ws != nullptr && ws_bytes >= need,
                       "rvgp_gemm: split-K needs %zu bytes of workspace", need);
        part = static_cast<Acc*>(ws);
    }
    if (pl.big) {
        dim3 grid(cdiv(M, BBM), cdiv(N, BBN), pl.split);
        hipLaunchKernelGGL(gemm_big_kernel<MD>, grid, dim3(GT), 0, s, transA, transB, M, N, K,
                           pl.kchunk, (Acc)alpha, static_cast<const In*>(A), lda,
                           static_cast<const In*>(B), ldb, kscale, (Acc)beta, static_cast<Acc*>(C),
                           ldc, part, upper, vec_ok(A, lda, B, ldb, (int)sizeof(In)));
    } else {
        dim3 grid(cdiv(M, BM), cdiv(N, BN), pl.split);
        hipLaunchKernelGGL(gemm_kernel<MD>, grid, dim3(GT), 0, s, transA, transB, M, N, K, pl.kchunk,
                           (Acc)alpha, static_cast<const In*>(A), lda, static_cast<const In*>(B), ldb,
                           kscale, (Acc)beta, static_cast<Acc*>(C), ldc, part, upper);
    }
    RVGP_LAUNCH_CHECK();
    if (pl.split > 1) {
        hipLaunchKernelGGL(splitk_reduce<Acc>, dim3(cdiv((size_t)M * N, 256)), dim3(256), 0, s, M, N,
                           pl.split, (Acc)alpha, part, (Acc)beta, static_cast<Acc*>(C), ldc);
        RVGP_LAUNCH_CHECK();
    }
    if (upper) {
        hipLaunchKernelGGL(mirror_lower<Acc>, dim3(cdiv((size_t)M * M, 256)), dim3(256), 0, s, M,
                           static_cast<Acc*>(C), ldc);
        RVGP_LAUNCH_CHECK();
    }
    return RVGP_OK;
}

// Symmetric Gram of FP32 blocks on the FP32 matrix cores (twice the FP64
// MFMA rate): K is cut into chunks of at most kF32Chunk rows, each chunk
// accumulated in FP32 by the 128 x 128 kernel (upper tiles only), and the
// chunk partials are summed in FP64 in chunk order.  A chunk of 4096 rows
// carries ~sqrt(4096) * 6e-8 = 4e-6 relative rounding, and the FP64 sum of
// the independent chunk errors ~1e-6 -- below the FP32 data's own level in
// a Rayleigh-Ritz step of an FP32 block.
constexpr int kF32Chunk = 4096;

static GemmPlan gram_f32_plan(int M, int K, int split_k) {
    GemmPlan p = gemm_plan(RVGP_GEMM_F32, M, M, K, Tr<Mode::F32>::BK, split_k);
    const int need = (K + kF32Chunk - 1) / kF32Chunk;
    if (p.small || p.split < need) {
        p.small = false;
        p.big = M >= 96;
        p.split = need < 1 ? 1 : need;
        if (p.big && split_k <= 0) {
            // Fill whole rounds of resident workgroups: with T upper tiles and
            // `slots` resident at once, S chunks take ceil(T S / slots) rounds
            // of K / S rows each, so S is raised to the largest count that
            // still fits the rounds the minimum count needs.
            const long nt = cdiv(M, BBM), T = nt * (nt + 1) / 2, slots = f32_big_slots();
            const long rounds = (T * p.split + slots - 1) / slots;
            const long s = slots * rounds / T;
            if (s > p.split && s <= 4 * (long)p.split) p.split = (int)s;
        }
        int kchunk = (K + p.split - 1) / p.split;
        kchunk = ((kchunk + Tr<Mode::F32>::BK - 1) / Tr<Mode::F32>::BK) * Tr<Mode::F32>::BK;
        p.kchunk = kchunk < Tr<Mode::F32>::BK ? Tr<Mode::F32>::BK : kchunk;
        p.split = K > 0 ? (K + p.kchunk - 1) / p.kchunk : 1;
    }
    return p;
}

__global__ void splitk_reduce_f32_f64(int M, int N, int S, const float* __restrict__ part,
                                      double* __restrict__ C, int ldc) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)M * N) return;
    double acc = 0.0;
    for (int z = 0; z < S; ++z) acc += (double)part[(size_t)z * M * N + e];
    const size_t gm = e / N, gn = e - gm * N;
    C[gm * ldc + gn] = acc;
}

static int gram_f32_chunked(int M, int K, const float* A, int lda, const float* B, int ldb,
                            double* C, int ldc, int split_k, void* ws, size_t ws_bytes,
                            void* stream) {
    RVGP_CHECK_ARG(M >= 0 && K >= 0 && ldc >= M && lda >= M && ldb >= M,
                   "rvgp_gemm_sym: bad sizes");
    if (M == 0) return RVGP_OK;
    hipStream_t s = as_stream(stream);
    const GemmPlan pl = gram_f32_plan(M, K, split_k);
    const size_t need = sizeof(float) * (size_t)pl.split * M * M;
    RVGP_CHECK_ARG(ws != nullptr && ws_bytes >= need,
                   "rvgp_gemm_sym: the FP32 Gram needs %zu bytes of workspace", need);
    float* part = static_cast<float*>(ws);
    if (pl.big) {
        dim3 grid(cdiv(M, BBM), cdiv(M, BBN), pl.split);
        hipLaunchKernelGGL(gemm_big_kernel<Mode::F32>, grid, dim3(GT), 0, s, 1, 0, M, M, K,
                           pl.kchunk, 1.0f, A, lda, B, ldb, nullptr, 0.0f, nullptr, M, part, 1,
                           vec_ok(A, lda, B, ldb, (int)sizeof(float)));
    } else {
        dim3 grid(cdiv(M, BM), cdiv(M, BN), pl.split);
        hipLaunchKernelGGL(gemm_kernel<Mode::F32>, grid, dim3(GT), 0, s, 1, 0, M, M, K, pl.kchunk,
                           1.0f, A, lda, B, ldb, nullptr, 0.0f, nullptr, M, part, 1);
    }
    RVGP_LAUNCH_CHECK();
    hipLaunchKernelGGL(splitk_reduce_f32_f64, dim3(cdiv((size_t)M * M, 256)), dim3(256), 0, s, M, M,
                       pl.split, part, C, ldc);
    RVGP_LAUNCH_CHECK();
    hipLaunchKernelGGL(mirror_lower<double>, dim3(cdiv((size_t)M * M, 256)), dim3(256), 0, s, M, C,
                       ldc);
    RVGP_LAUNCH_CHECK();
    return RVGP_OK;
}

}  // namespace rvgp

using namespace rvgp;

extern "C" int rvgp_gemm(int mode, int transA, int transB, int M, int N, int K, double alpha,
                         const void* A, int lda, const void* B, int ldb, const double* kscale,
                         double beta, void* C, int ldc, int split_k, void* ws, size_t ws_bytes,
                         void* stream) {
    switch (mode) {
        case RVGP_GEMM_F64:
            return gemm_launch<Mode::F64>(transA, transB, M, N, K, alpha, A, lda, B, ldb, kscale,
                                          beta, C, ldc, split_k, ws, ws_bytes, stream);
        case RVGP_GEMM_F32:
            return gemm_launch<Mode::F32>(transA, transB, M, N, K, alpha, A, lda, B, ldb, kscale,
                                          beta, C, ldc, split_k, ws, ws_bytes, stream);
        case RVGP_GEMM_F32_ACC64:
            return gemm_launch<Mode::F32ACC64>(transA, transB, M, N, K, alpha, A, lda, B, ldb,
                                               kscale, beta, C, ldc, split_k, ws, ws_bytes, stream);
        case RVGP_GEMM_BF16:
            return gemm_launch<Mode::BF16>(transA, transB, M, N, K, alpha, A, lda, B, ldb, kscale,
                                           beta, C, ldc, split_k, ws, ws_bytes, stream);
        default:
            set_error("rvgp_gemm: unknown mode %d", mode);
            return RVGP_EINVAL;
    }
}

extern "C" size_t rvgp_gemm_ws_bytes(int mode, int M, int N, int K, int split_k) {
    if (M <= 0 || N <= 0) return 0;
    const int BK = mode == RVGP_GEMM_BF16 ? Tr<Mode::BF16>::BK : Tr<Mode::F64>::BK;
    const size_t acc = (mode == RVGP_GEMM_F64 || mode == RVGP_GEMM_F32_ACC64) ? 8 : 4;
    const GemmPlan p = gemm_plan(mode, M, N, K, BK, split_k);
    size_t bytes = p.split > 1 ? acc * (size_t)p.split * M * N : 0;
    if (mode == RVGP_GEMM_F32 && M == N) {  // also covers rvgp_gemm_sym's chunked FP32 Gram
        const GemmPlan q = gram_f32_plan(M, K, split_k);
        const size_t b2 = sizeof(float) * (size_t)q.split * M * N;
        if (b2 > bytes) bytes = b2;
    }
    return bytes;
}

extern "C" int rvgp_gemm_sym(int mode, int M, int K, const void* A, int lda, const void* B,
                             int ldb, double* C, int ldc, int split_k, void* ws, size_t ws_bytes,
                             void* stream) {
    RVGP_CHECK_ARG(mode == RVGP_GEMM_F64 || mode == RVGP_GEMM_F32_ACC64 || mode == RVGP_GEMM_F32,
                   "rvgp_gemm_sym: mode must be F64, F32_ACC64 or F32");
    if (mode == RVGP_GEMM_F32)
        return gram_f32_chunked(M, K, static_cast<const float*>(A), lda,
                                static_cast<const float*>(B), ldb, C, ldc, split_k, ws, ws_bytes,
                                stream);
    if (mode == RVGP_GEMM_F64)
        return gemm_launch<Mode::F64>(1, 0, M, M, K, 1.0, A, lda, B, ldb, nullptr, 0.0, C, ldc,
                                      split_k, ws, ws_bytes, stream, 0, 1);
    return gemm_launch<Mode::F32ACC64>(1, 0, M, M, K, 1.0, A, lda, B, ldb, nullptr, 0.0, C, ldc,
                                       split_k, ws, ws_bytes, stream, 0, 1);
}

namespace rvgp {
// C = alpha A A^T + beta C on (at least) the lower triangle of the n x n C;
// A is n x k row-major.  Internal: the Cholesky trailing update.
int gemm_syrk_lower(int f64, int n, int k, double alpha, const void* A, int lda, double beta,
                    void* C, int ldc, void* stream) {
    if (f64)
        return gemm_launch<Mode::F64>(0, 1, n, n, k, alpha, A, lda, A, lda, nullptr, beta, C, ldc,
                                      1, nullptr, 0, stream, 1);
    return gemm_launch<Mode::F32>(0, 1, n, n, k, alpha, A, lda, A, lda, nullptr, beta, C, ldc, 1,
                                  nullptr, 0, stream, 1);
}
}  // namespace rvgp
