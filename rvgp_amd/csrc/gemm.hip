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
                                                  typename Tr<MD>::Acc* __restrict__ part) {
    using In = typename Tr<MD>::In;
    using Lds = typename Tr<MD>::Lds;
    using Acc = typename Tr<MD>::Acc;
    using Vec = typename Tr<MD>::Vec;
    constexpr int BK = Tr<MD>::BK;
    constexpr int PAD = (sizeof(Lds) == 8) ? 1 : 2;
    __shared__ Lds As[BK][BM + PAD];
    __shared__ Lds Bs[BK][BN + PAD];

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
__global__ __launch_bounds__(GT) void gemm_big_kernel(int transA, int transB, int M, int N, int K,
                                                      int kchunk, typename Tr<MD>::Acc alpha,
                                                      const typename Tr<MD>::In* __restrict__ A,
                                                      int lda,
                                                      const typename Tr<MD>::In* __restrict__ B,
                                                      int ldb, const double* __restrict__ kscale,
                                                      typename Tr<MD>::Acc beta,
                                                      typename Tr<MD>::Acc* __restrict__ C, int ldc,
                                                      typename Tr<MD>::Acc* __restrict__ part) {
    using In = typename Tr<MD>::In;
    using Lds = typename Tr<MD>::Lds;
    using Acc = typename Tr<MD>::Acc;
    using Vec = typename Tr<MD>::Vec;
    using Reg = typename std::conditional<MD == Mode::BF16, float, Lds>::type;
    constexpr int BK = Tr<MD>::BK;
    constexpr int PAD = (sizeof(Lds) == 8) ? 1 : 2;
    constexpr int EA = BBM * BK / GT;  // elements of each tile staged per thread
    __shared__ Lds As[BK][BBM + PAD];
    __shared__ Lds Bs[BK][BBN + PAD];

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
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < EA; ++q) {
            const int e = tid + GT * q;
            int mm, kk;
            if (tA) { kk = e / BBM; mm = e - kk * BBM; }
            else { mm = e / BK; kk = e - mm * BK; }
            const int gm = m0 + mm, gk = k0 + kk;
            double v = 0.0;
            if (gm < M && gk < kend) {
                v = (double)(tA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk]);
                if (kscale) v *= kscale[gk];
            }
            ra[q] = (Reg)v;
            int nn;
            if (tB) { nn = e / BK; kk = e - nn * BK; }
            else { kk = e / BBN; nn = e - kk * BBN; }
            const int gn = n0 + nn, gk2 = k0 + kk;
            In u = 0;
            if (gn < N && gk2 < kend) u = tB ? B[(size_t)gn * ldb + gk2] : B[(size_t)gk2 * ldb + gn];
            rb[q] = (Reg)u;
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int q = 0; q < EA; ++q) {
            const int e = tid + GT * q;
            int mm, kk;
            if (tA) { kk = e / BBM; mm = e - kk * BBM; }
            else { mm = e / BK; kk = e - mm * BK; }
            As[kk][mm] = static_cast<Lds>(ra[q]);
            int nn;
            if (tB) { nn = e / BK; kk = e - nn * BK; }
            else { kk = e / BBN; nn = e - kk * BBN; }
            Bs[kk][nn] = static_cast<Lds>(rb[q]);
        }
    };

    if (kbeg < kend) load(kbeg);
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        store();
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
struct GemmPlan {
    bool big;
    int split, kchunk;
};

static GemmPlan gemm_plan(int M, int N, int K, int BK, int split_k) {
    GemmPlan p;
    p.big = M >= 96 && N >= 96;
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
                       void* C, int ldc, int split, void* ws, size_t ws_bytes, void* stream) {
    using In = typename Tr<MD>::In;
    using Acc = typename Tr<MD>::Acc;
    RVGP_CHECK_ARG(M >= 0 && N >= 0 && K >= 0, "rvgp_gemm: negative size");
    RVGP_CHECK_ARG(ldc >= N, "rvgp_gemm: ldc < N");
    RVGP_CHECK_ARG(lda >= (transA ? M : K) && ldb >= (transB ? K : N), "rvgp_gemm: bad lda/ldb");
    if (M == 0 || N == 0) return RVGP_OK;
    hipStream_t s = as_stream(stream);
    const GemmPlan pl = gemm_plan(M, N, K, Tr<MD>::BK, split);
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
                           ldc, part);
    } else {
        dim3 grid(cdiv(M, BM), cdiv(N, BN), pl.split);
        hipLaunchKernelGGL(gemm_kernel<MD>, grid, dim3(GT), 0, s, transA, transB, M, N, K, pl.kchunk,
                           (Acc)alpha, static_cast<const In*>(A), lda, static_cast<const In*>(B), ldb,
                           kscale, (Acc)beta, static_cast<Acc*>(C), ldc, part);
    }
    RVGP_LAUNCH_CHECK();
    if (pl.split > 1) {
        hipLaunchKernelGGL(splitk_reduce<Acc>, dim3(cdiv((size_t)M * N, 256)), dim3(256), 0, s, M, N,
                           pl.split, (Acc)alpha, part, (Acc)beta, static_cast<Acc*>(C), ldc);
        RVGP_LAUNCH_CHECK();
    }
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
    const GemmPlan p = gemm_plan(M, N, K, BK, split_k);
    return p.split > 1 ? acc * (size_t)p.split * M * N : 0;
}
