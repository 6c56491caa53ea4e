"""Host wrappers of the matrix-core GEMM, Cholesky and triangular solves."""
import torch

from . import _lib, instrument

GEMM_F64, GEMM_F32, GEMM_F32_ACC64, GEMM_BF16 = 0, 1, 2, 3


class CholeskyError(RuntimeError):
    """Raised when a matrix handed to the Cholesky factorisation is not
    positive definite (tf.linalg.cholesky's failure inside GPflow)."""


def _ld(t):
    """Row stride of a row-major 2-D operand.  A single column may carry any
    column stride (torch treats it as contiguous); a row stride below the row
    width (an expanded / broadcast view, stride 0) would make the kernels read
    rows that are not stored, and is refused -- call .contiguous() first."""
    assert t.dim() == 2 and (t.stride(1) == 1 or t.shape[1] == 1), \
        "row-major 2-D tensor with unit column stride"
    if t.shape[0] > 1 and t.stride(0) < t.shape[1]:
        raise ValueError(f"operand rows overlap (shape {tuple(t.shape)}, row stride "
                         f"{t.stride(0)}): pass a .contiguous() tensor")
    return max(t.stride(0), t.shape[1])


def gemm(A, B, C=None, transA=False, transB=False, alpha=1.0, beta=0.0, kscale=None,
         mode=None, split_k=None, out_dtype=None):
    """C = alpha op(A) diag(kscale) op(B) + beta C on the matrix cores."""
    M = A.shape[1] if transA else A.shape[0]
    K = A.shape[0] if transA else A.shape[1]
    Kb = B.shape[1] if transB else B.shape[0]
    N = B.shape[0] if transB else B.shape[1]
    assert K == Kb, (A.shape, B.shape, transA, transB)
    if mode is None:
        if A.dtype == torch.float64:
            mode = GEMM_F64
        elif out_dtype == torch.float64 or (C is not None and C.dtype == torch.float64):
            mode = GEMM_F32_ACC64
        else:
            mode = GEMM_F32
    want = {GEMM_F64: (torch.float64, torch.float64), GEMM_F32: (torch.float32, torch.float32),
            GEMM_F32_ACC64: (torch.float32, torch.float64),
            GEMM_BF16: (torch.float32, torch.float32)}[mode]
    assert A.dtype == want[0] and B.dtype == want[0], (A.dtype, B.dtype, mode)
    if C is None:
        C = torch.empty((M, N), dtype=want[1], device=A.device)
        beta = 0.0
    assert C.dtype == want[1] and C.shape == (M, N)
    if kscale is not None:
        kscale = kscale.to(torch.float64).contiguous()
    if split_k is None:
        split_k = 0  # the library picks the K split from the shape
    ws = None
    ws_bytes = _lib.lib().rvgp_gemm_ws_bytes(mode, M, N, K, int(split_k))
    if ws_bytes > 0:
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=A.device)
    ev = instrument.begin()
    _lib.call("rvgp_gemm", mode, int(transA), int(transB), M, N, K, float(alpha), _lib.ptr(A),
              _ld(A), _lib.ptr(B), _ld(B), _lib.ptr(kscale), float(beta), _lib.ptr(C), _ld(C),
              int(split_k), _lib.ptr(ws), ws_bytes, _lib.stream())
    if ev is not None:
        instrument.end(ev, "gemm", MODE_NAMES[mode],
                       (M * K + K * N) * A.element_size() + M * N * C.element_size(),
                       2.0 * M * N * K)
    return C


MODE_NAMES = {GEMM_F64: "f64", GEMM_F32: "f32", GEMM_F32_ACC64: "f32acc64", GEMM_BF16: "bf16"}


def gemm_sym(A, B, fp32_mfma=True):
    """C = A^T B (f64) for a product known to be symmetric: upper tiles only,
    lower triangle mirrored (rvgp_gemm_sym).  A, B: (K, M), f64 or f32.  f32
    blocks run on the FP32 matrix cores in K-chunks of <= 4096 rows summed in
    FP64 (fp32_mfma=True), or widened to FP64 on the FP64 matrix cores."""
    K, M = A.shape
    assert B.shape == (K, M) and A.dtype == B.dtype
    if A.dtype == torch.float64:
        mode = GEMM_F64
    else:
        mode = GEMM_F32 if fp32_mfma else GEMM_F32_ACC64
    C = torch.empty((M, M), dtype=torch.float64, device=A.device)
    ws_bytes = _lib.lib().rvgp_gemm_ws_bytes(mode, M, M, K, 0)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=A.device) if ws_bytes > 0 else None
    ev = instrument.begin()
    _lib.call("rvgp_gemm_sym", mode, M, K, _lib.ptr(A), _ld(A), _lib.ptr(B), _ld(B), _lib.ptr(C),
              M, 0, _lib.ptr(ws), ws_bytes, _lib.stream())
    if ev is not None:
        instrument.end(ev, "gemm_sym", MODE_NAMES[mode], 2 * M * K * A.element_size() + M * M * 8,
                       2.0 * M * M * K)
    return C


def potrf(A, check=True):
    """In-place lower Cholesky of a square row-major tensor (f64 or f32)."""
    n = A.shape[0]
    st = torch.zeros(1, dtype=torch.int32, device=A.device)
    _lib.call("rvgp_potrf", int(A.dtype == torch.float64), n, _lib.ptr(A), _ld(A), _lib.ptr(st),
              _lib.stream())
    if check and int(st.item()) != 0:
        raise CholeskyError("Cholesky decomposition was not successful. The input might not be "
                            "valid (matrix not positive definite).")
    return A, st


def trsm(L, B, trans=False):
    """In place: B <- L^{-1} B (trans=False) or L^{-T} B (trans=True)."""
    B2 = B.reshape(-1, 1) if B.dim() == 1 else B
    _lib.call("rvgp_trsm", int(L.dtype == torch.float64), int(trans), L.shape[0], B2.shape[1],
              _lib.ptr(L), _ld(L), _lib.ptr(B2), _ld(B2), _lib.stream())
    return B


def colsumsq(Z):
    """out[j] = sum_i Z[i, j]^2 in FP64 (deterministic order), Z f32 or f64."""
    k = Z.shape[1]
    out = torch.empty(k, dtype=torch.float64, device=Z.device)
    ws_bytes = _lib.lib().rvgp_colsumsq_ws_bytes(Z.shape[0], k)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=Z.device)
    _lib.call("rvgp_colsumsq", int(Z.dtype == torch.float64), Z.shape[0], k, _lib.ptr(Z), _ld(Z),
              _lib.ptr(out), _lib.ptr(ws), ws_bytes, _lib.stream())
    return out


def ritz_residuals(X, AX, theta):
    """|AX[:, j] - theta[j] X[:, j]| / |X[:, j]| per column (f64), X and AX
    (M, k) f32 or f64 with unit column stride, theta (k,) f64."""
    M, k = X.shape
    if AX.shape != X.shape or AX.dtype != X.dtype:
        raise ValueError("ritz_residuals: X and AX differ in shape or dtype")
    th = theta.to(device=X.device, dtype=torch.float64).contiguous()
    out = torch.empty(k, dtype=torch.float64, device=X.device)
    ws_bytes = _lib.lib().rvgp_ritz_residuals_ws_bytes(M, k)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=X.device)
    _lib.call("rvgp_ritz_residuals", int(X.dtype == torch.float64), M, k, _lib.ptr(X), _ld(X),
              _lib.ptr(AX), _ld(AX), _lib.ptr(th), _lib.ptr(out), _lib.ptr(ws), ws_bytes,
              _lib.stream())
    return out


def evecs_finish(X, b, scale=1.0, perm=None):
    """Unit-norm columns of X (nb*b, k) times `scale`, block rows mapped back
    through perm (perm[new] = old block row; None = identity), as a new f64
    (nb*b, k) tensor."""
    rows, k = X.shape
    nb = rows // b
    if nb * b != rows:
        raise ValueError("evecs_finish: rows not a multiple of the block size")
    nrm2 = colsumsq(X)
    out = torch.empty((rows, k), dtype=torch.float64, device=X.device)
    pp = None
    if perm is not None:
        pp = perm.to(device=X.device, dtype=torch.int64).contiguous()
        if pp.numel() != nb:
            raise ValueError("evecs_finish: perm length differs from the block rows")
    _lib.call("rvgp_evecs_finish", int(X.dtype == torch.float64), nb, b, k, _lib.ptr(X), _ld(X),
              _lib.ptr(nrm2), _lib.ptr(pp), float(scale), _lib.ptr(out),
              k, _lib.stream())
    return out


def dot(x, y):
    """sum_i x_i y_i of two contiguous FP64 tensors, as a (1,) device
    tensor (fixed summation order)."""
    x = x.contiguous()
    y = y.contiguous()
    n = x.numel()
    if y.numel() != n:
        raise ValueError("dot: sizes differ")
    out = torch.empty(1, dtype=torch.float64, device=x.device)
    ws_bytes = _lib.lib().rvgp_dot_ws_bytes(n)
    ws = torch.empty(max(ws_bytes, 8), dtype=torch.uint8, device=x.device)
    _lib.call("rvgp_dot", n, _lib.ptr(x), _lib.ptr(y), _lib.ptr(out), _lib.ptr(ws), ws_bytes,
              _lib.stream())
    return out


def chol_logdet_half(L):
    """sum_i log L_ii (f64 device scalar)."""
    out = torch.empty(1, dtype=torch.float64, device=L.device)
    _lib.call("rvgp_chol_logdet", int(L.dtype == torch.float64), L.shape[0], _lib.ptr(L), _ld(L),
              _lib.ptr(out), _lib.stream())
    return out


def syev(A):
    """Eigenvalues (ascending) and eigenvectors (columns) of the symmetric
    FP64 matrix A on the device (rvgp_syev), as torch.linalg.eigh returns
    them.  A is not modified."""
    n = A.shape[0]
    assert A.shape == (n, n) and A.dtype == torch.float64
    work = A.contiguous()
    w = torch.empty(n, dtype=torch.float64, device=A.device)
    V = torch.empty((n, n), dtype=torch.float64, device=A.device)
    ws_bytes = _lib.lib().rvgp_syev_ws_bytes(n)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=A.device)
    ev = instrument.begin()
    _lib.call("rvgp_syev", n, _lib.ptr(work), n, _lib.ptr(w), _lib.ptr(V), n, _lib.ptr(ws),
              ws_bytes, _lib.stream())
    if ev is not None:
        instrument.end(ev, "syev", "f64", 8 * 3 * n * n, 4.0 * n ** 3 / 3.0 + 2.0 * n ** 3)
    return w, V
