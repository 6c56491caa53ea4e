"""Block-CSR operators on the device and their fused SpMM.

A :class:`BSR` holds the connection Laplacian (b = dim_man) or the scalar
Laplacian (b = 1) in HBM: block-row pointer and block-column indices (int32,
the symmetrised graph's CSR), and the b x b blocks row-major in f64 and f32.
"""
import numpy as np
import torch

from . import _lib, instrument


class BSR:
    def __init__(self, rowptr, cols, vals64, vals32, b):
        self.rowptr, self.cols, self.b = rowptr, cols, int(b)
        self.vals64, self.vals32 = vals64, vals32
        self.nb = rowptr.numel() - 1
        self.nnzb = cols.numel()
        self.shape = (self.nb * self.b, self.nb * self.b)

    def vals(self, dtype):
        if dtype == torch.float64:
            return self.vals64
        if self.vals32 is None:
            self.vals32 = self.vals64.to(torch.float32)
        return self.vals32

    def gershgorin_bound(self):
        """max_i sum_j |A_ij| >= lambda_max (host scalar): the row sums are
        |A| 1, one narrow block SpMM with the absolute values."""
        absA = BSR(self.rowptr, self.cols, self.vals64.abs(), None, self.b)
        ones = torch.ones(self.shape[0], dtype=torch.float64, device=self.vals64.device)
        return float(spmm(absA, ones).max().item())

    def spectral_upper_bound(self, steps=40, seed=20111):
        """Upper bound of lambda_max for the Chebyshev filters and the heat
        kernel series, much tighter than Gershgorin (which overestimates a
        kNN-graph (connection) Laplacian's lambda_max ~2x, and a Chebyshev
        filter's convergence rate per degree scales with 1/sqrt(bound)).

        k-step Lanczos in FP64 from a seeded random vector (no
        reorthogonalisation: only the largest Ritz value is used); the bound is
        theta_max + |beta_k z_k| -- the largest Ritz value plus the residual
        norm of its Ritz pair (Zhou & Li, "Bounding the spectrum of large
        Hermitian matrices", Linear Algebra Appl. 435, 2011) -- with a 2 %
        margin, never above the Gershgorin bound.  On kNN (connection)
        Laplacians 40 steps put theta_max + |beta_k z_k| within 1e-4 of
        lambda_max.  One host synchronisation (the k x k tridiagonal matrix);
        cached on the operator."""
        if getattr(self, "_ub", None) is not None:
            return self._ub
        dev = self.rowptr.device
        N = self.shape[0]
        steps = max(2, min(int(steps), N))
        g = torch.Generator(device=dev)
        g.manual_seed(int(seed))
        from . import dense
        v = torch.randn(N, generator=g, dtype=torch.float64, device=dev)
        v = v / torch.sqrt(dense.dot(v, v))
        v_prev = torch.zeros_like(v)
        beta = torch.zeros(1, dtype=torch.float64, device=dev)
        alphas, betas = [], []
        for _ in range(steps):
            w = spmm(self, v).reshape(-1)
            alpha = dense.dot(w, v)
            w = w - alpha * v - beta * v_prev
            beta = torch.sqrt(dense.dot(w, w))
            alphas.append(alpha)
            betas.append(beta)
            v_prev, v = v, w / torch.clamp(beta, min=1e-300)
        a = torch.cat(alphas).cpu().numpy()
        b = torch.cat(betas).cpu().numpy()
        T = np.diag(a) + np.diag(b[:-1], 1) + np.diag(b[:-1], -1)
        th, Zt = np.linalg.eigh(T)
        bound = float(th[-1]) + abs(float(b[-1]) * float(Zt[-1, -1]))
        self._ub = min(self.gershgorin_bound(), bound * 1.02)
        return self._ub

    def to_scipy(self):
        import scipy.sparse as sp
        rp, c = self.rowptr.cpu().numpy(), self.cols.cpu().numpy()
        v = self.vals64.cpu().numpy().reshape(-1, self.b, self.b)
        return sp.bsr_matrix((v, c, rp), shape=self.shape).tocsr()

    @staticmethod
    def from_scipy(A, b=1, device=None):
        import scipy.sparse as sp
        dev = device or _lib.device()
        B = sp.bsr_matrix(A, blocksize=(b, b))
        B.sort_indices()
        rp = torch.as_tensor(B.indptr.astype(np.int32), device=dev)
        c = torch.as_tensor(B.indices.astype(np.int32), device=dev)
        v = torch.as_tensor(np.ascontiguousarray(B.data, dtype=np.float64), device=dev).reshape(-1)
        return BSR(rp, c, v, None, b)


def spmm(A, X, Y=None, alpha=1.0, beta=0.0, Z=None, gamma=0.0, acc=None, delta=0.0):
    """Y = alpha A X + beta X + gamma Z ; acc += delta Y.  X is (nb*b, r) or
    (nb*b,); dtype float64 or float32 selects the kernel and A's values;
    float16 selects the half-storage form (f32 values and arithmetic)."""
    vec = X.dim() == 1
    X2 = X.reshape(-1, 1) if vec else X
    assert X2.shape[0] == A.shape[0], (X2.shape, A.shape)
    r = X2.shape[1]
    if X2.stride(1) != 1:
        X2 = X2.contiguous()
    if Y is None:
        Y = torch.empty_like(X2)
    Y2 = Y.reshape(-1, 1) if Y.dim() == 1 else Y
    Z2 = None if Z is None else (Z.reshape(-1, 1) if Z.dim() == 1 else Z)
    A2 = None if acc is None else (acc.reshape(-1, 1) if acc.dim() == 1 else acc)
    if X2.dtype == torch.float16:
        return _spmm_f16(A, X2, Y, Y2, alpha, beta, Z2, gamma, acc)
    f64 = X2.dtype == torch.float64
    vals = A.vals(X2.dtype)
    ev = instrument.begin()
    _lib.call("rvgp_bsr_spmm_f64" if f64 else "rvgp_bsr_spmm_f32", A.nb, A.b,
              _lib.ptr(A.rowptr), _lib.ptr(A.cols), _lib.ptr(vals), _lib.ptr(X2), X2.stride(0), r,
              float(alpha), float(beta), _lib.ptr(Z2), Z2.stride(0) if Z2 is not None else r,
              float(gamma), _lib.ptr(Y2), Y2.stride(0), _lib.ptr(A2),
              A2.stride(0) if A2 is not None else r, float(delta), _lib.stream())
    if ev is not None:
        s = X2.element_size()
        nbytes = spmm_bytes(A, r, s, Z is not None and gamma != 0.0, acc is not None)
        if r <= 4:
            kind = "spmm_narrow"
        elif not _vec_ok(r, X2, Y2, Z2, A2):
            kind = "spmm_wide"
        else:  # the buffer form addresses X with 32-bit offsets
            kind = "spmm_buf" if A.nb * A.b * X2.stride(0) * s < 2 ** 31 else "spmm_vec"
        instrument.end(ev, kind, (str(X2.dtype).replace("torch.", ""), A.b), nbytes,
                       2.0 * A.nnzb * A.b * A.b * r)
    return Y


def _spmm_f16(A, X2, Y, Y2, alpha, beta, Z2, gamma, acc):
    """Half-storage SpMM (binary16 X, Z, Y; f32 values and arithmetic):
    r and the row strides multiples of 8, no acc term."""
    if acc is not None:
        raise ValueError("spmm: the binary16 form has no acc term")
    ev = instrument.begin()
    r = X2.shape[1]
    _lib.call("rvgp_bsr_spmm_f16", A.nb, A.b, _lib.ptr(A.rowptr), _lib.ptr(A.cols),
              _lib.ptr(A.vals(torch.float32)), _lib.ptr(X2), X2.stride(0), r, float(alpha),
              float(beta), _lib.ptr(Z2), Z2.stride(0) if Z2 is not None else r, float(gamma),
              _lib.ptr(Y2), Y2.stride(0), _lib.stream())
    if ev is not None:
        nbytes = spmm_bytes(A, r, 2, Z2 is not None and gamma != 0.0) \
            + A.nnzb * A.b * A.b * 2  # values are f32: 4 bytes, not 2
        instrument.end(ev, "spmm_buf_h", ("float16", A.b), nbytes, 2.0 * A.nnzb * A.b * A.b * r)
    return Y


def graph_laplacian(G):
    """L = D - A of the symmetrised graph as a b = 1 BSR on the device
    (geometry.py:55-63): degree (row length minus the self loop) on the
    diagonal, -1 off it, in the graph's CSR order -- the same values the
    connections kernel writes."""
    rows = torch.repeat_interleave(torch.arange(G.n, device=G.rowptr.device),
                                   torch.diff(G.rowptr.long()))
    deg = (torch.diff(G.rowptr) - 1).to(torch.float64)
    l64 = torch.where(G.cols.long() == rows, deg[rows], torch.full_like(deg[rows], -1.0))
    return BSR(G.rowptr, G.cols, l64, l64.to(torch.float32), 1)


def _vec_ok(r, *ts):
    """Mirror of the kernel's test for the 16-byte vector form: r, every row
    stride and every base address multiples of 16 bytes."""
    for t in ts:
        if t is None:
            continue
        vw = 16 // t.element_size()
        if r % vw or t.stride(0) % vw or t.data_ptr() % 16:
            return False
    return True


def spmm_bytes(A, r, s, with_z=False, with_acc=False):
    """Algorithmic HBM bytes of one fused SpMM launch: block values and column
    indices once, row pointer once, X read once, Y written once, plus Z read
    and acc read + written when the epilogue uses them."""
    dense_rows = A.nb * A.b * r * s
    b = A.nnzb * (A.b * A.b * s + 4) + 4 * (A.nb + 1) + 2 * dense_rows
    if with_z:
        b += dense_rows
    if with_acc:
        b += 2 * dense_rows
    return b
