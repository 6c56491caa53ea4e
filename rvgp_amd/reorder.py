"""Locality renumbering of graph nodes for the block SpMM.

Points arrive in arbitrary order (the synthetic generators draw them
uniformly at random), so the block columns of a block row are scattered over
the whole operator and every gathered X row misses the caches.  Before the
eigensolve the nodes are renumbered along a Morton (Z-order) curve of their
leading principal coordinates: neighbours on the manifold get nearby
indices, consecutive block rows share most of their neighbours, and the
X rows a workgroup gathers stay resident in L2.

The permuted operator P A P^T has the same spectrum; eigenvector block rows
are mapped back through the permutation when the eigensolver writes its
output (rvgp_evecs_finish).  Node identity is never
renumbered anywhere else (the heap neighbourhoods, whose tie-breaking
depends on node order, run on the original numbering).
"""
import torch


def morton_perm(X, max_dims=6):
    """Permutation (new position -> old node) sorting nodes by the Morton
    code of their top-P principal coordinates (P = min(D, max_dims))."""
    from . import dense
    Xc = (X - X.mean(0, keepdim=True)).contiguous()
    C = dense.gemm_sym(Xc, Xc)
    w, V = dense.syev(C)
    P = min(X.shape[1], max_dims)
    Y = dense.gemm(Xc, V[:, -P:].contiguous())
    bits = min(20, 62 // P)
    lo = Y.min(0).values
    span = (Y.max(0).values - lo).clamp_min(1e-300)
    q = ((Y - lo) / span * ((1 << bits) - 1)).round().to(torch.int64)
    key = torch.zeros(X.shape[0], dtype=torch.int64, device=X.device)
    for b in range(bits - 1, -1, -1):
        for d in range(P):
            key = (key << 1) | ((q[:, d] >> b) & 1)
    return torch.sort(key, stable=True).indices


def bsr_permutation(rowptr, cols, perm):
    """The structure of P A P^T for a block-CSR pattern (rowptr, cols) and
    perm[new] = old block row: (new rowptr int32, new cols int32, order) with
    order[e] = the stored block that lands at position e.  Operators on one
    graph (the scalar and the connection Laplacian) share it."""
    dev = rowptr.device
    nb = rowptr.numel() - 1
    inv = torch.empty_like(perm)
    inv[perm] = torch.arange(nb, device=dev, dtype=perm.dtype)
    rp = rowptr.long()
    lens = rp[1:] - rp[:-1]
    rows = torch.repeat_interleave(torch.arange(nb, device=dev), lens)
    new_r = inv[rows]
    new_c = inv[cols.long()]
    order = torch.sort(new_r * nb + new_c, stable=True).indices
    new_rp = torch.zeros(nb + 1, dtype=torch.int64, device=dev)
    new_rp[1:] = torch.cumsum(lens[perm], 0)
    return new_rp.to(torch.int32), new_c[order].to(torch.int32), order


def permute_bsr(A, perm, plan=None):
    """B = P A P^T for a BSR operator; perm[new] = old block row.  plan: the
    result of bsr_permutation for A's pattern, when already computed."""
    from .sparse import BSR
    rp, cols, order = plan if plan is not None else bsr_permutation(A.rowptr, A.cols, perm)
    bb = A.b * A.b
    v64 = A.vals64.reshape(-1, bb)[order].reshape(-1).contiguous()
    v32 = None if A.vals32 is None else A.vals32.reshape(-1, bb)[order].reshape(-1).contiguous()
    return BSR(rp, cols, v64, v32, A.b)
