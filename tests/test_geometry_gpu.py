"""GPU parity of the geometry kernels (a1-a7) against the oracle.

Bit-exact for indices (FPS order, kNN sets, CSR, heap neighbourhoods) and
FPS lambdas; FP64 tolerances for SVD / polar outputs (gauge-invariant forms
where singular-vector signs are free).
"""
import glob
import os

import numpy as np
import pytest

from oracle import cpu
from tests.util import sphere, torus, flat_torus3, golden_cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def geo():
    from rvgp_amd import geometry
    return geometry


# ---------------------------------------------------------------- a1 FPS
@pytest.mark.parametrize("n,D,spacing,N,start", [
    (2000, 3, 0.1, None, 0),
    (2000, 3, 0.05, None, 17),
    (1500, 8, 0.2, None, 3),
    (2000, 3, 0.1, 500, 0),
    (300, 32, 0.1, 300, 5),
])
def test_fps_bit_exact(geo, n, D, spacing, N, start):
    x = np.random.default_rng(n + D).normal(size=(n, D))
    p_gpu, l_gpu = geo.furthest_point_sampling(x, N=N, spacing=spacing, start_idx=start)
    p_ref, l_ref = cpu.fps(x, N=N, spacing=spacing, start_idx=start)
    assert p_gpu.dtype == np.int32
    np.testing.assert_array_equal(p_gpu, p_ref)
    np.testing.assert_array_equal(l_gpu, l_ref)
    # against the reference's own formula (sklearn pairwise_distances)
    p_sk, l_sk = cpu.fps(x, N=N, spacing=spacing, start_idx=start, metric="sklearn")
    np.testing.assert_array_equal(p_gpu, p_sk)
    np.testing.assert_allclose(l_gpu, l_sk, rtol=1e-12, atol=1e-12)


def test_fps_torus_c2_size(geo):
    """Config-2 shape: 20,000 torus points, spacing 0.015."""
    x = torus(20000, 0)
    p_gpu, l_gpu = geo.furthest_point_sampling(x, spacing=0.015)
    p_ref, l_ref = cpu.fps(x, spacing=0.015)
    np.testing.assert_array_equal(p_gpu, p_ref)
    np.testing.assert_array_equal(l_gpu, l_ref)
    assert 3000 < len(p_gpu) < 5000
    assert np.all(np.diff(l_gpu[1:]) <= 0)  # lambdas non-increasing
    assert len(np.unique(p_gpu)) == len(p_gpu)


def test_fps_edge_cases(geo):
    x = np.random.default_rng(0).normal(size=(50, 3))
    p, l = geo.furthest_point_sampling(x, spacing=0.0)
    np.testing.assert_array_equal(p, np.arange(50))
    assert l is None
    # N > n: once every point is taken all distances are 0 and argmax -> 0
    p, l = geo.furthest_point_sampling(x, N=60)
    pr, lr = cpu.fps(x, N=60)
    np.testing.assert_array_equal(p, pr)
    np.testing.assert_array_equal(l, lr)
    # exact ties: duplicated points, lowest index wins
    xd = np.repeat(np.random.default_rng(1).normal(size=(20, 2)), 3, axis=0)
    p, l = geo.furthest_point_sampling(xd, N=40)
    pr, lr = cpu.fps(xd, N=40)
    np.testing.assert_array_equal(p, pr)
    # single point
    p, l = geo.furthest_point_sampling(x[:1])
    np.testing.assert_array_equal(p, [0])


# ---------------------------------------------------------------- a2 kNN + CSR
# n > 8192 takes the bounding pass over the first 8192 points
@pytest.mark.parametrize("n,D,k", [(2000, 3, 10), (3000, 32, 20), (1000, 8, 15), (500, 3, 40),
                                   (9000, 3, 10), (10000, 5, 30)])
def test_knn_bit_exact(geo, n, D, k):
    import torch
    x = np.random.default_rng(7 * n + D).normal(size=(n, D))
    xd = torch.as_tensor(x, device="cuda")
    nbr = geo.knn_device(xd, k).cpu().numpy()
    np.testing.assert_array_equal(nbr, cpu.knn(x, k))
    rp, c = geo.symmetrize_device(geo.knn_device(xd, k))
    rpo, co = cpu.sym_csr(cpu.knn(x, k))
    np.testing.assert_array_equal(rp.cpu().numpy(), rpo)
    np.testing.assert_array_equal(c.cpu().numpy(), co)


def test_knn_ties_lattice(geo):
    """Integer lattice: many exactly equal distances; lower index wins."""
    import torch
    g = np.stack(np.meshgrid(np.arange(12), np.arange(12), np.arange(3), indexing="ij"), -1)
    x = g.reshape(-1, 3).astype(np.float64)
    nbr = geo.knn_device(torch.as_tensor(x, device="cuda"), 6).cpu().numpy()
    np.testing.assert_array_equal(nbr, cpu.knn(x, 6))


@pytest.mark.parametrize("order", ["morton", "random", "reversed"])
def test_knn_query_order_does_not_change_result(geo, order):
    """The query order only packs queries into waves: any permutation gives
    the oracle's neighbour lists, on both sides of the bounding size."""
    import torch
    from rvgp_amd.reorder import morton_perm
    for n, D, k in [(3000, 6, 12), (9500, 32, 20)]:
        x = np.random.default_rng(n + D).normal(size=(n, D))
        xd = torch.as_tensor(x, device="cuda")
        if order == "morton":
            perm = morton_perm(xd)
        elif order == "random":
            perm = torch.as_tensor(np.random.default_rng(1).permutation(n), device="cuda")
        else:
            perm = torch.arange(n - 1, -1, -1, device="cuda")
        nbr = geo.knn_device(xd, k, query_order=perm).cpu().numpy()
        np.testing.assert_array_equal(nbr, cpu.knn(x, k))


def test_knn_duplicates_above_bounding_size(geo):
    """9000 points on a 20^3 integer lattice: every point has exact
    duplicates, so bounds of 0 and long runs of equal distances cross the
    bounding pass; lower index wins every tie."""
    import torch
    x = np.random.default_rng(3).integers(0, 20, size=(9000, 3)).astype(np.float64)
    for k in (4, 12):
        nbr = geo.knn_device(torch.as_tensor(x, device="cuda"), k).cpu().numpy()
        np.testing.assert_array_equal(nbr, cpu.knn(x, k))


def test_knn_rejects_non_finite(geo):
    import torch
    x = np.random.default_rng(0).normal(size=(50, 3))
    x[7, 1] = np.nan
    with pytest.raises(ValueError):
        geo.knn_device(torch.as_tensor(x, device="cuda"), 5)
    x[7, 1] = 1e200
    with pytest.raises(ValueError):
        geo.knn_device(torch.as_tensor(x, device="cuda"), 5)
    x[7, 1] = 0.0
    with pytest.raises(ValueError):  # an order must cover every point
        geo.knn_device(torch.as_tensor(x, device="cuda"), 5, query_order=np.arange(49))
    with pytest.raises(ValueError):  # ... exactly once
        geo.knn_device(torch.as_tensor(x, device="cuda"), 5, query_order=np.r_[0, np.arange(49)])


def test_symmetrize_out_of_range_ids_stay_in_bounds(geo):
    """Neighbour ids outside [0, n) (an nbr array not produced by a complete
    kNN) become self loops; the CSR equals the one without them."""
    import torch
    x = np.random.default_rng(2).normal(size=(300, 3))
    nbr = cpu.knn(x, 6)
    bad = nbr.copy()
    bad[5, 2] = -7
    bad[17, 0] = 10**6
    rp, c = geo.symmetrize_device(torch.as_tensor(bad, device="cuda"))
    keep = nbr.copy()
    keep[5, 2] = 5
    keep[17, 0] = 17
    rpo, co = cpu.sym_csr(keep)
    np.testing.assert_array_equal(rp.cpu().numpy(), rpo)
    np.testing.assert_array_equal(c.cpu().numpy(), co)


def test_manifold_graph_matches_networkx(geo):
    x = sphere(1200, 3)
    G = geo.manifold_graph(x, n_neighbors=10)
    rp, c, _ = cpu.graph_csr_networkx(x, 10)
    np.testing.assert_array_equal(G.rowptr.cpu().numpy(), rp)
    np.testing.assert_array_equal(G.cols.cpu().numpy(), c)
    Gx = G.to_networkx()
    assert Gx.number_of_nodes() == 1200 and all(Gx.has_edge(i, i) for i in range(1200))


def test_knn_rejects_bad_k(geo):
    with pytest.raises(ValueError):
        geo.manifold_graph(np.zeros((5, 3)), n_neighbors=5)


# ---------------------------------------------------------------- a3 heap nbh + gauges
@pytest.mark.parametrize("case", golden_cases())
def test_heap_neighbourhoods_and_gauges_vs_reference(case):
    """Golden vectors written by the reference's own compiled ptu_dijkstra."""
    import torch
    from rvgp_amd import ptu_dijkstra
    from rvgp_amd.graph import SymGraph
    z = np.load(case)
    X, k, K = z["X"], int(z["k"]), int(z["K"])
    rp, c = cpu.sym_csr(cpu.knn(X, k))
    g = SymGraph(torch.as_tensor(rp, device="cuda"), torch.as_tensor(c, device="cuda"), len(X))
    nbh = ptu_dijkstra.geo_nbh_device(g, K).cpu().numpy()
    nbh_ref, _ = cpu.geo_nbh(X, rp, c, K, centred=False)
    np.testing.assert_array_equal(nbh, nbh_ref)
    D = X.shape[1]
    T, S = ptu_dijkstra.tangent_frames_device(torch.as_tensor(X, device="cuda"), g, K, D, D, D)
    T, S = T.cpu().numpy(), S.cpu().numpy()
    np.testing.assert_allclose(S, z["Sigma"], rtol=1e-9, atol=1e-12)
    d = int(z["dim"])
    P = np.einsum("nij,nkj->nik", T[:, :, :d], T[:, :, :d])
    Pr = np.einsum("nij,nkj->nik", z["tangents"][:, :, :d], z["tangents"][:, :, :d])
    assert np.abs(P - Pr).max() < 1e-9
    # orthonormal frames
    I = np.einsum("nji,njk->nik", T, T)
    assert np.abs(I - np.eye(D)).max() < 1e-10


def test_heap_neighbourhoods_pool_overflow_second_pass():
    """A hub joined to 700 nodes: every search that pops it touches more
    nodes than the first pass's 512-node pool, so those sources are redone
    by the worst-case-pool pass; all neighbourhoods equal the oracle's."""
    import scipy.sparse as sp
    import torch
    from rvgp_amd import ptu_dijkstra
    from rvgp_amd.graph import SymGraph
    n, K = 2000, 30
    X = np.random.default_rng(11).normal(size=(n, 3))
    rp0, c0 = cpu.sym_csr(cpu.knn(X, 6))
    A = sp.csr_matrix((np.ones(len(c0)), c0, rp0), shape=(n, n))
    hub = sp.coo_matrix((np.ones(700), (np.zeros(700, int), np.arange(1, 701))), shape=(n, n))
    A = ((A + hub + hub.T) > 0).astype(np.float64).tocsr()
    A.sort_indices()
    rp, c = A.indptr.astype(np.int32), A.indices.astype(np.int32)
    g = SymGraph(torch.as_tensor(rp, device="cuda"), torch.as_tensor(c, device="cuda"), n)
    nbh = ptu_dijkstra.geo_nbh_device(g, K).cpu().numpy()
    nbh_ref, _ = cpu.geo_nbh(X, rp, c, K, centred=False)
    np.testing.assert_array_equal(nbh, nbh_ref)
    assert (nbh == 0).any(axis=1).sum() > 700  # the hub is in many neighbourhoods


def test_tangent_frames_dropin_api(golden_dir):
    from rvgp_amd import ptu_dijkstra, geometry
    z = np.load(os.path.join(golden_dir, "heap_sphere2000_k10.npz"))
    X = z["X"]
    G = geometry.manifold_graph(X, n_neighbors=10)
    T, S = ptu_dijkstra.tangent_frames(X, G, X.shape[1], 10 * 1.5)
    assert T.shape == (2000, 3, 3) and S.shape == (2000, 3)
    np.testing.assert_allclose(S, z["Sigma"], rtol=1e-9, atol=1e-12)
    with pytest.raises(ValueError):
        ptu_dijkstra.tangent_frames(X, G, 3, 2000)
    with pytest.raises(ValueError):
        ptu_dijkstra.tangent_frames(X, G, 3, 2.0)
    with pytest.raises(ValueError):
        ptu_dijkstra.tangent_frames(X, G, 4, 15)


def test_gauges_rank_deficient_raises():
    """Planar points: the third singular value is 0 -> RuntimeError (pyx:426-428)."""
    from rvgp_amd import ptu_dijkstra, geometry
    rng = np.random.default_rng(0)
    X = np.c_[rng.normal(size=(300, 2)), np.zeros(300)]
    G = geometry.manifold_graph(X, n_neighbors=10)
    with pytest.raises(RuntimeError):
        ptu_dijkstra.tangent_frames(X, G, 3, 15)


@pytest.mark.parametrize("k,K", [(10, 15), (20, 30)])
def test_gauges_wide_ambient_gram_path(k, K):
    """D > K+1 (the 3-torus in R^32 shape; k = 20 gives the headline's
    K = 30): one-sided Jacobi on the neighbourhood columns.  Singular values
    and tangent projectors vs the oracle's LAPACK SVD, and the frames
    orthonormal to rounding (the weight-space statistics use T^T T = I)."""
    import torch
    from rvgp_amd import ptu_dijkstra
    from rvgp_amd.graph import SymGraph
    X = flat_torus3(600, 5, D=32)
    rp, c = cpu.sym_csr(cpu.knn(X, k))
    g = SymGraph(torch.as_tensor(rp, device="cuda"), torch.as_tensor(c, device="cuda"), len(X))
    T, S = ptu_dijkstra.tangent_frames_device(torch.as_tensor(X, device="cuda"), g, K, K + 1, K + 1, K)
    T, S = T.cpu().numpy(), S.cpu().numpy()
    Tr, Sr = cpu.tangent_frames(X, rp, c, 32, K, allow_rank_deficient_k=True)
    np.testing.assert_allclose(S[:, :K], Sr[:, :K], rtol=1e-8, atol=1e-10)
    d = 3
    P = np.einsum("nij,nkj->nik", T[:, :, :d], T[:, :, :d])
    Pr = np.einsum("nij,nkj->nik", Tr[:, :, :d], Tr[:, :, :d])
    assert np.abs(P - Pr).max() < 1e-8
    # frames of the K singular directions with non-zero singular values
    TT = np.einsum("npi,npj->nij", T[:, :, :K], T[:, :, :K])
    assert np.abs(TT - np.eye(K)).max() < 1e-12


# ---------------------------------------------------------------- a5-a7 connections / Laplacians
@pytest.mark.parametrize("case", golden_cases())
def test_connections_vs_reference(case):
    import torch
    from rvgp_amd import ptu_dijkstra
    from rvgp_amd.graph import SymGraph
    z = np.load(case)
    X, k, d = z["X"], int(z["k"]), int(z["dim"])
    rp, c = cpu.sym_csr(cpu.knn(X, k))
    g = SymGraph(torch.as_tensor(rp, device="cuda"), torch.as_tensor(c, device="cuda"), len(X))
    # use the reference's own gauges so R is compared like for like
    T = torch.as_tensor(np.ascontiguousarray(z["tangents"][:, :, :d]), device="cuda")
    out = ptu_dijkstra.connections_device(T, g, d)
    R = out["R"].cpu().numpy()
    coo = ptu_dijkstra.blocks_to_coo(R, rp, c)
    np.testing.assert_array_equal(coo.row, z["R_row"])
    np.testing.assert_array_equal(coo.col, z["R_col"])
    np.testing.assert_allclose(coo.data, z["R_data"], rtol=0, atol=1e-10)
    # Laplacian blocks vs the restated scipy construction
    Lc = cpu.connection_laplacian(rp, c, R)
    from rvgp_amd.sparse import BSR
    B = BSR(g.rowptr, g.cols, out["lc64"].reshape(-1), out["lc32"].reshape(-1), d)
    assert abs(B.to_scipy() - Lc).max() < 1e-14
    L = cpu.laplacian(rp, c)
    assert np.abs(out["l64"].cpu().numpy() - L[np.repeat(np.arange(len(X)), np.diff(rp)), c].A1).max() == 0
    # the graph-only scalar Laplacian (built before the connections) is the same
    from rvgp_amd.sparse import graph_laplacian
    Lg = graph_laplacian(g)
    assert torch.equal(Lg.vals64, out["l64"]) and torch.equal(Lg.vals32, out["l32"])


# ---------------------------------------------------------------- SpMM
@pytest.mark.parametrize("b", [1, 2, 3])
@pytest.mark.parametrize("r", [1, 3, 7, 130, 136])
def test_bsr_spmm(b, r):
    import torch
    from rvgp_amd.sparse import BSR, spmm
    rng = np.random.default_rng(b * 100 + r)
    x = sphere(700, b + r)
    rp, c = cpu.sym_csr(cpu.knn(x, 8))
    nnz = len(c)
    vals = rng.normal(size=(nnz, b, b))
    A = BSR(torch.as_tensor(rp, device="cuda"), torch.as_tensor(c, device="cuda"),
            torch.as_tensor(vals.reshape(-1), device="cuda"), None, b)
    As = A.to_scipy()
    X = rng.normal(size=(700 * b, r))
    Z = rng.normal(size=(700 * b, r))
    ACC = rng.normal(size=(700 * b, r))
    ref = 0.7 * (As @ X) - 0.3 * X + 1.5 * Z
    for dt, tol in ((torch.float64, 1e-12), (torch.float32, 2e-5)):
        Xd = torch.as_tensor(X, device="cuda", dtype=dt)
        Zd = torch.as_tensor(Z, device="cuda", dtype=dt)
        accd = torch.as_tensor(ACC, device="cuda", dtype=dt)
        Y = spmm(A, Xd, alpha=0.7, beta=-0.3, Z=Zd, gamma=1.5, acc=accd, delta=0.25)
        scale = np.abs(ref).max()
        assert np.abs(Y.cpu().double().numpy() - ref).max() < tol * scale * 10
        assert np.abs(accd.cpu().double().numpy() - (ACC + 0.25 * ref)).max() < tol * scale * 10
        # plain product, in place over Z allowed
        Y2 = spmm(A, Xd, Y=Zd, alpha=1.0, Z=Zd, gamma=0.0)
        assert np.abs(Y2.cpu().double().numpy() - As @ X).max() < tol * np.abs(As @ X).max() * 10


@pytest.mark.parametrize("b", [1, 2, 3])
@pytest.mark.parametrize("r", [8, 136, 520])
def test_bsr_spmm_binary16(b, r):
    """rvgp_bsr_spmm_f16 (binary16 X, Z, Y; f32 block values and arithmetic)
    vs the FP64 product of the same binary16 inputs and f32 values.
    Tolerance: the output's one rounding to binary16 (2^-11 of |ref|) plus
    the f32 accumulation (1e-6 of the sum of the absolute terms).  Covers
    one and two column groups (r = 520: 65 8-column vectors), a column slice
    with the parent's row stride (the per-column-degree filter's form) and
    Y written over Z (the filter recurrence's form)."""
    import torch
    from rvgp_amd import _lib
    from rvgp_amd.sparse import BSR, spmm
    rng = np.random.default_rng(b * 1000 + r)
    n = 700
    x = sphere(n, b + r)
    rp, c = cpu.sym_csr(cpu.knn(x, 8))
    vals = rng.normal(size=(len(c), b, b))
    A = BSR(torch.as_tensor(rp, device="cuda"), torch.as_tensor(c, device="cuda"),
            torch.as_tensor(vals.reshape(-1), device="cuda"), None, b)
    A32 = BSR(A.rowptr, A.cols, A.vals(torch.float32).double(), None, b).to_scipy()
    Xh = torch.as_tensor(rng.normal(size=(n * b, r)), device="cuda").half()
    Zh = torch.as_tensor(rng.normal(size=(n * b, r)), device="cuda").half()
    X, Z = Xh.double().cpu().numpy(), Zh.double().cpu().numpy()

    def check(Y, ref, absterms):
        err = np.abs(Y.double().cpu().numpy() - ref)
        assert (err <= 2.0 ** -11 * np.abs(ref) + 1e-6 * absterms).all(), err.max()

    Y = spmm(A, Xh, alpha=0.7, beta=-0.3, Z=Zh, gamma=1.5)
    assert Y.dtype == torch.float16
    check(Y, 0.7 * (A32 @ X) - 0.3 * X + 1.5 * Z,
          0.7 * (abs(A32) @ np.abs(X)) + 0.3 * np.abs(X) + 1.5 * np.abs(Z))
    if r > 8:
        # trailing column slice (row stride r), written over Z's slice
        Zs = Zh.clone()
        Ys = spmm(A, Xh[:, 8:], Y=Zs[:, 8:], alpha=-0.4, beta=0.25, Z=Zs[:, 8:], gamma=-1.0)
        ref = -0.4 * (A32 @ X[:, 8:]) + 0.25 * X[:, 8:] - Z[:, 8:]
        check(Ys, ref, 0.4 * (abs(A32) @ np.abs(X[:, 8:])) + 0.25 * np.abs(X[:, 8:])
              + np.abs(Z[:, 8:]))
        assert torch.equal(Zs[:, :8], Zh[:, :8])
    # r not a multiple of 8 is refused (no silent widening)
    with pytest.raises(ValueError, match="multiple of 8"):
        spmm(A, Xh[:, :r - 4].contiguous())



def test_rw_connection_laplacian_uses_networkx_degrees():
    """compute_connection_laplacian(G, R, 'rw') (geometry.py:45-50): D^-1 Lc
    with networkx's degrees -- the kNN + I graph's self loop counts twice --
    for both a networkx graph and the device graph; compute_laplacian(G,
    True) is networkx's normalised Laplacian."""
    import networkx as nx
    import scipy.sparse as sp
    from sklearn.neighbors import kneighbors_graph
    from rvgp_amd import geometry
    X = sphere(150, 21)
    A = kneighbors_graph(X, 8, mode='connectivity', include_self=False) + sp.eye(150)
    Gx = nx.from_scipy_sparse_array(A)
    Gd = geometry.manifold_graph(X, n_neighbors=8)
    d = 2
    rng = np.random.default_rng(2)
    pat = sp.kron(nx.adjacency_matrix(Gx), np.ones((d, d))).tocsr()
    R = pat.multiply(rng.normal(size=pat.shape)).tocsr()
    Lref = sp.kron(sp.csr_matrix(nx.laplacian_matrix(Gx), dtype=np.float64), np.ones((d, d)))
    Lref = Lref.multiply(R)
    deg = np.array(list(dict(Gx.degree()).values()), dtype=np.float64)
    Lref = (sp.diags((1.0 / deg).repeat(d)) @ Lref).toarray()
    for G in (Gx, Gd):
        out = geometry.compute_connection_laplacian(G, R, normalization="rw").toarray()
        np.testing.assert_allclose(out, Lref, rtol=1e-15, atol=1e-15)
    Nref = nx.normalized_laplacian_matrix(Gx).toarray()
    for G in (Gx, Gd):
        np.testing.assert_allclose(geometry.compute_laplacian(G, True).toarray(), Nref, atol=1e-15)


def test_affinity_graph_matches_oracle():
    """manifold_graph(typ='affinity') (geometry.py:114-118): the Gaussian
    affinity of all pairs (sigma 0.1) as a weighted graph, against the
    oracle's sklearn pairwise_distances + nx.from_numpy_array; the heap
    neighbourhoods over those weights equal the oracle's."""
    import torch
    from rvgp_amd import geometry, ptu_dijkstra
    from rvgp_amd.graph import from_networkx
    X = sphere(300, 4) * 0.3
    g = geometry.manifold_graph(X, typ='affinity')
    Go, Wo = cpu.affinity_graph(X)
    W = g.to_scipy()
    assert W.nnz == Wo.nnz
    np.testing.assert_array_equal(W.indptr, Wo.indptr)
    np.testing.assert_array_equal(W.indices, Wo.indices)
    np.testing.assert_allclose(W.data, Wo.data, rtol=0, atol=1e-13)
    G = g.to_networkx()
    assert G.number_of_edges() == Go.number_of_edges()
    assert np.allclose(G.nodes[5]["pos"], X[5])
    # weighted Dijkstra neighbourhoods on the oracle's weights
    go = from_networkx(Go)
    K = 12
    nbh = ptu_dijkstra.geo_nbh_device(go, K).cpu().numpy()
    rp, c = Wo.indptr.astype(np.int32), Wo.indices.astype(np.int32)
    ref, _ = cpu.geo_nbh(X, rp, c, K, weights=Wo.data, centred=False)
    np.testing.assert_array_equal(nbh, ref)
    with pytest.raises(UnboundLocalError):
        geometry.manifold_graph(X, typ='radius')


@pytest.mark.parametrize("frac", [0.5, 0.8, 0.95, 1.0])
def test_quiet_dimension_equals_the_host_estimate(frac):
    """The data object's device reduction of the explained-variance profile
    picks the same dimension as manifold_dimension (geometry.py:83-97) on
    the host, for singular-value profiles of 1-, 2- and 3-dimensional
    point clouds."""
    import torch
    from rvgp_amd import geometry
    from rvgp_amd.dataclass import _quiet_dimension
    rng = np.random.default_rng(7)
    for d in (1, 2, 3):
        S = np.abs(rng.normal(scale=0.05, size=(5000, 12)))
        S[:, :d] += 1.0 + rng.uniform(size=(5000, d))
        S = -np.sort(-S, axis=1)
        want = geometry.manifold_dimension(S.copy(), frac_explained=frac)
        got = _quiet_dimension(torch.as_tensor(S, device="cuda"), frac)
        assert got == want, (d, frac, got, want)
