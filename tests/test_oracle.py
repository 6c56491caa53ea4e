"""CPU tests: pin the oracle against the reference's golden vectors, library
routines, closed forms, finite differences and brute force (no GPU)."""
import numpy as np
import pytest
import scipy.linalg
import scipy.sparse as sp

from oracle import cpu
from tests.util import golden_cases, sphere, torus


# ---------------------------------------------------------------- a1 FPS
@pytest.mark.parametrize("n,D,spacing,N,start", [(400, 3, 0.1, None, 0), (300, 8, 0.2, None, 5),
                                                 (500, 3, 0.1, 120, 3)])
def test_fps_direct_matches_sklearn_formula(n, D, spacing, N, start):
    x = np.random.default_rng(n).normal(size=(n, D))
    p1, l1 = cpu.fps(x, N, spacing, start)
    p2, l2 = cpu.fps(x, N, spacing, start, metric="sklearn")
    np.testing.assert_array_equal(p1, p2)
    np.testing.assert_allclose(l1, l2, rtol=1e-12, atol=1e-12)


def test_fps_brute_force_tiny():
    x = np.random.default_rng(0).normal(size=(12, 2))
    Dm = np.sqrt(((x[:, None] - x[None]) ** 2).sum(-1))
    p, l = cpu.fps(x, N=12)
    sel = [0]
    for i in range(1, 12):  # farthest from the selected set, lowest index on ties
        dmin = Dm[:, sel].min(1)
        sel.append(int(np.argmax(dmin)))
    np.testing.assert_array_equal(p, sel)
    assert np.all(np.diff(l[1:]) <= 1e-15)
    assert abs(cpu.diameter(x) - Dm.max()) < 1e-12


def test_fps_spacing_stop_rule():
    x = torus(3000, 1)
    p, l = cpu.fps(x, spacing=0.05)
    diam = cpu.diameter(x)
    assert np.all(l[1:] / diam >= 0.05)
    p_all, l_all = cpu.fps(x, N=len(p) + 1)
    assert l_all[len(p)] / diam < 0.05  # the dropped point triggered the stop


# ---------------------------------------------------------------- a2 graph
@pytest.mark.parametrize("n,D,k", [(500, 3, 10), (400, 32, 20), (300, 8, 15)])
def test_knn_graph_matches_sklearn_networkx(n, D, k):
    x = np.random.default_rng(n + k).normal(size=(n, D))
    rp, c = cpu.sym_csr(cpu.knn(x, k))
    rp2, c2, data = cpu.graph_csr_networkx(x, k)
    np.testing.assert_array_equal(rp, rp2)
    np.testing.assert_array_equal(c, c2)
    assert np.all(data == 1.0)


def test_knn_brute_force_ties():
    g = np.stack(np.meshgrid(np.arange(5), np.arange(5), indexing="ij"), -1).reshape(-1, 2) * 1.0
    nbr = cpu.knn(g, 4)
    for i in range(len(g)):
        d2 = ((g - g[i]) ** 2).sum(1)
        d2[i] = np.inf
        order = np.lexsort((np.arange(len(g)), d2))[:4]
        np.testing.assert_array_equal(nbr[i], order)


# ---------------------------------------------------------------- a3 heap + gauges
@pytest.mark.parametrize("case", golden_cases())
def test_heap_gauges_match_reference_golden(case):
    """Golden vectors written by the reference's own compiled ptu_dijkstra."""
    z = np.load(case)
    X, k, K = z["X"], int(z["k"]), int(z["K"])
    rp, c = cpu.sym_csr(cpu.knn(X, k))
    knn = sp.csr_matrix((np.ones(len(z["knn_indices"])), z["knn_indices"], z["knn_indptr"]))
    S = knn.maximum(knn.T).tocsr()
    S.sort_indices()
    np.testing.assert_array_equal(rp, S.indptr)
    np.testing.assert_array_equal(c, S.indices)
    T, Sig = cpu.tangent_frames(X, rp, c, X.shape[1], K)
    np.testing.assert_allclose(Sig, z["Sigma"], rtol=1e-10, atol=1e-13)
    d = int(z["dim"])
    assert cpu.manifold_dimension(Sig, 0.8) == d
    P = np.einsum("nij,nkj->nik", T[:, :, :d], T[:, :, :d])
    Pr = np.einsum("nij,nkj->nik", z["tangents"][:, :, :d], z["tangents"][:, :, :d])
    assert np.abs(P - Pr).max() < 1e-10
    # connections from the reference's gauges vs the reference's R
    R = cpu.connections(np.ascontiguousarray(z["tangents"][:, :, :d]), rp, c)
    coo = cpu.connections_coo(R, rp, c)
    order = np.lexsort((coo.col, coo.row))
    oref = np.lexsort((z["R_col"], z["R_row"]))
    np.testing.assert_array_equal(coo.row[order], z["R_row"][oref])
    np.testing.assert_array_equal(coo.col[order], z["R_col"][oref])
    np.testing.assert_allclose(coo.data[order], z["R_data"][oref], atol=1e-12)


def test_heap_neighbourhood_is_a_dijkstra_prefix():
    """Whatever the tie order, the K+1 popped nodes are a BFS-ball prefix."""
    X = sphere(300, 2)
    rp, c = cpu.sym_csr(cpu.knn(X, 6))
    nbh, _ = cpu.geo_nbh(X, rp, c, 12, centred=False)
    A = sp.csr_matrix((np.ones(len(c)), c, rp))
    from scipy.sparse.csgraph import shortest_path
    H = shortest_path(A, unweighted=True)
    for i in range(300):
        h = H[i, nbh[i]]
        assert nbh[i][0] == i and np.all(np.diff(h) >= 0)
        assert np.sum(H[i] < h.max()) <= 13  # every strictly closer node is included
        assert len(set(nbh[i].tolist())) == 13


def test_tangent_frames_validation():
    X = sphere(50, 0)
    rp, c = cpu.sym_csr(cpu.knn(X, 5))
    with pytest.raises(ValueError):
        cpu.tangent_frames(X, rp, c, 3, 50)
    with pytest.raises(ValueError):
        cpu.tangent_frames(X, rp, c, 3, 2)
    with pytest.raises(ValueError):
        cpu.tangent_frames(X, rp, c, 4, 10)
    P = np.c_[np.random.default_rng(0).normal(size=(80, 2)), np.zeros(80)]
    rp, c = cpu.sym_csr(cpu.knn(P, 8))
    with pytest.raises(RuntimeError):
        cpu.tangent_frames(P, rp, c, 3, 12)


# ---------------------------------------------------------------- a5-a7
def test_connections_polar_and_invariants():
    X = sphere(300, 5)
    rp, c = cpu.sym_csr(cpu.knn(X, 8))
    T, _ = cpu.tangent_frames(X, rp, c, 3, 12)
    T2 = np.ascontiguousarray(T[:, :, :2])
    R = cpu.connections(T2, rp, c)
    rows = np.repeat(np.arange(300), np.diff(rp))
    M = np.einsum("eki,ekj->eij", T2[rows], T2[c])
    for e in range(0, len(c), 97):
        U, _ = scipy.linalg.polar(M[e])
        np.testing.assert_allclose(R[e], U, atol=1e-12)
    np.testing.assert_allclose(np.einsum("eji,ejk->eik", R, R), np.broadcast_to(np.eye(2), R.shape),
                               atol=1e-12)
    diag = rows == c
    np.testing.assert_allclose(R[diag], np.broadcast_to(np.eye(2), R[diag].shape), atol=1e-12)
    # R_ji = R_ij^T
    key = {(i, j): e for e, (i, j) in enumerate(zip(rows, c))}
    for e in range(0, len(c), 53):
        i, j = rows[e], c[e]
        np.testing.assert_allclose(R[key[(j, i)]], R[e].T, atol=1e-12)


def test_laplacians_vs_networkx_and_invariants():
    import networkx as nx
    X = sphere(400, 6)
    rp, c = cpu.sym_csr(cpu.knn(X, 9))
    L = cpu.laplacian(rp, c)
    A = sp.csr_matrix((np.ones(len(c)), c, rp))
    G = nx.from_scipy_sparse_array(A)
    Lnx = sp.csr_matrix(nx.laplacian_matrix(G), dtype=np.float64)
    assert abs(L - Lnx).max() == 0
    assert np.abs(L @ np.ones(400)).max() == 0
    T, _ = cpu.tangent_frames(X, rp, c, 3, 13)
    R = cpu.connections(np.ascontiguousarray(T[:, :, :2]), rp, c)
    Lc = cpu.connection_laplacian(rp, c, R)
    assert abs(Lc - Lc.T).max() < 1e-12
    assert np.linalg.eigvalsh(Lc.toarray()).min() > -1e-10
    Bl = cpu.lc_blocks(rp, c, R)
    Lb = sp.bsr_matrix((Bl, c, rp), shape=Lc.shape).tocsr()
    assert abs(Lb - Lc).max() < 1e-14


# ---------------------------------------------------------------- a8 spectrum
def test_spectrum_ring_closed_form():
    n = 60
    rows = np.arange(n)
    A = sp.csr_matrix((np.ones(2 * n), (np.r_[rows, rows], np.r_[(rows + 1) % n, (rows - 1) % n])),
                      shape=(n, n))
    L = sp.diags(np.asarray(A.sum(1)).ravel()) - A
    w, V = cpu.compute_spectrum(sp.csr_matrix(L), 9)
    ref = np.sort(2 - 2 * np.cos(2 * np.pi * np.arange(n) / n))[:9]
    np.testing.assert_allclose(w, ref, atol=1e-10)
    np.testing.assert_allclose(np.linalg.norm(V, axis=0), np.sqrt(n))  # geometry.py:75
    with pytest.raises(TypeError):
        cpu.compute_spectrum(sp.csr_matrix(L), n - 1)


def test_spectrum_eigsh_vs_eigh():
    X = sphere(300, 7)
    rp, c = cpu.sym_csr(cpu.knn(X, 8))
    L = cpu.laplacian(rp, c)
    w1, _ = cpu.compute_spectrum(L, 20)
    w2, _ = cpu.compute_spectrum(L, 20, dense=True)
    np.testing.assert_allclose(w1, w2, atol=1e-9)


# ---------------------------------------------------------------- a11 smoothing
def test_heat_expm_equals_eigen_form_and_preserves_constants():
    X = sphere(150, 8)
    rp, c = cpu.sym_csr(cpu.knn(X, 8))
    L = cpu.laplacian(rp, c)
    x = np.random.default_rng(0).normal(size=(150, 2))
    np.testing.assert_allclose(cpu.heat(L, 3.0, x, "expm"), cpu.heat(L, 3.0, x, "eigh"), atol=1e-10)
    np.testing.assert_allclose(cpu.heat(L, 50.0, np.ones((150, 1)), "expm"), 1.0, atol=1e-10)


# ---------------------------------------------------------------- a12-a16 GP
def test_spectral_filter_finite_differences():
    lam = np.linspace(0, 6, 40)
    for typ in ("matern", "se"):
        th = np.array([1.7, 2.3, 0.9])
        S, dS = cpu.spectral_filter(lam, *th, 500.0, typ)
        assert abs(S.sum() - 500.0 * th[2]) < 1e-9
        for p in range(3):
            h = 1e-6
            tp, tm = th.copy(), th.copy()
            tp[p] += h
            tm[p] -= h
            fd = (cpu.spectral_filter(lam, *tp, 500.0, typ)[0] - cpu.spectral_filter(lam, *tm, 500.0, typ)[0]) / (2 * h)
            np.testing.assert_allclose(dS[p], fd, rtol=1e-6, atol=1e-8)


def test_gpr_nlml_vs_multivariate_normal_and_fd():
    from scipy.stats import multivariate_normal
    rng = np.random.default_rng(3)
    Phi = rng.normal(size=(60, 12))
    y = rng.normal(size=(60, 1))
    S = rng.uniform(0.1, 2.0, size=12)
    nlml = cpu.gpr_nlml(Phi, y, S, 0.3)
    ref = -multivariate_normal(mean=np.zeros(60), cov=(Phi * S) @ Phi.T + 0.3 * np.eye(60)).logpdf(y.ravel())
    assert abs(nlml - ref) < 1e-9 * abs(ref)
    gS, gn = cpu.gpr_grad(Phi, y, S, 0.3)
    h = 1e-6
    for j in (0, 5, 11):
        Sp, Sm = S.copy(), S.copy()
        Sp[j] += h
        Sm[j] -= h
        fd = (cpu.gpr_nlml(Phi, y, Sp, 0.3) - cpu.gpr_nlml(Phi, y, Sm, 0.3)) / (2 * h)
        assert abs(fd - gS[j]) < 1e-6
    fd = (cpu.gpr_nlml(Phi, y, S, 0.3 + h) - cpu.gpr_nlml(Phi, y, S, 0.3 - h)) / (2 * h)
    assert abs(fd - gn) < 1e-6


def test_gpr_multi_rhs_is_sum_of_columns():
    """R right-hand sides with shared hyperparameters: NLML = sum of the
    columns' -log densities (scipy), gradient = sum of the columns'
    gradients, and central differences of the summed NLML."""
    from scipy.stats import multivariate_normal
    rng = np.random.default_rng(13)
    Phi = rng.normal(size=(45, 10))
    Y = rng.normal(size=(45, 5))
    S = rng.uniform(0.1, 2.0, size=10)
    cov = (Phi * S) @ Phi.T + 0.2 * np.eye(45)
    ref = -sum(multivariate_normal(mean=np.zeros(45), cov=cov).logpdf(Y[:, r]) for r in range(5))
    assert abs(cpu.gpr_nlml(Phi, Y, S, 0.2) - ref) < 1e-9 * abs(ref)
    gS, gn = cpu.gpr_grad(Phi, Y, S, 0.2)
    cols = [cpu.gpr_grad(Phi, Y[:, r:r + 1], S, 0.2) for r in range(5)]
    np.testing.assert_allclose(gS, sum(c[0] for c in cols), rtol=1e-10)
    assert abs(gn - sum(c[1] for c in cols)) < 1e-10 * abs(gn)
    h = 1e-6
    Sp, Sm = S.copy(), S.copy()
    Sp[3] += h
    Sm[3] -= h
    fd = (cpu.gpr_nlml(Phi, Y, Sp, 0.2) - cpu.gpr_nlml(Phi, Y, Sm, 0.2)) / (2 * h)
    assert abs(fd - gS[3]) < 1e-5


def test_gpr_predict_conditioning_identity():
    """Predicting at the training inputs with noise -> 0 interpolates y."""
    rng = np.random.default_rng(4)
    Phi = rng.normal(size=(20, 40))
    y = rng.normal(size=(20, 1))
    S = np.ones(40)
    m, v = cpu.gpr_predict(Phi, y, S, 1e-10, Phi)
    np.testing.assert_allclose(m.ravel(), y.ravel(), atol=1e-6)
    assert np.abs(v).max() < 1e-6


def test_train_test_split_rule_matches_sklearn():
    from sklearn.model_selection import train_test_split
    for n, seed in ((10, 0), (1000, 0), (37, 5)):
        idx = np.arange(n)
        tr, te = train_test_split(idx, test_size=0.2, random_state=seed)
        tr2, te2 = cpu.train_test_split_indices(n, 0.2, seed)
        np.testing.assert_array_equal(tr, tr2)
        np.testing.assert_array_equal(te, te2)


def test_load_mesh_reads_the_reference_cube(golden_dir, tmp_path):
    """utils.load_mesh on the reference's cube.obj (Blender export: '#'
    comments, 'o' and 's' records, 8 vertices, 12 one-based faces)."""
    from rvgp_amd.utils import load_mesh
    v, f = load_mesh('cube', folder=golden_dir)
    assert v.shape == (8, 3) and v.dtype == np.float64
    assert f.shape == (12, 3)
    np.testing.assert_array_equal(v[0], [1.0, 1.0, -1.0])
    np.testing.assert_array_equal(v[7], [-1.0, -1.0, 1.0])
    np.testing.assert_array_equal(f[0], [4, 2, 0])
    np.testing.assert_array_equal(f[-1], [4, 0, 1])
    assert f.min() == 0 and f.max() == 7
    # every vertex of a closed triangulated cube is used; 12 faces x 3 = 36
    assert np.bincount(f.ravel()).sum() == 36
    # reference error behaviour: a blank line -> IndexError, 'f 1/1 2/2 3/3' -> ValueError
    import pytest
    (tmp_path / "blank.obj").write_text("v 0 0 0\n\nv 1 1 1\n")
    with pytest.raises(IndexError):
        load_mesh('blank', folder=str(tmp_path))
    (tmp_path / "tex.obj").write_text("v 0 0 0\nf 1/1 2/2 3/3\n")
    with pytest.raises(ValueError):
        load_mesh('tex', folder=str(tmp_path))


def test_rbf_gpr_oracle_vs_multivariate_normal_and_fd():
    """The RBF-branch oracle: summed NLML = sum of scipy log-densities,
    gradient = central differences, noise-free prediction interpolates."""
    from scipy.stats import multivariate_normal
    rng = np.random.default_rng(17)
    X = rng.normal(size=(40, 6))
    Y = rng.normal(size=(40, 3))
    th = (1.3, 2.1, 0.05)
    K = cpu.rbf_kernel(X, X, th[0], th[1]) + th[2] * np.eye(40)
    ref = -sum(multivariate_normal(mean=np.zeros(40), cov=K).logpdf(Y[:, r]) for r in range(3))
    assert abs(cpu.rbf_gpr_nlml(X, Y, *th) - ref) < 1e-10 * abs(ref)
    g = cpu.rbf_gpr_grad(X, Y, *th)
    h = 1e-6
    for p in range(3):
        tp, tm = list(th), list(th)
        tp[p] += h
        tm[p] -= h
        fd = (cpu.rbf_gpr_nlml(X, Y, *tp) - cpu.rbf_gpr_nlml(X, Y, *tm)) / (2 * h)
        assert abs(fd - g[p]) < 1e-5 * max(1.0, abs(fd)), (p, fd, g[p])
    m, v = cpu.rbf_gpr_predict(X, Y, 1.0, 1.5, 1e-10, X)
    np.testing.assert_allclose(m, Y, atol=1e-5)
    assert np.abs(v).max() < 1e-5


def test_sgpr_oracle_bound_properties():
    """The SGPR oracle (GPflow's SGPR.elbo formula): (1) with the training
    inputs themselves as inducing points (N < k, so K_uu is regular) the bound
    equals the exact GPR log marginal likelihood up to the jitter; (2) with
    fewer inducing points it is a lower bound (-ELBO >= NLML); (3) the
    predictive mean with Z = X equals the exact GPR mean."""
    rng = np.random.default_rng(23)
    Phi = rng.normal(size=(12, 20))
    y = rng.normal(size=(12, 1))
    S = rng.uniform(0.2, 2.0, 20)
    exact = cpu.gpr_nlml(Phi, y, S, 0.3)
    assert abs(cpu.sgpr_nlml(Phi, y, S, 0.3, Phi, jitter=1e-12) - exact) < 1e-7 * abs(exact)
    Phi2 = rng.normal(size=(80, 20))
    y2 = rng.normal(size=(80, 1))
    for m in (3, 10, 25):
        Z = Phi2[rng.choice(80, m, replace=False)]
        assert cpu.sgpr_nlml(Phi2, y2, S, 0.3, Z) >= cpu.gpr_nlml(Phi2, y2, S, 0.3) - 1e-9
    Ps = rng.normal(size=(7, 20))
    m1, v1 = cpu.sgpr_predict(Phi, y, S, 0.3, Phi, Ps, jitter=1e-12)
    m2, v2 = cpu.gpr_predict(Phi, y, S, 0.3, Ps)
    np.testing.assert_allclose(m1.ravel(), m2.ravel(), rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(v1.ravel(), v2.ravel(), rtol=1e-6, atol=1e-9)


# ---------------------------------------------------------------- a9-a11 frames, Phi, diffusion
def _orthonormal_frames(n, D, d, seed):
    rng = np.random.default_rng(seed)
    T = np.empty((n, D, d))
    for i in range(n):
        q, _ = np.linalg.qr(rng.normal(size=(D, d)))
        T[i] = q
    return T


def test_frame_maps_pinned_by_projector_identities():
    """project_to_manifold (geometry.py:165-168) is the orthogonal projector
    T T^T: idempotent, identity on the tangent plane, residual orthogonal to
    it; express_in_local_frame (geometry.py:171-175) and its reverse are
    T^T and T, so T^T (T c) = c and T (T^T x) = P x."""
    n, D, d = 50, 5, 2
    T = _orthonormal_frames(n, D, d, 0)
    rng = np.random.default_rng(1)
    x = rng.normal(size=(n, D))
    c = rng.normal(size=(n, d))
    Px = cpu.project_to_manifold(x, T)
    np.testing.assert_allclose(cpu.project_to_manifold(Px, T), Px, atol=1e-14)
    tang = np.einsum("nij,nj->ni", T, c)  # tangent vectors built without the oracle
    np.testing.assert_allclose(cpu.project_to_manifold(tang, T), tang, atol=1e-14)
    res = x - Px
    assert np.abs(np.einsum("nij,ni->nj", T, res)).max() < 1e-14
    assert np.all(np.linalg.norm(Px, axis=1) <= np.linalg.norm(x, axis=1) + 1e-14)
    loc = cpu.express_in_local_frame(tang, T)
    np.testing.assert_allclose(loc, c, atol=1e-14)
    back = cpu.express_in_local_frame(cpu.express_in_local_frame(x, T), T, reverse=True)
    np.testing.assert_allclose(back, Px, atol=1e-14)
    # local coordinates rotate with the gauge: T -> T O gives c -> O^T c
    O = _orthonormal_frames(n, d, d, 2)
    TO = np.einsum("nij,njk->nik", T, O)
    np.testing.assert_allclose(cpu.express_in_local_frame(x, TO),
                               np.einsum("nji,nj->ni", O, cpu.express_in_local_frame(x, T)),
                               atol=1e-14)


def test_data_phi_is_the_ambient_eigenbasis():
    """Data's Phi = einsum(gauges, U) (dataclass.py:57-59): Phi^T Phi = U^T U
    (orthonormal frames), every node's rows lie in its tangent plane, and
    Phi diagonalises the ambient connection Laplacian B Lc B^T
    (B = blockdiag(T_i)): B Lc B^T Phi = Phi diag(evals_Lc).  A transposed
    frame or a contraction over the wrong index fails all three."""
    X = sphere(300, 3)
    o = cpu.Data(X, n_neighbors=10, n_eigenpairs=12, dense_eigh=True)
    n, D, d, k = o.n, X.shape[1], o.dim_man, o.U.shape[1]
    Phi, U = o.evecs_Lc, o.U
    np.testing.assert_allclose(Phi.T @ Phi, U.T @ U, atol=1e-10 * n)
    P = np.einsum("nij,nkj->nik", o.gauges, o.gauges)
    Pr = Phi.reshape(n, D, k)
    np.testing.assert_allclose(np.einsum("nij,njk->nik", P, Pr), Pr, atol=1e-12)
    B = sp.block_diag([o.gauges[i] for i in range(n)], format="csr")
    Aamb = B @ o.Lc @ B.T
    np.testing.assert_allclose(Aamb @ Phi, Phi * o.evals_Lc, atol=1e-9 * np.abs(Phi).max())


def _trivial_connection(rp, cols, d):
    R = np.zeros((len(cols), d, d))
    R[:] = np.eye(d)
    return R


def test_vector_diffusion_pinned_by_closed_forms():
    """vector_diffusion (smoothing.py:37-64, normalise=True):
    * trivial connection (R = I, Lc = L (x) I) and a field a_i v of one
      direction with positive magnitudes a: the result is (e^{-tL} a)_i v,
      with e^{-tL} from an independent eigendecomposition of L -- a
      normalisation by |c| instead of |out|, or by e^{-tL}|c| without
      e^{-tL}1, fails it;
    * any connection: |result_i| = (e^{-tL}|x|)_i / (e^{-tL} 1)_i and
      result_i is parallel to (e^{-tLc} x)_i (Lc's heat kernel from its own
      eigendecomposition);
    * t -> 0 returns x."""
    X = torus(200, 4)
    rp, cols = cpu.sym_csr(cpu.knn(X, 8))
    n, d = len(rp) - 1, 2
    L = cpu.laplacian(rp, cols).toarray()
    w, V = np.linalg.eigh(L)
    def heatL(t, y):
        return (V @ (np.exp(-t * w)[:, None] * (V.T @ y.reshape(n, -1)))).reshape(y.shape)

    Lc_triv = cpu.connection_laplacian(rp, cols, _trivial_connection(rp, cols, d))
    np.testing.assert_allclose(Lc_triv.toarray(), np.kron(L, np.eye(d)), atol=1e-14)
    rng = np.random.default_rng(5)
    a = rng.uniform(0.2, 3.0, n)
    v = np.array([0.6, -0.8])
    x = a[:, None] * v
    for t in (0.3, 2.0):
        out = cpu.vector_diffusion(x, t, Lc_triv, cpu.laplacian(rp, cols), method="eigh")
        np.testing.assert_allclose(out, heatL(t, a)[:, None] * v, atol=1e-12)
    # a random orthogonal connection (R_ji = R_ij^T)
    Rb = np.empty((len(cols), d, d))
    rows = np.repeat(np.arange(n), np.diff(rp))
    ang = {}
    for e, (i, j) in enumerate(zip(rows, cols)):
        key = (min(i, j), max(i, j))
        if key not in ang:
            ang[key] = 0.0 if i == j else rng.uniform(-np.pi, np.pi)
        th = ang[key] if i <= j else -ang[key]
        Rb[e] = [[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]]
    Lc = cpu.connection_laplacian(rp, cols, Rb)
    wc, Vc = np.linalg.eigh(Lc.toarray())
    x = rng.normal(size=(n, d))
    t = 0.7
    out = cpu.vector_diffusion(x, t, Lc, cpu.laplacian(rp, cols), method="expm")
    hc = (Vc @ (np.exp(-t * wc)[:, None] * (Vc.T @ x.reshape(-1, 1)))).reshape(n, d)
    nrm = heatL(t, np.linalg.norm(x, axis=1)) / heatL(t, np.ones(n))
    np.testing.assert_allclose(np.linalg.norm(out, axis=1), nrm, rtol=1e-10)
    cosang = np.einsum("ni,ni->n", out, hc) / (np.linalg.norm(out, axis=1) * np.linalg.norm(hc, axis=1))
    np.testing.assert_allclose(cosang, 1.0, atol=1e-12)
    np.testing.assert_allclose(cpu.vector_diffusion(x, 1e-13, Lc, cpu.laplacian(rp, cols)), x,
                               atol=1e-11)


def test_fit_lbfgs_returns_a_stationary_minimum():
    """fit_lbfgs (main.py:87-95 via scipy L-BFGS-B): at the returned
    hyperparameters the gradient with respect to every parameter off its
    lower bound is ~0 against the gradient at the start, and no coordinate
    step of +-1e-3 (relative) lowers the NLML."""
    X = sphere(250, 6)
    o = cpu.Data(X, n_neighbors=10, n_eigenpairs=20, dense_eigh=True)
    rng = np.random.default_rng(7)
    nodes = rng.choice(o.n, size=120, replace=False)
    o.random_vector_field(seed=3)
    # a field smooth enough that the optimum of kappa is finite (a rough one
    # sends kappa up a flat ridge, where L-BFGS-B stops on its relative
    # decrease of f before the gradient is small)
    o.smooth_vector_field(t=100, method="eigh")
    k = o.evecs_Lc.shape[1]
    Phi = o.evecs_Lc.reshape(o.n, -1)[nodes].reshape(-1, k)
    y = o.vectors[nodes].reshape(-1, 1) + 0.01 * rng.normal(size=(len(nodes) * 3, 1))
    th, res = cpu.fit_lbfgs(Phi, y, o.evals_Lc, o.n * 3, maxiter=500)
    assert res.success, res.message

    def f_and_g(t):
        S, dS = cpu.spectral_filter(o.evals_Lc, t[0], t[1], t[2], o.n * 3)
        gS, gn = cpu.gpr_grad(Phi, y, S, t[3])
        return cpu.gpr_nlml(Phi, y, S, t[3]), np.array([dS[0] @ gS, dS[1] @ gS, dS[2] @ gS, gn])

    f0, g0 = f_and_g(np.array([1.5, 5.0, 1.0, 1.0]))
    fs, gs = f_and_g(th)
    assert fs < f0
    lower = np.array([0.0, 0.0, 0.0, 1e-6])
    free = th > lower * 1.01 + 1e-8
    assert np.all(np.abs(gs[free] * th[free]) < 1e-3 * np.abs(g0[free] * np.array([1.5, 5.0, 1.0, 1.0])[free]) + 1e-4)
    for i in range(4):
        for s in (-1e-3, 1e-3):
            tp = th.copy()
            tp[i] *= 1 + s
            if tp[i] <= lower[i]:
                continue
            assert f_and_g(tp)[0] >= fs - 1e-9 * abs(fs)


def test_affinity_graph_oracle_is_the_gaussian_of_all_pairs():
    """affinity_graph (geometry.py:114-118): every pair i, j (self loops
    included) carries exp(-|x_i - x_j|^2 / (2 * 0.1^2)), brute force; the
    graph is undirected and complete where no weight underflows."""
    x = np.random.default_rng(2).normal(size=(40, 3)) * 0.15
    G, W = cpu.affinity_graph(x)
    d2 = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)
    ref = np.exp(-d2 / (2 * 0.1 ** 2))
    np.testing.assert_allclose(W.toarray(), ref, rtol=1e-12, atol=1e-300)
    assert G.number_of_edges() == 40 * 41 // 2
    assert np.allclose(W.diagonal(), 1.0)


def test_spectral_diffusion_oracle_equals_expm_on_the_full_basis():
    """The "spectral" method (smoothing.py:17-32) with every eigenpair is the
    dense heat kernel (== "expm"); a common scale of the eigenvectors (the
    sqrt(rows) stored by compute_spectrum, geometry.py:75) cancels in the
    normalised vector diffusion; a truncated basis only approximates it."""
    X = sphere(150, 9)
    rp, cols = cpu.sym_csr(cpu.knn(X, 8))
    T, S = cpu.tangent_frames(X, rp, cols, 3, 12)
    R = cpu.connections(np.ascontiguousarray(T[:, :, :2]), rp, cols)
    L, Lc = cpu.laplacian(rp, cols), cpu.connection_laplacian(rp, cols, R)
    wl, Vl = np.linalg.eigh(L.toarray())
    wc, Vc = np.linalg.eigh(Lc.toarray())
    x = np.random.default_rng(1).normal(size=(150, 2))
    ref = cpu.vector_diffusion(x, 0.8, Lc, L, method="expm")
    full = cpu.vector_diffusion(x, 0.8, (wc, Vc), (wl, Vl), method="spectral")
    np.testing.assert_allclose(full, ref, atol=1e-10)
    scaled = cpu.vector_diffusion(x, 0.8, (wc, Vc * np.sqrt(300)), (wl, Vl * np.sqrt(150)),
                                  method="spectral")
    np.testing.assert_allclose(scaled, full, atol=1e-12)
    trunc = cpu.vector_diffusion(x, 0.8, (wc[:60], Vc[:, :60]), (wl[:30], Vl[:, :30]),
                                 method="spectral")
    assert np.abs(trunc - ref).max() > 1e-6
