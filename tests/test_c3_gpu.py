"""The headline configuration's shape at reduced n, against the FP64 oracle.

Flat 3-torus in R^32 (angles uniform, cos/sin pairs in R^6, a seeded
orthogonal map into R^32, N(0, 0.01^2) noise), k = 20 neighbours, so the heap
neighbourhoods hold K + 1 = 31 < D = 32 points (the documented D > K
extension, reference_semantics=False; the reference raises ValueError at
ptu_dijkstra.pyx:73-77), dim_man = 3 (3 x 3 connection blocks), about 100
eigenpairs, in the FP64 and the mixed form.  The oracle is cpu.Data with
allow_rank_deficient_k (dataclass.py:22-89: C kNN + heap, LAPACK SVD gauges,
polar connections, scipy Laplacians, eigsh 'SM' as geometry.py:66-80).

Bars (the north-star's):
* kNN graph: identical CSR;
* gauges: Sigma and the tangent projectors T T^T (gauge invariant);
* eigenvalues of Lc within 1e-6 relative (absolute floor 1e-9 of the
  spectral bound for eigenvalues near 0);
* eigenspaces: largest principal angle below 1e-4 rad over clusters of the
  oracle's eigenvalues delimited by relative gaps above 1e-5, compared in
  ambient coordinates Phi = T U (gauge invariant);
* predictions at test nodes within 1e-4 relative of the oracle's GPR at the
  same hyperparameters.
"""
import numpy as np
import pytest
import scipy.linalg

from oracle import cpu
from tests.util import flat_torus3_with_field

pytestmark = pytest.mark.gpu

N_POINTS, K_NN, K_TARGET = 4000, 20, 100


@pytest.fixture(scope="module")
def oracle_c3():
    X, V = flat_torus3_with_field(N_POINTS, 0)
    o = cpu.Data(X, n_neighbors=K_NN, n_eigenpairs=K_TARGET + 20, allow_rank_deficient_k=True)
    w = o.evals_Lc
    # cut the spectrum where it has a gap, so the first k eigenvectors span
    # the same space in every solver (a cut inside a near-degenerate
    # cluster would leave the span undetermined)
    lo, hi = K_TARGET - 10, K_TARGET + 10
    k = lo + int(np.argmax(w[lo:hi] - w[lo - 1:hi - 1]))
    print(f"[c3] n={N_POINTS} dim_man={o.dim_man} k={k} gap={w[k] - w[k - 1]:.3e} "
          f"lambda_1..3={w[:3]}", flush=True)
    return X, V, o, k


@pytest.fixture(scope="module", params=["fp64", "mixed"])
def gpu_c3(request, oracle_c3):
    import rvgp_amd as RVGP
    X, V, o, k = oracle_c3
    d = RVGP.create_data_object(X, vectors=V, n_neighbors=K_NN, n_eigenpairs=k,
                                precision=request.param, reference_semantics=False, verbose=False)
    return request.param, d


def test_graph_and_gauges_match_oracle(oracle_c3, gpu_c3):
    X, V, o, k = oracle_c3
    prec, d = gpu_c3
    rp, c = d._graph.csr_numpy()
    np.testing.assert_array_equal(rp, o.rowptr)
    np.testing.assert_array_equal(c, o.cols)
    assert d.dim_man == o.dim_man == 3
    T, To = d.gauges, o.gauges
    P = np.einsum("nij,nkj->nik", T, T)
    Po = np.einsum("nij,nkj->nik", To, To)
    assert np.abs(P - Po).max() < 1e-8
    # the K = 30 singular values the K + 1 centred points can carry
    S = d._Sigma.cpu().numpy()
    np.testing.assert_allclose(S[:, :30], o.Sigma[:, :30], rtol=1e-8, atol=1e-10)


def _clusters(w, k, scale, rel_gap=1e-5):
    """Index ranges of gap-delimited clusters lying wholly inside [0, k)."""
    out, start = [], 0
    for j in range(1, k + 1):
        if j == k or (w[j] - w[j - 1]) > rel_gap * max(abs(w[j]), scale):
            out.append((start, j))
            start = j
    return out


def test_lc_spectrum_and_eigenspaces_match_oracle(oracle_c3, gpu_c3):
    X, V, o, k = oracle_c3
    prec, d = gpu_c3
    w = np.asarray(d.evals_Lc)
    wo = o.evals_Lc[:k]
    ub = d._Lcb.gershgorin_bound()
    rel = np.abs(w - wo) / np.maximum(np.abs(wo), 1e-9 * ub)
    print(f"[c3 {prec}] eigenvalue rel err max {rel.max():.3e}", flush=True)
    assert rel.max() < 1e-6
    Phi = d.evecs_Lc                      # (n*D, k) ambient, gauge invariant
    Phio = o.evecs_Lc[:, :k]
    worst = 0.0
    for a0, a1 in _clusters(wo, k, 1e-9 * ub):
        ang = scipy.linalg.subspace_angles(Phio[:, a0:a1], Phi[:, a0:a1]).max()
        worst = max(worst, ang)
    print(f"[c3 {prec}] largest cluster principal angle {worst:.3e} rad", flush=True)
    assert worst < 1e-4


def test_predictions_match_oracle(oracle_c3, gpu_c3):
    import rvgp_amd as RVGP
    X, V, o, k = oracle_c3
    prec, d = gpu_c3
    rng = np.random.default_rng(11)
    train = rng.choice(N_POINTS, size=200, replace=False)
    gp = RVGP.fit(d, train_ind=train, epochs=20, optimizer="adam", lr=0.05,
                  reference_semantics=False, precision=prec, verbose=False)
    test = [int(i) for i in rng.choice(np.setdiff1d(np.arange(N_POINTS), train), 500,
                                       replace=False)]
    mean, _ = gp.transform(d, test)
    th = gp.engine.theta_host()
    D = X.shape[1]
    Phi = o.evecs_Lc[:, :k].reshape(N_POINTS, -1)[train].reshape(-1, k)
    S, _ = cpu.spectral_filter(o.evals_Lc[:k], th[0], th[1], th[2], N_POINTS * D)
    Ps = o.evecs_Lc[:, :k].reshape(N_POINTS, -1)[test].reshape(-1, k)
    mr, _ = cpu.gpr_predict(Phi, V[train].reshape(-1, 1), S, th[3], Ps)
    err = np.linalg.norm(mean.ravel() - mr.ravel()) / np.linalg.norm(mr)
    print(f"[c3 {prec}] prediction rel err {err:.3e} at theta={th}", flush=True)
    assert err < 1e-4
