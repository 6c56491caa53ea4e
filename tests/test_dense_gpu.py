"""Matrix-core GEMM, Cholesky and triangular solves vs torch / numpy FP64."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _t(x, dt):
    import torch
    return torch.as_tensor(np.ascontiguousarray(x), device="cuda", dtype=dt)


# shapes on both sides of the small-GEMM rule (K <= 1024 and few 128-tiles):
# matrix-vector (500, 1, 500), a Cholesky-update shape (436, 436, 64), and
# K = 1000 / 1100 around the threshold
@pytest.mark.parametrize("M,N,K", [(70, 33, 45), (64, 64, 16), (1, 5, 3), (130, 200, 257),
                                   (500, 1, 500), (70, 1, 3000), (436, 436, 64), (96, 300, 1000),
                                   (96, 300, 1100)])
@pytest.mark.parametrize("tA", [False, True])
@pytest.mark.parametrize("tB", [False, True])
def test_gemm_modes(M, N, K, tA, tB):
    import torch
    from rvgp_amd import dense
    rng = np.random.default_rng(M * 7 + N + K)
    A = rng.normal(size=(K, M) if tA else (M, K))
    B = rng.normal(size=(N, K) if tB else (K, N))
    C0 = rng.normal(size=(M, N))
    s = rng.uniform(0.5, 2.0, size=K)
    opA = A.T if tA else A
    opB = B.T if tB else B
    ref = 0.75 * (opA * s) @ opB - 0.5 * C0
    # f64
    C = _t(C0, torch.float64)
    dense.gemm(_t(A, torch.float64), _t(B, torch.float64), C, tA, tB, 0.75, -0.5, _t(s, torch.float64))
    np.testing.assert_allclose(C.cpu().numpy(), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    # f32 in, f64 accumulate: exact products of the f32-rounded inputs
    A32, B32 = A.astype(np.float32).astype(np.float64), B.astype(np.float32).astype(np.float64)
    ref32 = 0.75 * ((A32.T if tA else A32) * s) @ (B32.T if tB else B32) - 0.5 * C0
    C = _t(C0, torch.float64)
    dense.gemm(_t(A, torch.float32), _t(B, torch.float32), C, tA, tB, 0.75, -0.5, _t(s, torch.float64))
    np.testing.assert_allclose(C.cpu().numpy(), ref32, rtol=1e-10, atol=1e-10 * np.abs(ref).max())
    # f32
    C = _t(C0, torch.float32)
    dense.gemm(_t(A, torch.float32), _t(B, torch.float32), C, tA, tB, 0.75, -0.5, _t(s, torch.float64))
    assert np.abs(C.cpu().numpy() - ref).max() < 1e-5 * np.abs(ref).max() * np.sqrt(K)
    # bf16 tiles, f32 accumulation
    C = _t(C0, torch.float32)
    dense.gemm(_t(A, torch.float32), _t(B, torch.float32), C, tA, tB, 0.75, -0.5,
               _t(s, torch.float64), mode=dense.GEMM_BF16)
    assert np.abs(C.cpu().numpy() - ref).max() < 2e-2 * np.abs(ref).max()


def test_gemm_split_k_tall_skinny():
    import torch
    from rvgp_amd import dense
    rng = np.random.default_rng(0)
    X = rng.normal(size=(40000, 90))
    Y = rng.normal(size=(40000, 70))
    ref = X.T @ Y
    for split in (None, 1, 7):
        C = dense.gemm(_t(X, torch.float64), _t(Y, torch.float64), transA=True, split_k=split)
        np.testing.assert_allclose(C.cpu().numpy(), ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())
    C1 = dense.gemm(_t(X, torch.float64), _t(Y, torch.float64), transA=True, split_k=9)
    C2 = dense.gemm(_t(X, torch.float64), _t(Y, torch.float64), transA=True, split_k=9)
    assert torch.equal(C1, C2)  # deterministic split-K reduction


@pytest.mark.parametrize("n", [1, 5, 63, 64, 65, 200, 333])
@pytest.mark.parametrize("f64", [True, False])
def test_potrf_trsm_logdet(n, f64):
    import torch
    from rvgp_amd import dense
    rng = np.random.default_rng(n)
    G = rng.normal(size=(n, n + 3))
    A = G @ G.T + n * np.eye(n)
    dt = torch.float64 if f64 else torch.float32
    tol = 1e-11 if f64 else 2e-4
    L = _t(A, dt)
    dense.potrf(L)
    Lh = np.tril(L.cpu().double().numpy())
    Lr = np.linalg.cholesky(A)
    np.testing.assert_allclose(Lh, Lr, rtol=0, atol=tol * np.abs(Lr).max())
    Bm = rng.normal(size=(n, 7))
    X = _t(Bm, dt)
    dense.trsm(L, X)
    np.testing.assert_allclose(X.cpu().double().numpy(), np.linalg.solve(Lr, Bm), rtol=0,
                               atol=tol * 10 * np.abs(np.linalg.solve(Lr, Bm)).max())
    dense.trsm(L, X, trans=True)
    sol = np.linalg.solve(A, Bm)
    np.testing.assert_allclose(X.cpu().double().numpy(), sol, rtol=0, atol=tol * 10 * np.abs(sol).max())
    ld = dense.chol_logdet_half(L).item()
    assert abs(2 * ld - np.linalg.slogdet(A)[1]) < (1e-10 if f64 else 1e-3) * max(1.0, abs(ld))


def test_potrf_not_spd():
    import torch
    from rvgp_amd import dense
    A = np.eye(70)
    A[40, 40] = -1.0
    with pytest.raises(dense.CholeskyError):
        dense.potrf(_t(A, torch.float64))


@pytest.mark.parametrize("K,M", [(40000, 300), (5000, 576), (3000, 90), (257, 130)])
def test_gemm_sym_matches_full_product(K, M):
    """Upper-tile symmetric Gram: equals A^T B on the upper triangle and is
    exactly symmetric; f64 and f32-in/f64-accumulate, with and without split-K."""
    import torch
    from rvgp_amd import dense
    rng = np.random.default_rng(K + M)
    Y = rng.normal(size=(K, M))
    W = Y + 0.1 * rng.normal(size=(K, M))
    for A, B in ((Y, Y), (Y, W)):
        ref = A.T @ B
        # FP32 blocks: widened to FP64 on the FP64 matrix cores (exact
        # products, FP64 sums), or FP32 MFMA in <= 4096-row chunks summed in
        # FP64 (~sqrt(4096) * 6e-8 per chunk)
        for dt, f32mfma, tol in ((torch.float64, False, 1e-11), (torch.float32, False, 1e-10),
                                 (torch.float32, True, 1e-5)):
            if dt == torch.float32:
                A32, B32 = A.astype(np.float32).astype(np.float64), B.astype(np.float32).astype(np.float64)
                ref = A32.T @ B32
            C = dense.gemm_sym(_t(A, dt), _t(B, dt), fp32_mfma=f32mfma).cpu().numpy()
            iu = np.triu_indices(M)
            np.testing.assert_allclose(C[iu], ref[iu], rtol=0, atol=tol * np.abs(ref).max())
            assert np.array_equal(C, C.T)


def test_gemm_sym_f32_raised_split_matches_full_product():
    """The FP32 Gram's K split is raised past ceil(K/4096) chunks to fill
    whole rounds of resident workgroups (gram_f32_plan): at K = 65536,
    M = 576 the workspace query reports the raised split, and the product
    computed with it equals A^T B to the FP32-chunk tolerance."""
    import torch
    from rvgp_amd import _lib, dense
    K, M = 65536, 576
    ws = _lib.lib().rvgp_gemm_ws_bytes(dense.GEMM_F32, M, M, K, 0)
    assert ws > -(-K // 4096) * M * M * 4, ws
    rng = np.random.default_rng(3)
    Y = rng.normal(size=(K, M)).astype(np.float32)
    C = dense.gemm_sym(_t(Y, torch.float32), _t(Y, torch.float32)).cpu().numpy()
    Yd = Y.astype(np.float64)
    ref = Yd.T @ Yd
    np.testing.assert_allclose(C, ref, rtol=0, atol=1e-5 * np.abs(ref).max())
    assert np.array_equal(C, C.T)


@pytest.mark.parametrize("f64", [True, False])
@pytest.mark.parametrize("M,k,ld", [(1, 1, 1), (37, 5, 8), (3000, 300, 304), (70001, 64, 64)])
def test_ritz_residuals_match_the_formed_residual_block(M, k, ld, f64):
    """rvgp_ritz_residuals (one pass, f64 sums) against |AX - X diag(theta)|
    / |X| formed column by column in FP64 numpy; strided rows (ld > k)."""
    import torch
    from rvgp_amd import dense
    rng = np.random.default_rng(M + k)
    dt = torch.float64 if f64 else torch.float32
    Xs = rng.normal(size=(M, ld))
    AXs = rng.normal(size=(M, ld))
    theta = rng.uniform(0.0, 3.0, size=k)
    X, AX = _t(Xs, dt)[:, :k], _t(AXs, dt)[:, :k]
    Xh, AXh = X.cpu().double().numpy(), AX.cpu().double().numpy()
    ref = np.linalg.norm(AXh - Xh * theta, axis=0) / np.linalg.norm(Xh, axis=0)
    got = dense.ritz_residuals(X, AX, _t(theta, torch.float64)).cpu().numpy()
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=0)


@pytest.mark.parametrize("f64", [True, False])
@pytest.mark.parametrize("b", [1, 3])
@pytest.mark.parametrize("with_perm", [False, True])
def test_evecs_finish_normalises_scales_and_unpermutes(b, f64, with_perm):
    """rvgp_evecs_finish: out[perm[i] b + a] = scale X[i b + a] / |X[:, j]|,
    against the same map written with torch indexing in FP64 (a strided
    column slice as input, as the eigensolver passes it)."""
    import torch
    from rvgp_amd import dense
    nb, k, ld = 4001, 77, 96
    rng = np.random.default_rng(b * 10 + int(f64))
    dt = torch.float64 if f64 else torch.float32
    X = _t(rng.normal(size=(nb * b, ld)), dt)[:, :k]
    perm = torch.as_tensor(rng.permutation(nb), device="cuda") if with_perm else None
    scale = float(np.sqrt(nb * b))
    got = dense.evecs_finish(X, b, scale, perm)
    Xd = X.double()
    ref = Xd / torch.linalg.vector_norm(Xd, dim=0) * scale
    if with_perm:
        out = torch.empty_like(ref)
        out.view(nb, b, k)[perm] = ref.view(nb, b, k)
        ref = out
    np.testing.assert_allclose(got.cpu().numpy(), ref.cpu().numpy(), rtol=1e-13, atol=1e-13)
