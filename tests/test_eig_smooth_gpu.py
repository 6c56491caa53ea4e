"""Eigensolver (a8) and heat-kernel smoothing (a11) on the GPU vs FP64 oracles.

Eigenvalues: relative error <= 1e-6 (absolute floor 1e-9 * ub for ~0
eigenvalues); eigenspaces: largest principal angle <= 1e-4 rad over clusters
delimited by spectral gaps (pairs / multiplets are compared as subspaces,
never vector by vector).
"""
import numpy as np
import pytest
import scipy.linalg
import scipy.sparse as sp

from oracle import cpu
from tests.util import flat_torus3, sphere, torus

pytestmark = pytest.mark.gpu


def _bsr_from_blocks(rp, c, blocks, b):
    import torch
    from rvgp_amd.sparse import BSR
    return BSR(torch.as_tensor(rp, device="cuda"), torch.as_tensor(c, device="cuda"),
               torch.as_tensor(np.ascontiguousarray(blocks).reshape(-1), device="cuda"), None, b)


def ring_connection_laplacian(n, theta):
    """Cycle graph with self loops; transport R(theta/n) along i -> i+1."""
    rp = [0]
    cols, blocks = [], []
    phi = theta / n
    Rf = np.array([[np.cos(phi), -np.sin(phi)], [np.sin(phi), np.cos(phi)]])
    for i in range(n):
        nb = sorted({(i - 1) % n, i, (i + 1) % n})
        for j in nb:
            cols.append(j)
            if j == i:
                blocks.append(2.0 * np.eye(2))
            elif j == (i + 1) % n:
                blocks.append(-Rf)
            else:
                blocks.append(-Rf.T)
        rp.append(len(cols))
    return np.array(rp, np.int32), np.array(cols, np.int32), np.array(blocks)


def cluster_angles(w_ref, V_ref, V, k, rel_gap=1e-5):
    """Max principal angle over gap-delimited clusters fully inside [0, k)."""
    scale = max(abs(w_ref[k - 1]), 1e-12)
    worst = 0.0
    start = 0
    for j in range(1, k + 1):
        if j == k or (w_ref[j] - w_ref[j - 1]) > rel_gap * scale:
            if j == k and (k < len(w_ref)) and (w_ref[k] - w_ref[k - 1]) <= rel_gap * scale:
                break  # cluster straddles the cut at k
            ang = scipy.linalg.subspace_angles(V_ref[:, start:j], V[:, start:j]).max()
            worst = max(worst, ang)
            start = j
    return worst


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ring_closed_forms(precision):
    from rvgp_amd.eigsolve import chfsi
    n, theta, k = 300, 1.3, 24
    rp, c, blocks = ring_connection_laplacian(n, theta)
    A = _bsr_from_blocks(rp, c, blocks, 2)
    w, V, st = chfsi(A, k, precision=precision)
    # real 2n x 2n form: each 2 - 2cos((2 pi j + theta)/n), j in Z, appears twice
    j = np.arange(n)
    ref = 2 - 2 * np.cos((2 * np.pi * j + theta) / n)
    ref = np.sort(np.concatenate([ref, ref]))[:k]
    w = w.cpu().numpy()
    np.testing.assert_allclose(w, ref, rtol=1e-6, atol=1e-9 * 4)
    # scalar cycle: 2 - 2 cos(2 pi j / n)
    rp1 = rp
    A1 = _bsr_from_blocks(rp1, c, np.array([2.0 if (i == cc) else -1.0 for i, cc in
                                            zip(np.repeat(np.arange(n), np.diff(rp1)), c)]), 1)
    w1, V1, _ = chfsi(A1, 9, precision=precision)
    ref1 = np.sort(2 - 2 * np.cos(2 * np.pi * np.arange(n) / n))[:9]
    np.testing.assert_allclose(w1.cpu().numpy(), ref1, rtol=1e-6, atol=1e-9 * 4)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sphere_connection_laplacian_vs_eigh(precision):
    """C1-style operator (S^2, k=10, d=2) at n=1200: eigenvalues and
    cluster subspaces vs numpy.linalg.eigh of the dense FP64 matrix."""
    import torch
    from rvgp_amd.eigsolve import chfsi
    X = sphere(1200, 0)
    rp, c = cpu.sym_csr(cpu.knn(X, 10))
    T, S = cpu.tangent_frames(X, rp, c, 3, 15)
    R = cpu.connections(np.ascontiguousarray(T[:, :, :2]), rp, c)
    blocks = cpu.lc_blocks(rp, c, R)
    A = _bsr_from_blocks(rp, c, blocks, 2)
    Lc = cpu.connection_laplacian(rp, c, R)
    w_ref, V_ref = np.linalg.eigh(Lc.toarray())
    k = 50
    w, V, st = chfsi(A, k, precision=precision)
    w, V = w.cpu().numpy(), V.cpu().numpy()
    ub = A.gershgorin_bound()
    err = np.abs(w - w_ref[:k]) / np.maximum(np.abs(w_ref[:k]), 1e-3 * ub)
    assert err.max() < 1e-6, (err.max(), st)
    assert cluster_angles(w_ref, V_ref, V, k) < 1e-4
    assert (st["half_passes"] >= 1) == (precision == "mixed"), st
    # scalar Laplacian of the same graph
    L = cpu.laplacian(rp, c)
    rows = np.repeat(np.arange(1200), np.diff(rp))
    A1 = _bsr_from_blocks(rp, c, np.asarray(L[rows, c]).ravel(), 1)
    w1, V1, _ = chfsi(A1, k, precision=precision)
    wr1, Vr1 = np.linalg.eigh(L.toarray())
    w1 = w1.cpu().numpy()
    assert abs(w1[0]) < 1e-9 * ub
    np.testing.assert_allclose(w1[1:], wr1[1:k], rtol=1e-6)
    assert cluster_angles(wr1, Vr1, V1.cpu().numpy(), k) < 1e-4


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_flat_3torus_connection_laplacian_b3_vs_eigh(precision):
    """The headline operator's kind (3-torus in R^32, k = 20, K = 30 < D,
    dim_man = 3: 3 x 3 blocks) at n = 700: eigenvalues and cluster subspaces
    of the b = 3 block solver vs numpy.linalg.eigh of the dense FP64 matrix."""
    from rvgp_amd.eigsolve import chfsi
    X = flat_torus3(700, 2)
    rp, c = cpu.sym_csr(cpu.knn(X, 20))
    T, S = cpu.tangent_frames(X, rp, c, 32, 30, allow_rank_deficient_k=True)
    assert cpu.manifold_dimension(S, 0.8) == 3
    R = cpu.connections(np.ascontiguousarray(T[:, :, :3]), rp, c)
    A = _bsr_from_blocks(rp, c, cpu.lc_blocks(rp, c, R), 3)
    Lc = cpu.connection_laplacian(rp, c, R)
    w_ref, V_ref = np.linalg.eigh(Lc.toarray())
    # a cut at a spectral gap (the first k eigenvectors then span one space)
    gaps = np.diff(w_ref[50:70])
    k = 51 + int(np.argmax(gaps))
    w, V, st = chfsi(A, k, precision=precision)
    w, V = w.cpu().numpy(), V.cpu().numpy()
    ub = A.gershgorin_bound()
    err = np.abs(w - w_ref[:k]) / np.maximum(np.abs(w_ref[:k]), 1e-9 * ub)
    assert err.max() < 1e-6, (err.max(), st)
    assert cluster_angles(w_ref, V_ref, V, k) < 1e-4
    # the mixed form's opening filter passes ran on binary16 blocks
    assert (st["half_passes"] >= 1) == (precision == "mixed"), st


def test_eig_rejects_k_ge_n_minus_1():
    from rvgp_amd.eigsolve import chfsi
    rp, c, blocks = ring_connection_laplacian(10, 0.5)
    A = _bsr_from_blocks(rp, c, blocks, 2)
    with pytest.raises(TypeError):
        chfsi(A, 19)


# ---------------------------------------------------------------- smoothing
@pytest.mark.parametrize("t", [0.3, 5.0, 100.0])
def test_expm_multiply_vs_scipy_expm(t):
    import torch
    from rvgp_amd.smoothing import expm_multiply
    X = sphere(250, 3)
    rp, c = cpu.sym_csr(cpu.knn(X, 8))
    T, S = cpu.tangent_frames(X, rp, c, 3, 12)
    R = cpu.connections(np.ascontiguousarray(T[:, :, :2]), rp, c)
    A = _bsr_from_blocks(rp, c, cpu.lc_blocks(rp, c, R), 2)
    Lc = cpu.connection_laplacian(rp, c, R)
    x = np.random.default_rng(1).normal(size=(500, 3))
    ref = scipy.linalg.expm(-t * Lc.toarray()) @ x
    out = expm_multiply(A, torch.as_tensor(x, device="cuda"), t).cpu().numpy()
    assert np.abs(out - ref).max() < 1e-12 * max(1.0, np.abs(x).max()) * 50


def test_heat_kernel_preserves_constants():
    """exp(-tL) 1 = 1 for a graph Laplacian (row sums zero), any t."""
    import torch
    from rvgp_amd.smoothing import expm_multiply
    X = torus(5000, 2)
    rp, c = cpu.sym_csr(cpu.knn(X, 15))
    L = cpu.laplacian(rp, c)
    rows = np.repeat(np.arange(5000), np.diff(rp))
    A = _bsr_from_blocks(rp, c, np.asarray(L[rows, c]).ravel(), 1)
    one = torch.ones(5000, dtype=torch.float64, device="cuda")
    for t in (1.0, 50.0, 500.0):
        out = expm_multiply(A, one, t).cpu().numpy()
        assert np.abs(out - 1.0).max() < 1e-11


def test_vector_diffusion_vs_oracle():
    """smoothing.vector_diffusion (normalised) vs the dense-expm oracle."""
    import torch
    from rvgp_amd.smoothing import vector_diffusion_device
    X = sphere(300, 4)
    rp, c = cpu.sym_csr(cpu.knn(X, 10))
    T, S = cpu.tangent_frames(X, rp, c, 3, 15)
    T2 = np.ascontiguousarray(T[:, :, :2])
    R = cpu.connections(T2, rp, c)
    Lc = cpu.connection_laplacian(rp, c, R)
    L = cpu.laplacian(rp, c)
    A = _bsr_from_blocks(rp, c, cpu.lc_blocks(rp, c, R), 2)
    rows = np.repeat(np.arange(300), np.diff(rp))
    A1 = _bsr_from_blocks(rp, c, np.asarray(L[rows, c]).ravel(), 1)
    v = np.random.default_rng(5).normal(size=(300, 3))
    cl = cpu.express_in_local_frame(v, T2)
    for t in (1.0, 100.0):
        ref = cpu.vector_diffusion(cl, t, Lc, L, "expm")
        out = vector_diffusion_device(torch.as_tensor(cl, device="cuda"), t, A, A1).cpu().numpy()
        # the renormalisation divides by |exp(-t Lc) c|, which has decayed by
        # amp = |c| / |exp(-t Lc) c|; FP64 rounding is amplified by that factor
        raw = cpu.heat(Lc, t, cl.reshape(-1, 1), "eigh")
        amp = np.linalg.norm(cl) / np.linalg.norm(raw)
        assert np.abs(out - ref).max() < 1e-13 * amp * np.abs(ref).max()
    # batched fields = one multi-column pass
    cl2 = np.stack([cl, 2 * cl[::-1].copy()])
    outb = vector_diffusion_device(torch.as_tensor(cl2, device="cuda"), 3.0, A, A1).cpu().numpy()
    for f in range(2):
        ref = cpu.vector_diffusion(cl2[f], 3.0, Lc, L, "expm")
        assert np.abs(outb[f] - ref).max() < 1e-10 * np.abs(ref).max()


def test_vector_diffusion_per_field_times_vs_oracle():
    """One diffusion time per field (the batched-fits config draws them
    log-uniformly): one shared Chebyshev recurrence, per-column coefficients,
    each field equal to the dense-expm oracle at its own time."""
    import torch
    from rvgp_amd.smoothing import vector_diffusion_device
    X = sphere(250, 6)
    rp, c = cpu.sym_csr(cpu.knn(X, 10))
    T, S = cpu.tangent_frames(X, rp, c, 3, 15)
    T2 = np.ascontiguousarray(T[:, :, :2])
    R = cpu.connections(T2, rp, c)
    Lc = cpu.connection_laplacian(rp, c, R)
    L = cpu.laplacian(rp, c)
    A = _bsr_from_blocks(rp, c, cpu.lc_blocks(rp, c, R), 2)
    rows = np.repeat(np.arange(250), np.diff(rp))
    A1 = _bsr_from_blocks(rp, c, np.asarray(L[rows, c]).ravel(), 1)
    rng = np.random.default_rng(8)
    ts = np.array([0.5, 3.0, 12.0, 40.0])
    cl = np.stack([cpu.express_in_local_frame(rng.normal(size=(250, 3)), T2) for _ in ts])
    out = vector_diffusion_device(torch.as_tensor(cl, device="cuda"), ts, A, A1).cpu().numpy()
    for f, t in enumerate(ts):
        ref = cpu.vector_diffusion(cl[f], t, Lc, L, "expm")
        raw = cpu.heat(Lc, t, cl[f].reshape(-1, 1), "eigh")
        amp = np.linalg.norm(cl[f]) / np.linalg.norm(raw)
        assert np.abs(out[f] - ref).max() < 1e-12 * amp * np.abs(ref).max(), (f, t)


def test_spectral_upper_bound_brackets_lambda_max():
    """The Lanczos bound used by the filters and the heat-kernel series lies
    between lambda_max (dense eigh) and the Gershgorin bound, and within 3 %
    of lambda_max (2 % margin)."""
    X = sphere(400, 11)
    rp, c = cpu.sym_csr(cpu.knn(X, 10))
    T, S = cpu.tangent_frames(X, rp, c, 3, 15)
    T2 = np.ascontiguousarray(T[:, :, :2])
    R = cpu.connections(T2, rp, c)
    L = cpu.laplacian(rp, c)
    rows = np.repeat(np.arange(400), np.diff(rp))
    for A, dense_A in ((_bsr_from_blocks(rp, c, cpu.lc_blocks(rp, c, R), 2),
                        cpu.connection_laplacian(rp, c, R).toarray()),
                       (_bsr_from_blocks(rp, c, np.asarray(L[rows, c]).ravel(), 1),
                        L.toarray())):
        lmax = np.linalg.eigvalsh(dense_A)[-1]
        ub = A.spectral_upper_bound()
        assert lmax <= ub <= A.gershgorin_bound() + 1e-12
        assert ub <= 1.03 * lmax, (ub, lmax)


def test_spectral_smoothing_vs_oracle():
    """smooth_vector_field(method='spectral') and the drop-in
    smoothing.vector_diffusion(method='spectral') (smoothing.py:17-64): the
    heat kernel restricted to the computed eigenpairs, against the oracle's
    restatement on its own eigenpairs (cut at a spectral gap, so the
    truncated kernel does not depend on the basis); gauge-invariant ambient
    output."""
    import rvgp_amd as RVGP
    from rvgp_amd import smoothing
    X = sphere(500, 8)
    o = cpu.Data(X, n_neighbors=10, n_eigenpairs=80, dense_eigh=True)
    w, wl = o.evals_Lc, o.evals_L
    # a cut where both spectra have a gap (the sphere's L has near-multiplets)
    gaps = np.minimum(np.diff(w[40:71]) / w[41:71], np.diff(wl[40:71]) / wl[41:71])
    k = 41 + int(np.argmax(gaps))
    kl = 15 + int(np.argmax(np.diff(wl[15:40])))
    d = RVGP.create_data_object(X, n_neighbors=10, n_eigenpairs=k, verbose=False)
    d.random_vector_field(seed=2)
    o.random_vector_field(seed=2)
    v0 = o.vectors.copy()
    # the data object's own spectra: k for Lc, and k for L
    d.smooth_vector_field(t=3.0, method="spectral")
    c = cpu.express_in_local_frame(v0, o.gauges)
    U, Ul = o.U[:, :k], o.evecs_L[:, :k]
    ref = cpu.vector_diffusion(c, 3.0, (w[:k], U), (wl[:k], Ul), method="spectral")
    ref = cpu.express_in_local_frame(ref, o.gauges, reverse=True)
    assert wl[k] - wl[k - 1] > 1e-3 * wl[k] and w[k] - w[k - 1] > 1e-3 * w[k]
    np.testing.assert_allclose(d.vectors, ref, rtol=0, atol=1e-8 * np.abs(ref).max())
    # drop-in function on host eigenpairs (ambient rows and the scalar pairs)
    out = smoothing.vector_diffusion(c, 3.0, (w[:k], U), (wl[:kl], o.evecs_L[:, :kl]))
    ref2 = cpu.vector_diffusion(c, 3.0, (w[:k], U), (wl[:kl], o.evecs_L[:, :kl]), method="spectral")
    np.testing.assert_allclose(out, ref2, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("dtype", ["float64", "float32", "float16"])
def test_per_column_degree_filter_equals_uniform_filters_per_group(dtype):
    """_cheb_filter_cols with degrees (4, 4, ..., 9, ..., 13) per group of 8
    columns equals, group by group, the one-degree filter _cheb_filter of
    that group's degree on that group alone (the recurrence is column-wise;
    the parity bookkeeping of the two iterate buffers is what is checked),
    on the b = 2 ring connection Laplacian."""
    import torch
    from rvgp_amd.eigsolve import _cheb_filter, _cheb_filter_cols
    n, theta = 400, 0.7
    rp, c, blocks = ring_connection_laplacian(n, theta)
    A = _bsr_from_blocks(rp, c, blocks, 2)
    dt = getattr(torch, dtype)
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    X0 = torch.randn((2 * n, 48), generator=g, device="cuda", dtype=torch.float64).to(dt)
    ms = np.repeat([4, 4, 9, 10, 10, 13], 8)
    a, b, a0 = 0.05, 4.0, 0.0
    Y = _cheb_filter_cols(A, X0.clone(), ms, a, b, a0)
    for grp in range(6):
        sl = slice(8 * grp, 8 * grp + 8)
        ref = _cheb_filter(A, X0[:, sl].contiguous(), int(ms[8 * grp]), a, b, a0)
        tol = {"float64": 1e-12, "float32": 1e-5, "float16": 2e-3}[dtype]
        err = float(((Y[:, sl] - ref).abs().max() / ref.abs().max()).item())
        assert err < tol, (grp, err)



def test_binary16_filter_pass_is_the_fp32_filter_up_to_column_scale():
    """_filter_pass(half=True) scales each column to unit RMS, stores the block
    as binary16 and normalises the Chebyshev filter at 0 instead of a0: the
    same polynomial of A up to a constant per column.  So every output column
    must be parallel to the FP32 pass's column (sine of the angle within a few
    binary16 roundings per step: 2e-2 at degree 13), on the b = 2 ring with
    a0 = 0.5 above the lowest eigenvalues, where the FP32 filter amplifies."""
    import torch
    from rvgp_amd.eigsolve import _filter_pass
    n, theta = 400, 0.7
    rp, c, blocks = ring_connection_laplacian(n, theta)
    A = _bsr_from_blocks(rp, c, blocks, 2)
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    X = torch.randn((2 * n, 48), generator=g, device="cuda", dtype=torch.float64).float()
    ms = np.repeat([4, 4, 9, 10, 10, 13], 8)
    a, b, a0 = 0.05, 4.0, 0.5
    Y32 = _filter_pass(A, X.clone(), ms, 13, a, b, a0, half=False).double()
    Y16 = _filter_pass(A, X.clone(), ms, 13, a, b, a0, half=True)
    assert Y16.dtype == torch.float32 and torch.isfinite(Y16).all()
    Y16 = Y16.double()
    cos = (Y32 * Y16).sum(0).abs() / (Y32.norm(dim=0) * Y16.norm(dim=0))
    sin = torch.sqrt(torch.clamp(1.0 - cos * cos, min=0.0))
    assert float(sin.max()) < 2e-2, sin.max()
    # and the FP32 pass itself grows the directions below a0 (what the
    # normalisation at 0 keeps out of binary16's range)
    assert float(Y32.norm(dim=0).max()) > float(X.double().norm(dim=0).max())
