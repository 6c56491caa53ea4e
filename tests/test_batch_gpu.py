"""Batched independent fits (one launch per optimiser step for all fields)
vs the oracle and vs the single-fit engine.

Each fit f has its own node subset (sizes from 1 node to most of the
manifold) and its own hyperparameters; the oracle is the textbook
function-space GPR (oracle/cpu.py) on the ambient eigenvector rows of the
fit's nodes, so the weight-space statistics gathered from intrinsic rows, the
fused Cholesky / L^-1 step, the adjoint and Adam are all checked against
independent arithmetic.
"""
import numpy as np
import pytest

from oracle import cpu

pytestmark = pytest.mark.gpu


def sphere_data(n=700, k_eig=40, seed=0):
    import rvgp_amd as RVGP
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 3))
    X /= np.linalg.norm(X, axis=1, keepdims=True)
    return RVGP.create_data_object(X, n_neighbors=10, n_eigenpairs=k_eig, verbose=False)


def torus3_data(n=1200, k_eig=45, seed=1):
    import rvgp_amd as RVGP
    rng = np.random.default_rng(seed)
    a = rng.uniform(0, 2 * np.pi, (n, 3))
    X = np.concatenate([np.cos(a), np.sin(a)], 1)
    return RVGP.create_data_object(X, n_neighbors=12, n_eigenpairs=k_eig, verbose=False,
                                   reference_semantics=False)


@pytest.fixture(scope="module")
def sphere():
    return sphere_data()


@pytest.fixture(scope="module")
def torus3():
    return torus3_data()


def fields_and_sets(d, F, sizes, seed=3):
    import torch
    rng = np.random.default_rng(seed)
    V = rng.normal(size=(F, d.n, d._X.shape[1]))
    Vd = torch.as_tensor(V, device="cuda")
    Vd = d._frame(2, Vd, V.shape[2])  # tangent fields
    sets = [rng.choice(d.n, size=s, replace=False) for s in sizes]
    thetas = np.stack([rng.uniform(1.0, 2.5, F), rng.uniform(2.0, 6.0, F),
                       rng.uniform(0.5, 2.0, F), rng.uniform(0.01, 0.5, F)], 1)
    return Vd, sets, thetas


def oracle_fit(d, Vd, nodes, theta):
    import torch
    Phi = d.phi_rows(torch.as_tensor(nodes.astype(np.int32), device="cuda")).cpu().numpy()
    y = Vd[:, torch.as_tensor(nodes, device="cuda").long()].cpu().numpy()
    lam = d.spectrum_device().cpu().numpy()
    D = Vd.shape[2]
    S, dS = cpu.spectral_filter(lam, theta[0], theta[1], theta[2], d.n * D)
    return Phi, y, S, dS


@pytest.mark.parametrize("tile", [32, 64, 128])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_batch_statistics_equal_ambient_products(sphere, tile, prec):
    """G_f = Phi_f^T Phi_f, b_f = Phi_f^T y_f, yy_f = |y_f|^2 over the fit's
    AMBIENT rows, although the kernels read only intrinsic rows."""
    import torch
    from rvgp_amd.batch import BatchedGPR
    d = sphere
    Vd, sets, th = fields_and_sets(d, 5, [1, 9, 64, 200, 650])
    gp = BatchedGPR(d, Vd, sets, theta0=th, gram_precision=prec, gram_tile=tile, panel=32)
    for f, s in enumerate(sets):
        Phi = d.phi_rows(torch.as_tensor(s.astype(np.int32), device="cuda"))
        y = Vd[f, torch.as_tensor(s, device="cuda").long()].reshape(-1, 1)
        G = (Phi.T @ Phi).cpu().numpy()
        rtol = 1e-12 if prec == "fp64" else 2e-6
        np.testing.assert_allclose(gp.G[f].cpu().numpy(), G, rtol=rtol,
                                   atol=rtol * np.abs(G).max())
        b = (Phi.T @ y).cpu().numpy().ravel()
        np.testing.assert_allclose(gp.b[f].cpu().numpy(), b, rtol=1e-10,
                                   atol=1e-11 * np.abs(b).max())
        assert abs(gp.yy[f].item() - float((y * y).sum())) < 1e-12 * float((y * y).sum())


@pytest.mark.parametrize("panel", [16, 32, (16, 1024), (32, 1024)])
@pytest.mark.parametrize("which", ["sphere", "torus3"])
def test_batch_nlml_and_adjoint_vs_oracle(request, panel, which):
    """One launch evaluates every fit: NLML and gradient w.r.t. (nu, kappa,
    sigma_f, noise) vs the textbook FP64 GPR on the fit's ambient rows."""
    from rvgp_amd.batch import BatchedGPR
    d = request.getfixturevalue(which)
    sizes = [1, 7, 33, 120, 400] if which == "sphere" else [2, 50, 333, 900]
    Vd, sets, th = fields_and_sets(d, len(sizes), sizes, seed=5)
    gp = BatchedGPR(d, Vd, sets, theta0=th, panel=panel, gram_tile=64)
    out = gp.evaluate().cpu().numpy()
    gp.check_status()
    thh = gp.theta_host()
    for f, s in enumerate(sets):
        Phi, y, S, dS = oracle_fit(d, Vd, s, thh[f])
        yf = y[f].reshape(-1, 1)
        ref = cpu.gpr_nlml(Phi, yf, S, thh[f, 3])
        gS, gn = cpu.gpr_grad(Phi, yf, S, thh[f, 3])
        g = np.array([dS[0] @ gS, dS[1] @ gS, dS[2] @ gS, gn])
        assert abs(out[f, 0] - ref) < 1e-9 * max(1.0, abs(ref)), (f, out[f, 0], ref)
        np.testing.assert_allclose(out[f, 1:5], g, rtol=1e-7, atol=1e-8 * np.abs(g).max())


def test_batch_adam_matches_single_fit_engine(sphere):
    """12 graph-replayed batched Adam steps == 12 Adam steps of the
    single-fit weight-space engine, fit by fit."""
    import torch
    from rvgp_amd.batch import BatchedGPR
    from rvgp_amd.gp import GPREngine
    d = sphere
    Vd, sets, th = fields_and_sets(d, 4, [5, 40, 150, 500], seed=7)
    gp = BatchedGPR(d, Vd, sets, theta0=th, gram_tile=64, panel=32)
    gp.fit(epochs=12, lr=0.05, use_graph=True)
    gp.check_status()
    thb = gp.theta_host()
    lower = gp.lower.cpu().numpy()
    for f, s in enumerate(sets):
        Phi = d.phi_rows(torch.as_tensor(s.astype(np.int32), device="cuda"))
        y = Vd[f, torch.as_tensor(s, device="cuda").long()].reshape(-1)
        e = GPREngine(Phi, y, d.spectrum_device(), d.n * 3, form="weight", lower=lower)
        e.set_theta(th[f])
        e.adam_setup()
        for _ in range(12):
            e.adam_step(lr=0.05)
        np.testing.assert_allclose(thb[f], e.theta_host(), rtol=1e-7)


def test_batch_graph_replays_match_eager(sphere):
    from rvgp_amd.batch import BatchedGPR
    d = sphere
    Vd, sets, th = fields_and_sets(d, 3, [20, 100, 300], seed=9)
    a = BatchedGPR(d, Vd, sets, theta0=th, gram_tile=64, panel=16)
    b = BatchedGPR(d, Vd, sets, theta0=th, gram_tile=64, panel=16)
    for _ in range(6):
        a.adam_step(0.05)
    g = b.capture_adam_step(0.05)
    for _ in range(6):
        g.replay()
        b.evaluate()  # eager work between the replays
    np.testing.assert_array_equal(a.u.cpu().numpy(), b.u.cpu().numpy())
    np.testing.assert_array_equal(a.t.cpu().numpy(), b.t.cpu().numpy())


def test_fit_batched_and_transform_vs_oracle(torus3):
    """The drop-in entry point: fit, then predict held-out nodes of one fit
    vs the oracle's predictive equations at the fitted hyperparameters."""
    import rvgp_amd as RVGP
    d = torus3
    Vd, sets, _ = fields_and_sets(d, 6, [30, 80, 150, 220, 400, 700], seed=11)
    gp = RVGP.fit_batched(d, sets, vectors=Vd, epochs=25, lr=0.05)
    gp.check_status()
    th = gp.theta_host()
    assert np.all(np.isfinite(th))
    f = 4
    test = np.setdiff1d(np.arange(d.n), sets[f])[:150]
    mean, var = gp.transform(d, f, test)
    Phi, y, S, _ = oracle_fit(d, Vd, sets[f], th[f])
    import torch
    Ps = d.phi_rows(torch.as_tensor(test.astype(np.int32), device="cuda")).cpu().numpy()
    mr, vr = cpu.gpr_predict(Phi, y[f].reshape(-1, 1), S, th[f, 3], Ps)
    assert mean.shape == (len(test), Vd.shape[2])
    assert np.linalg.norm(mean.ravel() - mr.ravel()) / np.linalg.norm(mr) < 1e-7
    np.testing.assert_allclose(var.ravel(), vr.ravel(), rtol=1e-6, atol=1e-10)
    # the NLML of every fit went down from the initial hyperparameters
    gp0 = RVGP.fit_batched(d, sets, vectors=Vd, epochs=0)
    assert np.all(gp.training_loss() < gp0.training_loss())


def test_batch_persistent_grid_smaller_than_fit_count(sphere):
    """With fewer workgroups than fits each workgroup loops over fits: the
    result equals the one-workgroup-per-fit launch."""
    from rvgp_amd.batch import BatchedGPR
    d = sphere
    Vd, sets, th = fields_and_sets(d, 5, [3, 30, 90, 200, 410], seed=13)
    a = BatchedGPR(d, Vd, sets, theta0=th, gram_tile=64, panel=32)
    b = BatchedGPR(d, Vd, sets, theta0=th, gram_tile=64, panel=32, grid=2)
    a.fit(epochs=4, use_graph=False)
    b.fit(epochs=4, use_graph=False)
    np.testing.assert_array_equal(a.out.cpu().numpy(), b.out.cpu().numpy())
    np.testing.assert_array_equal(a.u.cpu().numpy(), b.u.cpu().numpy())
