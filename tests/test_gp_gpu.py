"""GP layer on the GPU (a12, a14, a16) vs the FP64 textbook oracle.

The spectral filter, NLML, its adjoint gradient and the predictive
mean / variance are checked at fixed hyperparameters on the same Phi, y;
function-space and weight-space forms must agree with each other; the
gradient is also checked against central finite differences of the GPU NLML.
"""
import numpy as np
import pytest

from oracle import cpu

pytestmark = pytest.mark.gpu


def problem(M_nodes=300, D=3, k=40, seed=0):
    rng = np.random.default_rng(seed)
    lam = np.sort(rng.uniform(0.0, 5.0, size=k))
    lam[0] = 0.0
    Phi = rng.normal(size=(M_nodes * D, k))
    y = rng.normal(size=(M_nodes * D, 1))
    return Phi, y, lam


def engine(Phi, y, lam, n_rows, form, precision="fp64", theta=(1.5, 5.0, 1.0, 0.3)):
    import torch
    from rvgp_amd.gp import GPREngine
    e = GPREngine(torch.as_tensor(Phi, device="cuda"), torch.as_tensor(y, device="cuda"),
                  torch.as_tensor(lam, device="cuda"), n_rows, form=form, precision=precision)
    e.set_theta(theta)
    return e


@pytest.mark.parametrize("typ", ["matern", "se"])
def test_spectral_filter(typ):
    import torch
    from rvgp_amd import _lib
    lam = np.linspace(0, 7, 60)
    th = (1.3, 2.7, 0.8)
    S = torch.empty(60, dtype=torch.float64, device="cuda")
    dS = torch.empty(180, dtype=torch.float64, device="cuda")
    thd = torch.tensor(th, dtype=torch.float64, device="cuda")
    _lib.call("rvgp_spectral_filter", 0 if typ == "matern" else 1, 60,
              _lib.ptr(torch.as_tensor(lam, device="cuda")), _lib.ptr(thd), 1234.0, _lib.ptr(S),
              _lib.ptr(dS), _lib.stream())
    S_ref, dS_ref = cpu.spectral_filter(lam, *th, 1234.0, typ)
    np.testing.assert_allclose(S.cpu().numpy(), S_ref, rtol=1e-13)
    np.testing.assert_allclose(dS.cpu().numpy().reshape(3, 60), dS_ref, rtol=1e-11, atol=1e-13)
    assert abs(S.sum().item() - 1234.0 * th[2]) < 1e-9  # normalisation (kernels.py:46)


@pytest.mark.parametrize("form", ["function", "weight"])
def test_nlml_and_gradient_vs_oracle(form):
    Phi, y, lam = problem()
    n_rows = 2000.0
    th = (1.5, 5.0, 1.0, 0.3)
    e = engine(Phi, y, lam, n_rows, form, theta=th)
    out = e.evaluate().cpu().numpy()
    S, dS = cpu.spectral_filter(lam, th[0], th[1], th[2], n_rows)
    ref = cpu.gpr_nlml(Phi, y, S, th[3])
    gS, gn = cpu.gpr_grad(Phi, y, S, th[3])
    gref = np.array([dS[0] @ gS, dS[1] @ gS, dS[2] @ gS, gn])
    assert abs(out[0] - ref) < 1e-10 * abs(ref)
    np.testing.assert_allclose(out[1:5], gref, rtol=1e-8, atol=1e-9 * np.abs(gref).max())


def test_function_and_weight_space_agree():
    Phi, y, lam = problem(M_nodes=400, k=60, seed=3)
    a = engine(Phi, y, lam, 5000.0, "function")
    b = engine(Phi, y, lam, 5000.0, "weight")
    oa, ob = a.evaluate().cpu().numpy(), b.evaluate().cpu().numpy()
    np.testing.assert_allclose(oa, ob, rtol=1e-9)
    Ps = np.random.default_rng(9).normal(size=(50, 60))
    import torch
    ma, va = a.predict(torch.as_tensor(Ps, device="cuda"))
    mb, vb = b.predict(torch.as_tensor(Ps, device="cuda"))
    # FP64 on both sides; K_y has condition ~1e5-1e6 here, so the two
    # algebraically equal forms agree to ~1e-8 relative, not to 1e-15
    ma, mb = ma.cpu().numpy(), mb.cpu().numpy()
    assert np.linalg.norm(ma - mb) / np.linalg.norm(ma) < 1e-7
    np.testing.assert_allclose(va.cpu().numpy(), vb.cpu().numpy(), rtol=1e-7, atol=1e-12)


@pytest.mark.parametrize("form", ["function", "weight"])
def test_gradient_vs_finite_differences(form):
    Phi, y, lam = problem(M_nodes=200, k=30, seed=5)
    e = engine(Phi, y, lam, 900.0, form)
    u0 = e.u.cpu().numpy().copy()
    _, g, _ = e.nlml_grad_u(u0)
    # central differences with h = 1e-4: truncation ~1e-8, rounding noise of an
    # O(1e3) NLML ~1e-12 relative -> ~1e-5 absolute; the exact adjoint is
    # pinned to 1e-8 against the oracle's closed form elsewhere
    h = 1e-4
    for p in range(4):
        up, um = u0.copy(), u0.copy()
        up[p] += h
        um[p] -= h
        fd = (e.nlml_grad_u(up)[0] - e.nlml_grad_u(um)[0]) / (2 * h)
        assert abs(fd - g[p]) < 2e-5 * max(1.0, abs(fd)), (p, fd, g[p])


def test_predict_vs_oracle():
    import torch
    Phi, y, lam = problem(M_nodes=250, k=35, seed=7)
    th = (2.0, 3.0, 1.5, 0.05)
    e = engine(Phi, y, lam, 777.0, "function", theta=th)
    Ps = np.random.default_rng(1).normal(size=(40, 35))
    m, v = e.predict(torch.as_tensor(Ps, device="cuda"))
    S, _ = cpu.spectral_filter(lam, th[0], th[1], th[2], 777.0)
    mr, vr = cpu.gpr_predict(Phi, y, S, th[3], Ps)
    # FP64 both sides; condition of K_y ~1e6 bounds the agreement
    m = m.cpu().numpy()
    assert np.linalg.norm(m - mr.ravel()) / np.linalg.norm(mr) < 1e-7
    np.testing.assert_allclose(v.cpu().numpy(), vr.ravel(), rtol=1e-6, atol=1e-10)


def test_mixed_precision_function_space_within_1e4():
    """Mixed function-space form (FP32-tile Gram accumulated in FP64, Woodbury
    alpha refined against the FP64 K_y) vs the FP64 form."""
    import torch
    Phi, y, lam = problem(M_nodes=500, k=50, seed=11)
    a = engine(Phi, y, lam, 3000.0, "function")
    b = engine(Phi, y, lam, 3000.0, "function", precision="mixed")
    oa, ob = a.evaluate().cpu().numpy(), b.evaluate().cpu().numpy()
    assert abs(oa[0] - ob[0]) < 1e-8 * abs(oa[0])
    np.testing.assert_allclose(ob[1:5], oa[1:5], rtol=1e-4, atol=1e-6 * np.abs(oa[1:5]).max())
    Ps = torch.as_tensor(np.random.default_rng(2).normal(size=(30, 50)), device="cuda")
    ma, _ = a.predict(Ps)
    mb, _ = b.predict(Ps)
    assert (torch.linalg.norm(ma - mb) / torch.linalg.norm(ma)).item() < 1e-4


def test_adam_step_is_device_resident_and_decreases_nlml():
    Phi, y, lam = problem(M_nodes=200, k=30, seed=13)
    e = engine(Phi, y, lam, 900.0, "weight")
    f0 = e.nlml_grad_u(e.u.cpu().numpy())[0]
    e.adam_setup()
    for _ in range(50):
        e.adam_step(lr=0.05)
    f1 = e.nlml_grad_u(e.u.cpu().numpy())[0]
    assert f1 < f0


def test_lbfgs_reaches_oracle_optimum():
    Phi, y, lam = problem(M_nodes=150, k=25, seed=17)
    e = engine(Phi, y, lam, 600.0, "function", theta=(1.5, 5.0, 1.0, 1.0))
    res = e.fit_lbfgs(maxiter=200)
    th = e.theta_host()
    th_ref, res_ref = cpu.fit_lbfgs(Phi, y, lam, 600.0, maxiter=200)
    assert abs(res.fun - res_ref.fun) < 1e-6 * abs(res_ref.fun)
    np.testing.assert_allclose(th, th_ref, rtol=1e-3)


@pytest.mark.parametrize("form", ["weight", "function"])
def test_adam_graph_replays_match_eager(form):
    """Every replay of the captured optimiser step equals the eager step,
    with eager work and synchronisations between the replays."""
    import torch
    Phi, y, lam = problem(M_nodes=120, k=20, seed=19)
    a = engine(Phi, y, lam, 500.0, form)
    b = engine(Phi, y, lam, 500.0, form)
    a.adam_setup()
    b.adam_setup()
    g = b.capture_adam_step(lr=0.05)
    for step in range(12):
        a.adam_step(lr=0.05)
        g.replay()
        torch.cuda.synchronize()
        _ = (torch.ones(1000, device="cuda") * step).sum().item()  # unrelated eager work
        np.testing.assert_allclose(b.u.cpu().numpy(), a.u.cpu().numpy(), rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(b.out.cpu().numpy(), a.out.cpu().numpy(), rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("form,precision", [("function", "fp64"), ("weight", "fp64"),
                                            ("function", "mixed"), ("weight", "mixed")])
def test_multi_rhs_vs_oracle(form, precision):
    """R = 9 right-hand sides sharing the hyperparameters (the multi-RHS
    config): summed NLML, its gradient and the per-column predictive means vs
    the oracle; the variance does not depend on Y."""
    import torch
    rng = np.random.default_rng(21)
    Phi, _, lam = problem(M_nodes=64, k=30, seed=21)
    Y = rng.normal(size=(Phi.shape[0], 9))
    th = (1.7, 4.0, 1.2, 0.08)
    e = engine(Phi, Y, lam, 1500.0, form, precision=precision, theta=th)
    out = e.evaluate().cpu().numpy()
    S, dS = cpu.spectral_filter(lam, th[0], th[1], th[2], 1500.0)
    ref = cpu.gpr_nlml(Phi, Y, S, th[3])
    gS, gn = cpu.gpr_grad(Phi, Y, S, th[3])
    gref = np.array([dS[0] @ gS, dS[1] @ gS, dS[2] @ gS, gn])
    tol = 1e-10 if precision == "fp64" else 1e-8
    assert abs(out[0] - ref) < tol * abs(ref)
    np.testing.assert_allclose(out[1:5], gref, rtol=1e-7 if precision == "fp64" else 1e-5,
                               atol=1e-9 * np.abs(gref).max())
    Ps = rng.normal(size=(25, 30))
    m, v = e.predict(torch.as_tensor(Ps, device="cuda"))
    assert tuple(m.shape) == (25, 9)
    for r in (0, 4, 8):
        mr, vr = cpu.gpr_predict(Phi, Y[:, r:r + 1], S, th[3], Ps)
        mc = m[:, r].cpu().numpy()
        assert np.linalg.norm(mc - mr.ravel()) / np.linalg.norm(mr) < 1e-6
    np.testing.assert_allclose(v.cpu().numpy(), vr.ravel(), rtol=1e-6, atol=1e-10)


@pytest.mark.parametrize("form", ["function", "weight"])
def test_autograd_function_uses_the_adjoint(form):
    """NLMLFunction: the value is the engine's NLML, torch.autograd returns
    the adjoint gradient (constrained and unconstrained parameters), and the
    backward passes torch's own finite-difference gradcheck."""
    import torch
    Phi, y, lam = problem(M_nodes=80, k=20, seed=13)
    th = (1.4, 3.5, 0.9, 0.2)
    e = engine(Phi, y, lam, 700.0, form, theta=th)
    ref = e.evaluate().cpu().numpy().copy()
    x = torch.tensor(th, dtype=torch.float64, device="cuda", requires_grad=True)
    val = e.nlml(x)
    (g,) = torch.autograd.grad(val, x)
    assert abs(val.item() - ref[0]) < 1e-12 * abs(ref[0])
    np.testing.assert_allclose(g.cpu().numpy(), ref[1:5], rtol=1e-12)
    u = e.u.clone().requires_grad_(True)
    (gu,) = torch.autograd.grad(e.nlml(u, unconstrained=True), u)
    np.testing.assert_allclose(gu.cpu().numpy(), ref[5:9], rtol=1e-12)
    assert torch.autograd.gradcheck(lambda t: e.nlml(t), (x,), eps=1e-6, atol=1e-4, rtol=1e-4)


@pytest.mark.parametrize("M,R", [(60, 3), (300, 2)])
def test_rbf_engine_vs_oracle(M, R):
    """kernel='rbf' engine: summed NLML over the output columns, the
    adjoint w.r.t. (variance, lengthscale, noise) and the predictive mean /
    variance vs the textbook RBF GPR of the oracle."""
    import torch
    from rvgp_amd.rbf import RBFEngine
    rng = np.random.default_rng(M)
    X = rng.normal(size=(M, 12))
    Y = rng.normal(size=(M, R))
    th = (1.4, 2.3, 0.07)
    e = RBFEngine(torch.as_tensor(X, device="cuda"), torch.as_tensor(Y, device="cuda"))
    e.set_theta(th)
    out = e.evaluate().cpu().numpy()
    ref = cpu.rbf_gpr_nlml(X, Y, *th)
    g = cpu.rbf_gpr_grad(X, Y, *th)
    assert abs(out[0] - ref) < 1e-10 * abs(ref)
    np.testing.assert_allclose(out[1:4], g, rtol=1e-8, atol=1e-10 * np.abs(g).max())
    Xs = rng.normal(size=(25, 12))
    m, v = e.predict(torch.as_tensor(Xs, device="cuda"))
    mr, vr = cpu.rbf_gpr_predict(X, Y, *th, Xs)
    np.testing.assert_allclose(m.cpu().numpy(), mr, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(v.cpu().numpy(), vr, rtol=1e-9, atol=1e-12)


def test_fit_rbf_kernel_vs_oracle_lbfgs():
    """RVGP.fit(d, kernel='rbf') end to end: evecs_L rows in, the D
    components as output columns, L-BFGS-B from (1, 1, noise 1.0): the
    optimum matches the oracle's L-BFGS-B on the same rows, and a fixed
    kernel_variance stays fixed."""
    import rvgp_amd as RVGP
    from tests.util import sphere
    X = sphere(400, 31)
    d = RVGP.create_data_object(X, n_neighbors=10, n_eigenpairs=20, verbose=False)
    d.random_vector_field(seed=2)
    d.smooth_vector_field(t=10)
    train = np.random.default_rng(4).choice(400, size=250, replace=False)
    gp = RVGP.fit(d, train_ind=train, kernel='rbf', epochs=200, verbose=False)
    tr, te = cpu.train_test_split_indices(len(train), 0.2, 0)
    nodes = train[tr]
    Xtr = np.asarray(d.evecs_L)[nodes]
    Ytr = np.asarray(d.vectors)[nodes]
    import scipy.optimize

    # the positivity bounds in force when the kernel was created (earlier fits
    # in the process may have set the default minimum, main.py:52)
    lower = gp.engine.lower.cpu().numpy()

    def f(u):
        th = cpu.softplus(u) + lower
        return (cpu.rbf_gpr_nlml(Xtr, Ytr, *th),
                cpu.rbf_gpr_grad(Xtr, Ytr, *th) / (1.0 + np.exp(-u)))
    u0 = cpu.inv_softplus(np.array([1.0, 1.0, 1.0]) - lower)
    res = scipy.optimize.minimize(f, u0, jac=True, method="L-BFGS-B", options={"maxiter": 200})
    th_ref = cpu.softplus(res.x) + lower
    th = gp.engine.theta_host()
    np.testing.assert_allclose(th, th_ref, rtol=2e-3)
    assert abs(gp.training_loss() - res.fun) < 1e-6 * abs(res.fun)
    # held-out error on the re-split nodes, from the fitted model
    Xte = np.asarray(d.evecs_L)[train[te]]
    m, _ = cpu.rbf_gpr_predict(Xtr, Ytr, *th, Xte)
    err = np.linalg.norm(np.asarray(d.vectors)[train[te]] - m, axis=1).mean()
    assert abs(gp.heldout_error - err) < 1e-8
    gp2 = RVGP.fit(d, train_ind=train, kernel='rbf', kernel_variance=0.5, epochs=50, verbose=False)
    assert gp2.hyperparameters["variance"] == pytest.approx(0.5, rel=1e-12)


def test_fps_on_high_dimensional_rows_matches_oracle():
    """FPS over spectral rows (D = n_eigenpairs, here 200 > 64) as main.py:59
    runs it to pick SGPR inducing points: bit-exact indices and distances."""
    from rvgp_amd import geometry
    x = np.random.default_rng(41).normal(size=(3000, 200))
    perm, lam = geometry.furthest_point_sampling(x, N=60, start_idx=7)
    perm_ref, lam_ref = cpu.fps(x, N=60, start_idx=7)
    np.testing.assert_array_equal(perm, perm_ref)
    np.testing.assert_array_equal(lam, lam_ref)


@pytest.mark.parametrize("m", [5, 20])
def test_sgpr_engine_vs_oracle(m):
    """SGPR bound and adjoint vs the oracle's GPflow-formula bound (value,
    central differences of the oracle for the gradient) and predictions."""
    import torch
    from rvgp_amd.sgpr import SGPREngine
    rng = np.random.default_rng(m)
    Phi, y, lam = problem(M_nodes=100, k=25, seed=m)
    Z = Phi[rng.choice(Phi.shape[0], m, replace=False)]
    th = np.array([1.6, 3.0, 1.1, 0.2])
    e = SGPREngine(torch.as_tensor(Phi, device="cuda"), torch.as_tensor(y, device="cuda"),
                   torch.as_tensor(Z, device="cuda"), torch.as_tensor(lam, device="cuda"), 900.0)
    e.set_theta(th)
    out = e.evaluate().cpu().numpy()

    def ref(t):
        S, _ = cpu.spectral_filter(lam, t[0], t[1], t[2], 900.0)
        return cpu.sgpr_nlml(Phi, y, S, t[3], Z)
    r0 = ref(th)
    assert abs(out[0] - r0) < 1e-9 * abs(r0)
    h = 1e-5
    for p in range(4):
        tp, tm = th.copy(), th.copy()
        tp[p] += h
        tm[p] -= h
        fd = (ref(tp) - ref(tm)) / (2 * h)
        assert abs(fd - out[1 + p]) < 1e-5 * max(1.0, abs(fd)), (p, fd, out[1 + p])
    Ps = rng.normal(size=(30, 25))
    mean, var = e.predict(torch.as_tensor(Ps, device="cuda"))
    S, _ = cpu.spectral_filter(lam, *th[:3], 900.0)
    mr, vr = cpu.sgpr_predict(Phi, y, S, th[3], Z, Ps)
    np.testing.assert_allclose(mean.cpu().numpy(), mr.ravel(), rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(var.cpu().numpy(), vr, rtol=1e-8, atol=1e-10)


def test_fit_with_inducing_points_vs_oracle():
    """RVGP.fit(..., n_inducing_points=m): inducing rows from FPS of the
    training rows (bit-exact vs the oracle's FPS on the same rows); after
    L-BFGS-B iterations the loss is the oracle's SGPR bound at the fitted
    hyperparameters and below its starting value; held-out predictions vs the
    oracle's SGPR predictive equations."""
    import torch
    import rvgp_amd as RVGP
    from tests.util import sphere
    X = sphere(500, 43)
    d = RVGP.create_data_object(X, n_neighbors=10, n_eigenpairs=25, verbose=False)
    d.random_vector_field(seed=3)
    d.smooth_vector_field(t=10)
    train = np.random.default_rng(5).choice(500, size=300, replace=False)
    # m = 15 inducing rows < k = 25 features: Kuu = Z S Z^T is regular (with
    # m > k it is singular up to GPflow's jitter and the bound is fragile)
    gp = RVGP.fit(d, train_ind=train, n_inducing_points=15, epochs=10, verbose=False)
    tr, te = cpu.train_test_split_indices(len(train), 0.2, 0)
    nodes = train[tr]
    Phi = d.phi_rows(torch.as_tensor(nodes.astype(np.int32), device="cuda")).cpu().numpy()
    y = np.asarray(d.vectors)[nodes].reshape(-1, 1)
    ind, _ = cpu.fps(Phi, N=15)
    np.testing.assert_array_equal(gp.engine.inducing_rows, ind)
    Z = Phi[ind]
    lam = np.asarray(d.evals_Lc)
    # the positivity bounds the kernel was created with (set_default_positive_
    # minimum is applied after the kernel exists, main.py:30 vs 52)
    lower = gp.engine.lower

    def f(u):
        t = cpu.softplus(u) + lower
        S, _ = cpu.spectral_filter(lam, t[0], t[1], t[2], 500 * 3)
        return cpu.sgpr_nlml(Phi, y, S, t[3], Z)
    u0 = cpu.inv_softplus(np.array([1.5, 5.0, 1.0, 1.0]) - lower)
    loss = gp.training_loss()
    assert abs(loss - f(gp.engine.u)) < 1e-8 * abs(loss)
    assert loss < f(u0)
    te_nodes = [int(i) for i in np.setdiff1d(np.arange(500), train)[:50]]
    mean, var = gp.transform(d, te_nodes)
    th = gp.engine.theta_host()
    S, _ = cpu.spectral_filter(lam, th[0], th[1], th[2], 500 * 3)
    Ps = d.phi_rows(torch.as_tensor(np.asarray(te_nodes, np.int32), device="cuda")).cpu().numpy()
    mr, vr = cpu.sgpr_predict(Phi, y, S, th[3], Z, Ps)
    assert mean.shape == (50, 3)
    np.testing.assert_allclose(mean.ravel(), mr.ravel(), rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(var.ravel(), vr.ravel(), rtol=1e-7, atol=1e-9)


@pytest.mark.parametrize("k", [1, 31, 32, 33, 100, 500])
def test_fused_step_equals_multi_launch_step(k):
    """The single-launch step (rvgp_ws_step: tile dataflow of Cholesky, L^-1,
    NLML, adjoint and Adam) against the multi-kernel weight-space evaluate
    and Adam step: NLML and gradients at a fixed theta, then the Adam
    trajectory over 25 steps."""
    Phi, y, lam = problem(M_nodes=max(60, k), k=k, seed=100 + k)
    th = (1.7, 4.0, 1.3, 0.2)
    n_rows = 3.0 * Phi.shape[0]
    a = engine(Phi, y, lam, n_rows, "weight", theta=th)
    b = engine(Phi, y, lam, n_rows, "weight", theta=th)
    a.fused_step = False
    assert b.fused() and not a.fused()
    a.adam_setup()
    b.adam_setup()
    out_a = a.evaluate().cpu().numpy()
    out_b = b._ws_step(0).cpu().numpy()
    assert int(b.status.item()) == 0
    np.testing.assert_allclose(out_b, out_a, rtol=1e-10, atol=1e-11 * np.abs(out_a).max())
    S, dS = cpu.spectral_filter(lam, th[0], th[1], th[2], n_rows)
    ref = cpu.gpr_nlml(Phi, y, S, th[3])
    assert abs(out_b[0] - ref) < 1e-10 * abs(ref)
    for _ in range(25):
        a.adam_step(lr=0.05)
        b.adam_step(lr=0.05)
    assert int(b.status.item()) == 0
    np.testing.assert_allclose(b.u.cpu().numpy(), a.u.cpu().numpy(), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(b.out.cpu().numpy(), a.out.cpu().numpy(), rtol=1e-8,
                               atol=1e-10 * np.abs(out_a).max())
    np.testing.assert_array_equal(b.t.cpu().numpy(), a.t.cpu().numpy())


def test_fused_step_at_the_size_limit_and_past_it():
    """k = 2048 (64 x 64 tiles, 4225 workgroups) runs fused; k = 2049 takes
    the multi-launch path."""
    import torch
    from rvgp_amd.gp import GPREngine
    rng = np.random.default_rng(5)
    for k, fused in ((2048, True), (2049, False)):
        lam = np.sort(rng.uniform(0.0, 8.0, size=k))
        Phi = rng.normal(size=(k + 200, k)) / np.sqrt(k)
        y = rng.normal(size=(k + 200, 1))
        Pd = torch.as_tensor(Phi, device="cuda")
        yd = torch.as_tensor(y, device="cuda")
        a = GPREngine(Pd, yd, torch.as_tensor(lam, device="cuda"), 3.0 * k, form="weight")
        a.set_theta((1.5, 3.0, 1.0, 0.5))
        a.adam_setup()
        assert a.fused() == fused
        ref = a.evaluate().cpu().numpy().copy() if fused else None
        if fused:
            out = a._ws_step(0).cpu().numpy()
            assert int(a.status.item()) == 0
            np.testing.assert_allclose(out, ref, rtol=1e-9, atol=1e-10 * np.abs(ref).max())
        a.adam_step(lr=0.05)
        assert int(a.status.item()) == 0


def test_fused_step_reports_a_non_positive_definite_system():
    """A = G + s2 diag(S)^-1 with G = -10 I is not positive definite: the
    fused step reports RVGP_ST_NOT_SPD like the multi-launch path, and the
    next call on a good system reports RVGP_ST_OK again."""
    import torch
    from rvgp_amd import _lib
    from rvgp_amd.gp import GPREngine
    k = 70
    lam = torch.linspace(0.0, 4.0, k, dtype=torch.float64, device="cuda")
    b = torch.ones(k, 1, dtype=torch.float64, device="cuda")
    yy = torch.full((1,), 5.0, dtype=torch.float64, device="cuda")
    for G, want in ((-10.0 * torch.eye(k, dtype=torch.float64, device="cuda"), _lib.ST_NOT_SPD),
                    (10.0 * torch.eye(k, dtype=torch.float64, device="cuda"), _lib.ST_OK)):
        e = GPREngine(None, None, lam, 3.0 * k, form="weight", stats=(G, b, yy, 4 * k))
        e.set_theta((1.5, 3.0, 1.0, 1e-3))
        e.adam_setup()
        e._ws_step(0)
        assert int(e.status.item()) == want
        e.fused_step = False
        e.evaluate()
        assert int(e.status.item()) == want


def test_fused_step_graph_replays_match_eager():
    """The captured single-launch step (one graph node per step) replayed
    with eager work in between equals the eager fused step every time."""
    import torch
    Phi, y, lam = problem(M_nodes=400, k=200, seed=23)
    a = engine(Phi, y, lam, 1200.0, "weight")
    b = engine(Phi, y, lam, 1200.0, "weight")
    a.adam_setup()
    b.adam_setup()
    assert a.fused() and b.fused()
    g = b.capture_adam_step(lr=0.05)
    for step in range(15):
        a.adam_step(lr=0.05)
        g.replay()
        torch.cuda.synchronize()
        _ = (torch.ones(1000, device="cuda") * step).sum().item()
        np.testing.assert_array_equal(b.u.cpu().numpy(), a.u.cpu().numpy())
        np.testing.assert_array_equal(b.out.cpu().numpy(), a.out.cpu().numpy())
