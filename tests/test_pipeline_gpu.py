"""End-to-end drop-in pipeline on the GPU vs the FP64 oracle pipeline:
create_data_object -> random_vector_field -> smooth_vector_field -> fit ->
transform, on a C1-style sphere at reduced n so the dense-expm oracle runs
in seconds."""
import numpy as np
import pytest

from oracle import cpu
from tests.util import sphere

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pair():
    import rvgp_amd as RVGP
    X = sphere(800, 0)
    d = RVGP.create_data_object(X, n_neighbors=10, n_eigenpairs=30, verbose=False)
    o = cpu.Data(X, n_neighbors=10, n_eigenpairs=30)
    return d, o, X


def test_data_object_matches_oracle(pair):
    d, o, X = pair
    assert d.n == 800 and d.dim_man == o.dim_man == 2
    np.testing.assert_allclose(d.evals_Lc, o.evals_Lc, rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(d.evals_L[1:], o.evals_L[1:], rtol=1e-6)
    # gauge-invariant: projectors onto tangent planes
    P = np.einsum("nij,nkj->nik", d.gauges, d.gauges)
    Po = np.einsum("nij,nkj->nik", o.gauges, o.gauges)
    assert np.abs(P - Po).max() < 1e-10
    assert d.evecs_Lc.shape == (800 * 3, 30)
    # parallel transport in ambient coordinates, T_i R_ij T_j^T, is gauge invariant
    rows = np.repeat(np.arange(800), np.diff(o.rowptr))
    Rg = d._R.cpu().numpy()
    amb = np.einsum("eij,ejk,elk->eil", d.gauges[rows], Rg, d.gauges[o.cols])
    amb_o = np.einsum("eij,ejk,elk->eil", o.gauges[rows], o.Rblocks, o.gauges[o.cols])
    assert np.abs(amb - amb_o).max() < 1e-9
    assert d.R.shape == (1600, 1600)
    assert d.G.number_of_nodes() == 800


def test_random_and_smoothed_field_match_oracle(pair):
    d, o, X = pair
    # projections and the smoothed ambient field are gauge invariant
    d.random_vector_field(seed=1)
    o.random_vector_field(seed=1)
    np.testing.assert_allclose(d.vectors, o.vectors, rtol=1e-12, atol=1e-13)
    d.smooth_vector_field(t=10)
    o.smooth_vector_field(t=10, method="expm")
    np.testing.assert_allclose(d.vectors, o.vectors, rtol=0, atol=1e-9 * np.abs(o.vectors).max())


def test_fit_and_transform_match_oracle(pair):
    import rvgp_amd as RVGP
    d, o, X = pair
    rng = np.random.default_rng(4)
    train_ind = rng.choice(np.arange(800), size=400, replace=False)
    gp = RVGP.fit(d, train_ind=train_ind, epochs=30, verbose=False)
    test_ind = [int(i) for i in range(800) if i not in set(train_ind.tolist())]
    mean, var = gp.transform(d, test_ind)
    assert mean.shape == (len(test_ind), 3) and var.shape == (len(test_ind), 3)
    # the oracle at the GPU's fitted hyperparameters, on the oracle's own
    # eigenpairs and vectors (gauge-invariant ambient predictions)
    th = gp.engine.theta_host()
    tr, _ = cpu.train_test_split_indices(400, 0.2, 0)
    nodes = train_ind[tr]
    Phi = o.evecs_Lc.reshape(800, -1)[nodes].reshape(-1, 30)
    y = o.vectors[nodes].reshape(-1, 1)
    S, _ = cpu.spectral_filter(o.evals_Lc, th[0], th[1], th[2], 800 * 3)
    Ps = o.evecs_Lc.reshape(800, -1)[test_ind].reshape(-1, 30)
    mr, vr = cpu.gpr_predict(Phi, y, S, th[3], Ps)
    rel = np.linalg.norm(mean.ravel() - mr.ravel()) / np.linalg.norm(mr)
    assert rel < 1e-4, rel
    with pytest.raises(UnboundLocalError):
        gp.transform(d, np.asarray(test_ind))  # np.int64 dispatch (main.py:104-109)


def test_fit_reference_quirks(pair):
    import rvgp_amd as RVGP
    d, _, _ = pair
    with pytest.raises(AttributeError):
        RVGP.fit(d, train_ind=np.arange(100), kernel_variance=1.0, epochs=1, verbose=False)
    # a kernel object leaves `input` unbound in the reference (main.py:29-37)
    with pytest.raises(UnboundLocalError):
        RVGP.fit(d, train_ind=np.arange(100), kernel=object(), epochs=1, verbose=False)
    with pytest.raises(NotImplementedError):  # SGPR with the RBF kernel is not built
        RVGP.fit(d, train_ind=np.arange(100), kernel='rbf', n_inducing_points=10, epochs=1,
                 verbose=False)


def test_mixed_precision_pipeline_within_1e4(pair):
    """The mixed form (FP32 eigensolver body, default residual tolerance
    1e-7 of the spectral bound, FP64 k x k Woodbury system) against
    the FP64 form on the same points, field and training split."""
    import rvgp_amd as RVGP
    _, _, X = pair
    out = {}
    for prec in ("fp64", "mixed"):
        d = RVGP.create_data_object(X, n_neighbors=10, n_eigenpairs=30, precision=prec,
                                    verbose=False)
        d.random_vector_field(seed=3)
        d.smooth_vector_field(t=10)
        train_ind = np.random.default_rng(5).choice(np.arange(800), size=400, replace=False)
        gp = RVGP.fit(d, train_ind=train_ind, epochs=30, precision=prec, verbose=False)
        test_ind = [int(i) for i in range(800) if i not in set(train_ind.tolist())]
        mean, _ = gp.transform(d, test_ind)
        out[prec] = (np.asarray(d.evals_Lc), np.asarray(d.vectors), np.asarray(mean))
    (e64, v64, m64), (em, vm, mm) = out["fp64"], out["mixed"]
    # |theta - lambda| <= residual <= 1e-7 x the Gershgorin bound (< 4 * 2 * 20 here)
    np.testing.assert_allclose(em, e64, rtol=0, atol=1e-5)
    assert np.abs(vm - v64).max() < 1e-4 * np.abs(v64).max()
    assert np.linalg.norm(mm - m64) / np.linalg.norm(m64) < 1e-4


@pytest.mark.parametrize("fields", [1, 3])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_weight_statistics_from_intrinsic_rows(pair, fields, prec):
    """Phi^T Phi = U^T U and Phi^T y = U^T (T^T y) (orthonormal frames): the
    engine built from the intrinsic statistics gives the NLML and gradient
    of the one built from the ambient rows."""
    import torch
    from rvgp_amd.gp import GPREngine, weight_space_stats
    d, _, _ = pair
    rng = np.random.default_rng(8)
    nodes = torch.as_tensor(rng.choice(800, size=500, replace=False), device="cuda")
    V = torch.as_tensor(rng.normal(size=(fields, 800, 3)), device="cuda")
    V = V[0] if fields == 1 else V
    Phi = d.phi_rows(nodes.int())
    if fields == 1:
        y = V[nodes].reshape(-1)
        c = d._frame(0, V, d.dim_man)[nodes].reshape(-1, 1).contiguous()
        yv = V[nodes]
    else:
        y = V[:, nodes].permute(1, 2, 0).reshape(-1, fields).contiguous()
        c = d._frame(0, V, d.dim_man)[:, nodes].permute(1, 2, 0).reshape(-1, fields).contiguous()
        yv = V[:, nodes]
    lam = d.spectrum_device()
    a = GPREngine(Phi, y, lam, 800 * 3, form="weight", precision=prec)
    G, b = weight_space_stats(d.intrinsic_rows(nodes), c, prec)
    e = GPREngine(None, None, lam, 800 * 3, form="weight", precision=prec,
                  stats=(G, b, (yv * yv).sum().reshape(1), 1500))
    for th in [(1.5, 5.0, 1.0, 0.3), (2.2, 3.0, 0.7, 0.05)]:
        a.set_theta(th)
        e.set_theta(th)
        oa, oe = a.evaluate().cpu().numpy(), e.evaluate().cpu().numpy()
        tol = 1e-10 if prec == "fp64" else 1e-6
        assert abs(oa[0] - oe[0]) < tol * abs(oa[0])
        np.testing.assert_allclose(oe[1:], oa[1:], rtol=tol, atol=tol * np.abs(oa[1:]).max())


def test_weight_form_fit_vs_oracle(pair):
    """RVGP.fit through the intrinsic statistics (form='weight'), then the
    ambient predictions vs the oracle at the fitted hyperparameters."""
    import rvgp_amd as RVGP
    d, o, X = pair
    d.random_vector_field(seed=2)
    o.random_vector_field(seed=2)
    d.smooth_vector_field(t=10)
    o.smooth_vector_field(t=10, method="expm")
    train_ind = np.random.default_rng(6).choice(np.arange(800), size=400, replace=False)
    gp = RVGP.fit(d, train_ind=train_ind, epochs=30, form="weight", verbose=False)
    assert gp.engine.Phi is None and gp.engine.M == 320 * 3
    test_ind = [int(i) for i in range(800) if i not in set(train_ind.tolist())]
    mean, _ = gp.transform(d, test_ind)
    th = gp.engine.theta_host()
    tr, _ = cpu.train_test_split_indices(400, 0.2, 0)
    nodes = train_ind[tr]
    Phi = o.evecs_Lc.reshape(800, -1)[nodes].reshape(-1, 30)
    y = o.vectors[nodes].reshape(-1, 1)
    S, _ = cpu.spectral_filter(o.evals_Lc, th[0], th[1], th[2], 800 * 3)
    Ps = o.evecs_Lc.reshape(800, -1)[test_ind].reshape(-1, 30)
    mr, _ = cpu.gpr_predict(Phi, y, S, th[3], Ps)
    assert np.linalg.norm(mean.ravel() - mr.ravel()) / np.linalg.norm(mr) < 1e-4


@pytest.mark.parametrize("fields", [1, 2])
def test_predict_nodes_equals_ambient_rows(pair, fields):
    """Mean and variance at test nodes from intrinsic rows and frames equal
    predict() on the nodes' ambient rows Phi = T U."""
    import torch
    from rvgp_amd.gp import GPREngine, weight_space_stats
    d, _, _ = pair
    rng = np.random.default_rng(12)
    nodes = torch.as_tensor(rng.choice(800, size=300, replace=False), device="cuda")
    V = torch.as_tensor(rng.normal(size=(fields, 800, 3)), device="cuda")
    c = d._frame(0, V, d.dim_man)[:, nodes].permute(1, 2, 0).reshape(-1, fields).contiguous()
    yv = V[:, nodes]
    G, b = weight_space_stats(d.intrinsic_rows(nodes), c, "fp64")
    e = GPREngine(None, None, d.spectrum_device(), 800 * 3, form="weight",
                  stats=(G, b, (yv * yv).sum().reshape(1), 900))
    e.set_theta((1.7, 4.0, 0.9, 0.2))
    test = torch.as_tensor(np.setdiff1d(np.arange(800), nodes.cpu().numpy())[:200], device="cuda")
    m1, v1 = e.predict(d.phi_rows(test.int()))
    T = d._gauges.index_select(0, test)[:, :, :d.dim_man]
    m2, v2 = e.predict_nodes(d.intrinsic_rows(test), T)
    m1, m2 = m1.cpu().numpy(), m2.cpu().numpy()
    np.testing.assert_allclose(m2, m1, rtol=1e-10, atol=1e-12 * np.abs(m1).max())
    np.testing.assert_allclose(v2.cpu().numpy(), v1.cpu().numpy(), rtol=1e-9, atol=1e-14)


def test_transform_float_positional_encodings(pair):
    """transform with floats (main.py:108-109): spectral rows given directly
    (here the ambient rows of evecs_Lc of some nodes, flattened) predict the
    same values as the int branch at those nodes, one row per spectral row."""
    import rvgp_amd as RVGP
    d, o, X = pair
    d.random_vector_field(seed=5)
    train_ind = np.random.default_rng(13).choice(np.arange(800), size=300, replace=False)
    gp = RVGP.fit(d, train_ind=train_ind, epochs=15, verbose=False)
    nodes = [3, 17, 400, 799]
    m_int, v_int = gp.transform(d, nodes)
    rows = d.evecs_Lc.reshape(800, -1)[nodes].reshape(-1, d.evecs_Lc.shape[1])
    floats = [float(z) for z in rows.ravel()]
    m_f, v_f = gp.transform(d, floats)
    assert m_f.shape == (len(nodes) * 3, 1) and v_f.shape == (len(nodes) * 3, 1)
    np.testing.assert_allclose(m_f.ravel(), m_int.ravel(), rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(v_f.ravel(), v_int.ravel(), rtol=1e-9, atol=1e-13)
