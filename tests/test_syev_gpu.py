"""rvgp_syev (the Rayleigh-Ritz step's dense symmetric eigenproblem, which
ARPACK solves inside eigsh, RVGP/geometry.py:73) against numpy.linalg.eigh:
eigenvalues, residuals |A V - V diag(w)| and orthogonality |V^T V - I|, on
random, clustered, exactly degenerate, diagonal, tiny and Rayleigh-Ritz-like
matrices."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sym_with_spectrum(w, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(len(w), len(w))))
    return (Q * w) @ Q.T


def _cases():
    rng = np.random.default_rng(0)
    out = []
    for n in (1, 2, 3, 5, 33, 64, 65, 100, 257, 576, 1000, 1100):
        B = rng.normal(size=(n, n))
        out.append((f"random{n}", 0.5 * (B + B.T)))
    # clustered: multiplets like the flat-torus connection Laplacian, split by
    # 1e-6 .. 1e-12 relative, plus exact repeats
    base = np.repeat(np.linspace(0.01, 40.0, 48), 12)[:576]
    jit = np.random.default_rng(1).normal(size=576) * np.repeat([1e-6, 1e-9, 1e-12, 0.0], 144)
    out.append(("clustered576", _sym_with_spectrum(np.sort(base + jit * base), 2)))
    out.append(("degenerate200", _sym_with_spectrum(np.r_[np.ones(100), 2 * np.ones(100)], 3)))
    out.append(("diagonal50", np.diag(np.linspace(-3, 7, 50))))
    out.append(("zero7", np.zeros((7, 7))))
    # nearly diagonal, as the later Rayleigh-Ritz steps see
    d = np.sort(np.random.default_rng(4).uniform(0, 40, 400))
    E = np.random.default_rng(5).normal(size=(400, 400)) * 1e-5
    out.append(("neardiag400", np.diag(d) + 0.5 * (E + E.T)))
    return out


@pytest.mark.parametrize("name,A", _cases(), ids=[c[0] for c in _cases()])
def test_syev_matches_numpy_eigh(name, A):
    import torch
    from rvgp_amd import dense
    n = A.shape[0]
    w, V = dense.syev(torch.as_tensor(A, device="cuda"))
    w, V = w.cpu().numpy(), V.cpu().numpy()
    wr = np.linalg.eigvalsh(A)
    scale = max(np.abs(wr).max(), 1e-300)
    assert np.all(np.diff(w) >= -1e-14 * scale)
    np.testing.assert_allclose(w, wr, rtol=0, atol=1e-12 * scale * max(1, np.sqrt(n)) + 1e-300)
    res = np.abs(A @ V - V * w).max()
    assert res <= 1e-12 * scale * max(1.0, np.sqrt(n)) + 1e-300, res
    orth = np.abs(V.T @ V - np.eye(n)).max()
    assert orth <= 1e-10, orth
