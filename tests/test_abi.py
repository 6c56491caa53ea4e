"""CPU tests of the C ABI and host logic (no GPU compute calls)."""
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "rvgp.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rvgp_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from rvgp_amd import _lib
    assert os.path.exists(_lib.LIB_PATH), "run __graft_entry__.build() first"
    L = _lib.load_symbols_only()
    names = declared_symbols()
    assert len(names) >= 25
    for name in names:
        assert hasattr(L, name), name
        assert name in _lib.SIGNATURES, f"{name} missing from the ctypes table"
    assert L.rvgp_version() >= 1


def test_ctypes_table_has_no_undeclared_entries():
    from rvgp_amd import _lib
    assert set(_lib.SIGNATURES) <= set(declared_symbols())


def test_arg_validation_reports_value_errors_without_a_gpu():
    """Argument checks run before any device work: null pointers are fine."""
    from rvgp_amd import _lib
    L = _lib.load_symbols_only()
    rc = L.rvgp_knn(None, 10, 3, 10, None, None, None, 0, None)  # k >= n
    assert rc == _lib.RVGP_EINVAL
    assert b"0 < k < n" in L.rvgp_last_error()
    rc = L.rvgp_bsr_spmm_f64(10, 4, None, None, None, None, 1, 1, 1.0, 0.0, None, 1, 0.0, None, 1,
                             None, 1, 0.0, None)
    assert rc == _lib.RVGP_EINVAL
    # binary16 SpMM: r a multiple of 8, 16-byte aligned operands
    rc = L.rvgp_bsr_spmm_f16(10, 3, None, None, None, None, 12, 12, 1.0, 0.0, None, 12, 0.0,
                             None, 12, None)
    assert rc == _lib.RVGP_EINVAL and b"multiple of 8" in L.rvgp_last_error()
    rc = L.rvgp_bsr_spmm_f16(10, 3, None, None, None, 8, 16, 16, 1.0, 0.0, None, 16, 0.0,
                             64, 16, None)
    assert rc == _lib.RVGP_EINVAL and b"aligned" in L.rvgp_last_error()
    rc = L.rvgp_gemm(7, 0, 0, 1, 1, 1, 1.0, None, 1, None, 1, None, 0.0, None, 1, 1, None, 0, None)
    assert rc == _lib.RVGP_EINVAL
    rc = L.rvgp_gauges(None, 10, 65, None, 10, 1, 1, 1, None, None, None, None)
    assert rc == _lib.RVGP_EINVAL
    # the single-launch weight-space step: k range, workspace and flag sizes
    def ws_step(k, ws_bytes, sync_ints):
        # k, typ, evals, n_rows, G, ldg, b, yy, M, lower, u, trainable, m1, m2, t,
        # do_adam, lr, b1, b2, eps, ws, ws_bytes, sync, sync_ints, out, status, stream
        return L.rvgp_ws_step(k, 0, None, 1.0, None, max(k, 1), None, None, 1.0, None, None, None,
                              None, None, None, 0, 0.0, 0.9, 0.999, 1e-8, None, ws_bytes, None,
                              sync_ints, None, None, None)
    assert ws_step(0, 1 << 40, 1 << 20) == _lib.RVGP_EINVAL
    assert ws_step(2049, 1 << 40, 1 << 20) == _lib.RVGP_EINVAL
    assert ws_step(100, L.rvgp_ws_step_ws_bytes(100) - 8,
                   L.rvgp_ws_step_sync_ints(100)) == _lib.RVGP_EINVAL
    assert b"workspace" in L.rvgp_last_error()
    # Rayleigh-Ritz residuals and the eigenvector output pass
    rc = L.rvgp_ritz_residuals(1, 10, 5, None, 4, None, 5, None, None, None, 0, None)  # ldx < k
    assert rc == _lib.RVGP_EINVAL
    rc = L.rvgp_ritz_residuals(1, 10, 5, 1, 5, 1, 5, 1, 1, 1, 8, None)  # workspace too small
    assert rc == _lib.RVGP_EINVAL and b"workspace" in L.rvgp_last_error()
    rc = L.rvgp_evecs_finish(1, 10, 4, 5, None, 5, None, None, 1.0, None, 5, None)  # b > 3
    assert rc == _lib.RVGP_EINVAL
    rc = L.rvgp_evecs_finish(1, 10, 3, 5, 1, 5, 1, None, 1.0, 1, 5, None)  # out aliases X
    assert rc == _lib.RVGP_EINVAL and b"alias" in L.rvgp_last_error()


def test_workspace_queries():
    from rvgp_amd import _lib
    L = _lib.load_symbols_only()
    assert L.rvgp_graph_symmetrize_ws_bytes(100, 10) >= 4 * 100 * 21
    # pools up to 512 nodes: one pass, workspace linear in the lanes in flight
    assert L.rvgp_geo_nbh_ws_bytes(1000, 2, 100) == 2 * L.rvgp_geo_nbh_ws_bytes(1000, 1, 100)
    # a larger worst-case pool adds the second pass's lanes and the overflow list
    assert L.rvgp_geo_nbh_ws_bytes(1000, 4096, 1200) > 4096 * L.rvgp_geo_nbh_ws_bytes(1000, 1, 512)
    assert L.rvgp_colsumsq_ws_bytes(1000, 10) >= 8 * 10
    assert L.rvgp_colsumsq_ws_bytes(10**6, 10) == 8 * 10 * 1024  # at most 1024 row slices
    # two f64 partial sums per (32-row slice, column), at most 1024 slices
    assert L.rvgp_ritz_residuals_ws_bytes(100, 7) == 2 * 8 * 4 * 7
    assert L.rvgp_ritz_residuals_ws_bytes(10**6, 7) == 2 * 8 * 1024 * 7
    # fused step: T = ceil(k/32) tile rows; two packed triangles of 32 x 32
    # tiles (L and L^-1) plus t, w, diag(A^-1), S and dS (3) of length 32 T,
    # 8 partial sums per tile row and 8 scalars; flags: T(T-1)/2 + 1 L tasks,
    # T diagonal, T(T-1)/2 X tasks, T rows, T columns, the filter, and 4
    # control words
    for k, T in ((1, 1), (32, 1), (33, 2), (500, 16), (2048, 64)):
        assert L.rvgp_ws_step_ws_bytes(k) == 8 * (T * (T + 1) * 1024 + 7 * 32 * T + 8 * T + 8)
        assert L.rvgp_ws_step_sync_ints(k) == T * T + 2 * T + 6
    assert L.rvgp_ws_step_ws_bytes(2049) == 0 and L.rvgp_ws_step_sync_ints(0) == 0


def test_host_split_rule_and_softplus():
    from rvgp_amd.main import train_test_split_indices
    from rvgp_amd.gp import inv_softplus, softplus
    from sklearn.model_selection import train_test_split
    tr, te = train_test_split(np.arange(101), test_size=0.2, random_state=3)
    tr2, te2 = train_test_split_indices(101, 0.2, 3)
    np.testing.assert_array_equal(tr, tr2)
    np.testing.assert_array_equal(te, te2)
    x = np.array([1e-6, 0.3, 5.0, 40.0])
    np.testing.assert_allclose(softplus(inv_softplus(x)), x, rtol=1e-12)


def test_chebyshev_heat_coefficients():
    """The truncated Chebyshev series of exp(-t x) on [0, ub] is exact to
    ~1e-16 at sample points (host-side coefficient logic)."""
    from rvgp_amd.smoothing import cheb_coeffs
    for t, ub in ((0.5, 10.0), (100.0, 20.0), (500.0, 45.0)):
        c = cheb_coeffs(t, ub)
        x = np.linspace(0, ub, 41)
        y = np.polynomial.chebyshev.chebval(2 * x / ub - 1, c)
        np.testing.assert_allclose(y, np.exp(-t * x), atol=1e-14)


def test_package_imports_without_gpu():
    import rvgp_amd
    from rvgp_amd import geometry, dataclass, main, kernels, smoothing, eigsolve  # noqa: F401
    assert callable(rvgp_amd.create_data_object) and callable(rvgp_amd.fit)
