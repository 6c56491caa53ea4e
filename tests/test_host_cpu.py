"""CPU tests of host logic that must act before any kernel launch."""
import types

import numpy as np
import pytest


def test_checked_nodes_follows_numpy_indexing():
    from rvgp_amd.main import checked_nodes
    n = 10
    a = np.arange(n) * 3.0
    for idx in ([0, 3, 9], [-1, -10, 4], np.array([2, -3], np.int32), [True] * 4 + [False] * 6):
        got = checked_nodes(idx, n)
        np.testing.assert_array_equal(a[got], a[np.asarray(idx)])
        assert got.min() >= 0 and got.max() < n
    for bad in ([10], [-11], [0, 5, 12]):
        with pytest.raises(IndexError):
            a[np.asarray(bad)]  # numpy's own behaviour
        with pytest.raises(IndexError, match="out of bounds"):
            checked_nodes(bad, n)
    with pytest.raises(IndexError):
        checked_nodes([0.5, 1.0], n)
    with pytest.raises(IndexError):
        checked_nodes([True, False], n)  # mask of the wrong length
    assert checked_nodes([], n).size == 0


def test_out_of_range_train_and_test_nodes_raise_before_any_launch():
    """train_gp and transform validate the node indices first: with a stand-in
    data object and engine that hold no device state, an out-of-range index
    raises IndexError (nothing reaches a kernel)."""
    import rvgp_amd as RVGP
    from rvgp_amd.main import manifold_GPR
    data = types.SimpleNamespace(n=20)
    for bad in ([20], [-21], np.array([3, 40])):
        with pytest.raises(IndexError):
            RVGP.fit(data, train_ind=bad, epochs=1, verbose=False)
    eng = types.SimpleNamespace(dev="cpu", R=1, k=5, form="weight", precision="fp64")
    gp = manifold_GPR(data, None, np.arange(3), eng)
    with pytest.raises(IndexError):
        gp.transform(data, [0, 25])
    with pytest.raises(IndexError):
        gp.predict_nodes(data, np.array([-30]))


def test_dense_operand_stride_rules():
    """Leading dimensions handed to the GEMM / Cholesky kernels: a single
    column may carry any column stride, a broadcast (stride-0) view is
    refused instead of letting the kernel read rows that are not stored."""
    import torch
    from rvgp_amd.dense import _ld
    a = torch.zeros(6, 4, dtype=torch.float64)
    assert _ld(a) == 4
    assert _ld(a[:, 1:3]) == 4
    assert _ld(a[:, 2:3]) == 4
    assert _ld(torch.zeros(1, 5)) == 5
    with pytest.raises(ValueError):
        _ld(torch.zeros(1, 4).expand(6, 4))
    with pytest.raises(ValueError):
        _ld(torch.zeros(1, 1).expand(6, 1))


def test_column_degrees_are_monotone_grouped_and_bounded():
    """_column_degrees: non-decreasing along the block, constant over groups
    of 8 columns, within [4, m], and m for the guard columns beyond kk; a
    column far below the cut with a small residual needs fewer steps than one
    near the cut."""
    from rvgp_amd.eigsolve import _column_degrees
    theta = np.linspace(0.0, 0.95, 40)
    res = np.full(40, 1e-3)
    res[:8] = 1e-6
    ms = _column_degrees(theta, res, 1e-3, 56, 40, 1.0, 20.0, 30, 1e-7)
    assert ms.shape == (56,)
    assert np.all(np.diff(ms) >= 0) and ms.min() >= 4 and ms.max() <= 30
    assert np.all(ms.reshape(7, 8) == ms.reshape(7, 8)[:, :1])
    assert np.all(ms[40:] == 30) and ms[0] < ms[39]
