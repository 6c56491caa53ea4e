"""Host AddressSanitizer / UndefinedBehaviorSanitizer run of the heap
neighbourhood search (rvgp_amd/csrc/nbh_heap.h, the code every lane of the
geo_nbh kernel runs).  The reference's heap runs with Cython bounds checks
off (RVGP/lib/ptu_dijkstra.pyx:217-220); here the same search is built for
the host with both sanitizers (tests/sanitize/nbh_host.hip, per-lane buffers
as separate exact-size allocations), run over the golden graphs, a lattice
full of distance ties and a pool too small to finish, and its neighbourhoods
are compared with the oracle's C restatement."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import cpu
from tests.util import golden_cases

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if not os.path.exists(HIPCC) and shutil.which("hipcc") is None:
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("asan") / "nbh_host_asan")
    cmd = [HIPCC, "-x", "hip", "--offload-host-only", "-std=c++17", "-g", "-O1",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-sanitize-recover=all",
           os.path.join(HERE, "sanitize", "nbh_host.hip"), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("sanitizer build unavailable: " + r.stderr[-500:])
    return out


def _run(driver, tmp_path, rowptr, cols, K, pool):
    n = len(rowptr) - 1
    w = np.ones(len(cols))
    inp, outp = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    with open(inp, "wb") as f:
        np.array([n, K, pool, len(cols)], np.int32).tofile(f)
        np.asarray(rowptr, np.int32).tofile(f)
        np.asarray(cols, np.int32).tofile(f)
        w.astype(np.float64).tofile(f)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=99",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=98")
    r = subprocess.run([driver, inp, outp], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr
    raw = np.fromfile(outp, np.int32)
    return int(raw[0]), raw[1:].reshape(n, K + 1)


def _lattice(m):
    g = np.stack(np.meshgrid(np.arange(m), np.arange(m), indexing="ij"), -1).reshape(-1, 2) * 1.0
    return g


@pytest.mark.parametrize("case", golden_cases() + ["lattice"])
def test_heap_search_is_clean_under_asan_and_ubsan(driver, tmp_path, case):
    if case == "lattice":
        X, k = _lattice(14), 8
    else:
        z = np.load(case)
        X, k = z["X"], int(z["k"])
    rp, c = cpu.sym_csr(cpu.knn(X, k))
    K = int(1.5 * k)
    ref, _ = cpu.geo_nbh(X, rp, c, K, centred=False)
    maxrow = int(np.diff(rp).max())
    st, nbh = _run(driver, tmp_path, rp, c, K, 1 + K * maxrow)
    assert st == 0
    np.testing.assert_array_equal(nbh, ref)
    # a pool too small for any search to finish: the overflow path reports
    # it without touching memory past the pool
    st, _ = _run(driver, tmp_path, rp, c, K, K + 2)
    assert st == 1
