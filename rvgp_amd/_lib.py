"""ctypes binding of librvgp.so (the C ABI declared in include/rvgp.h).

The library is built in-tree by ``python -m rvgp_amd._build``.  There is no
fallback: if the library or a GPU is missing, the first call raises.
torch is imported (and its HIP runtime loaded) before the library, so the
library binds to the same libamdhip64 instance and torch's streams and
device pointers are valid in it.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librvgp.so")
_LIB = None

P = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double
F = ctypes.c_float
S = ctypes.c_size_t

# name -> (restype, argtypes); mirrors include/rvgp.h
SIGNATURES = {
    "rvgp_version": (I, []),
    "rvgp_last_error": (ctypes.c_char_p, []),
    "rvgp_fps_diam2": (I, [P, I, I, P, P]),
    "rvgp_fps": (I, [P, I, I, I, I, D, I, P, P, P, P, P, P]),
    "rvgp_knn_ws_bytes": (S, [I, I]),
    "rvgp_knn": (I, [P, I, I, I, P, P, P, S, P]),
    "rvgp_affinity": (I, [P, I, I, D, P, I, P]),
    "rvgp_graph_symmetrize_ws_bytes": (S, [I, I]),
    "rvgp_graph_symmetrize": (I, [P, I, I, P, P, P, S, P]),
    "rvgp_geo_nbh_ws_bytes": (S, [I, I, I]),
    "rvgp_geo_nbh": (I, [P, P, P, I, I, I, I, P, P, S, P, P]),
    "rvgp_gauges": (I, [P, I, I, P, I, I, I, I, P, P, P, P]),
    "rvgp_connections": (I, [P, I, I, I, I, P, P, I, P, P, P, P, P, P]),
    "rvgp_bsr_spmm_f64": (I, [I, I, P, P, P, P, I, I, D, D, P, I, D, P, I, P, I, D, P]),
    "rvgp_bsr_spmm_f32": (I, [I, I, P, P, P, P, I, I, F, F, P, I, F, P, I, P, I, F, P]),
    "rvgp_bsr_spmm_f16": (I, [I, I, P, P, P, P, I, I, F, F, P, I, F, P, I, P]),
    "rvgp_gemm": (I, [I, I, I, I, I, I, D, P, I, P, I, P, D, P, I, I, P, S, P]),
    "rvgp_gemm_ws_bytes": (S, [I, I, I, I, I]),
    "rvgp_potrf": (I, [I, I, P, I, P, P]),
    "rvgp_trsm": (I, [I, I, I, I, P, I, P, I, P]),
    "rvgp_chol_logdet": (I, [I, I, P, I, P, P]),
    "rvgp_frame_apply": (I, [I, I, I, I, I, P, I, P, P, P]),
    "rvgp_diffusion_normalise": (I, [S, I, P, I, P, I, P, I, P]),
    "rvgp_phi_rows": (I, [I, I, I, I, I, P, I, I, P, I, P, P, I, P]),
    "rvgp_gemm_sym": (I, [I, I, I, P, I, P, I, P, I, I, P, S, P]),
    "rvgp_syev_ws_bytes": (S, [I]),
    "rvgp_syev": (I, [I, P, I, P, P, I, P, S, P]),
    "rvgp_colsumsq_ws_bytes": (S, [I, I]),
    "rvgp_colsumsq": (I, [I, I, I, P, I, P, P, S, P]),
    "rvgp_colsumsq_noalloc": (I, [I, I, I, P, I, P, P]),
    "rvgp_ritz_residuals_ws_bytes": (S, [I, I]),
    "rvgp_ritz_residuals": (I, [I, I, I, P, I, P, I, P, P, P, S, P]),
    "rvgp_evecs_finish": (I, [I, I, I, I, P, I, P, P, D, P, I, P]),
    "rvgp_spectral_filter": (I, [I, I, P, P, D, P, P, P]),
    "rvgp_ws_form": (I, [I, P, I, P, P, P, I, P]),
    "rvgp_ws_finish": (I, [I, D, I, P, I, P, P, P, P, P, P, P, P, P, P, P, P]),
    "rvgp_fs_finish": (I, [I, D, I, P, P, P, P, P, P, P, P, P, P, P, P]),
    "rvgp_dot_ws_bytes": (S, [S]),
    "rvgp_dot": (I, [S, P, P, P, P, S, P]),
    "rvgp_lemma_terms": (I, [I, D, P, P, P, P, I, P, P, P]),
    "rvgp_gp_params": (I, [P, P, P, P, P]),
    "rvgp_adam_step": (I, [P, P, P, P, P, P, D, D, D, D, P]),
    "rvgp_ws_step_ws_bytes": (S, [I]),
    "rvgp_ws_step_sync_ints": (S, [I]),
    "rvgp_ws_step": (I, [I, I, P, D, P, I, P, P, D, P, P, P, P, P, P, I, D, D, D, D, P, S, P, S,
                         P, P, P]),
    "rvgp_batch_gram": (I, [I, I, I, I, P, I, P, P, I, P, P]),
    "rvgp_batch_rhs": (I, [I, I, I, P, I, P, P, P, P, P, P, P]),
    "rvgp_batch_step_ws_bytes": (S, [I, I]),
    "rvgp_rbf_params": (I, [P, P, P, P]),
    "rvgp_sgpr_finish": (I, [I, I, D, P, P, P, P, P, P, P, P, P, P, P, P, P, D, P, P, P, P]),
    "rvgp_rbf_kernel": (I, [I, I, P, I, P, P, P, I, I, P, I, P]),
    "rvgp_rbf_grad_ws_bytes": (S, [I]),
    "rvgp_rbf_grad": (I, [I, I, P, I, P, I, P, I, P, P, P, P, P, P, S, P, P]),
    "rvgp_batch_step": (I, [I, I, I, P, D, P, P, P, P, P, P, P, P, P, P, I, D, D, D, D, I, I, I,
                            P, S, P, P, P]),
}

RVGP_EINVAL, RVGP_ERUNTIME, RVGP_ENOMEM, RVGP_EHIP = -1, -2, -3, -4
ST_OK, ST_RANK_DEFICIENT, ST_SHORT_COMPONENT, ST_POOL_OVERFLOW, ST_NOT_SPD, ST_TIMEOUT = \
    0, 1, 2, 3, 4, 5


def load_symbols_only():
    """Load the library without touching the GPU (used by CPU tests)."""
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype, fn.argtypes = res, args
    return L


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m rvgp_amd._build`")
        if not torch.cuda.is_available():
            raise RuntimeError("rvgp_amd requires a ROCm GPU (torch.cuda.is_available() is False)")
        torch.cuda.init()
        _LIB = load_symbols_only()
    return _LIB


def call(name, *args):
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        msg = (lib().rvgp_last_error() or b"").decode(errors="replace")
        if rc == RVGP_EINVAL:
            raise ValueError(msg)
        if rc == RVGP_ENOMEM:
            raise MemoryError(msg)
        raise RuntimeError(f"{name} failed ({rc}): {msg}")
    return rc


def ptr(t):
    """Device pointer of a tensor (or None)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    return ctypes.c_void_p(t.data_ptr())


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def device():
    lib()
    return torch.device("cuda", torch.cuda.current_device())
