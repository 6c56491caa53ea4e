"""Optional per-launch timing of the hot kernels with HIP events.

When a Recorder is active, every BSR SpMM and GEMM launch issued through
sparse.spmm / dense.gemm is bracketed by events on the stream it is launched
on (torch's current stream), together with its algorithmic byte and flop
counts.  Used by bench.py to compute roofline fractions; inactive otherwise.
"""
import torch

_ACTIVE = None


class Recorder:
    def __init__(self):
        self.records = []  # (kind, key, bytes, flops, ev0, ev1)

    def __enter__(self):
        global _ACTIVE
        self._prev = _ACTIVE
        _ACTIVE = self
        return self

    def __exit__(self, *a):
        global _ACTIVE
        _ACTIVE = self._prev

    def summary(self):
        torch.cuda.synchronize()
        out = {}
        for kind, key, b, f, e0, e1 in self.records:
            ms = e0.elapsed_time(e1)
            s = out.setdefault((kind, key), {"count": 0, "ms": 0.0, "bytes": 0.0, "flops": 0.0})
            s["count"] += 1
            s["ms"] += ms
            s["bytes"] += b
            s["flops"] += f
        return out


def active():
    return _ACTIVE


def begin():
    if _ACTIVE is None:
        return None
    e = torch.cuda.Event(enable_timing=True)
    e.record()
    return e


def end(e0, kind, key, nbytes, flops):
    if e0 is None or _ACTIVE is None:
        return
    e1 = torch.cuda.Event(enable_timing=True)
    e1.record()
    _ACTIVE.records.append((kind, key, float(nbytes), float(flops), e0, e1))
