"""Launch-shape autotuner for the batched-fit kernels.

A kernel with several valid launch variants (tile edge of the ragged Gram,
panel width of the fused Cholesky step) is timed once per problem shape on
the caller's own operands, with HIP events on the launch stream; the fastest
variant is remembered for the rest of the process under a key made of the
kernel name, the device name and the shape.  Setting RVGP_AUTOTUNE_CACHE to a
file path also keeps the table on disk (JSON), so later processes skip the
timing runs.  A variant the library rejects for the shape (for instance a
panel whose LDS footprint exceeds 160 KiB) is skipped.
"""
import json
import os
import threading

import torch

_TABLE = {}
_LOCK = threading.Lock()
_LOADED = False


def _path():
    return os.environ.get("RVGP_AUTOTUNE_CACHE")


def _load():
    global _LOADED
    if _LOADED:
        return
    _LOADED = True
    p = _path()
    if p and os.path.exists(p):
        try:
            with open(p) as fh:
                _TABLE.update(json.load(fh))
        except (OSError, ValueError):
            pass


def _save():
    p = _path()
    if not p:
        return
    tmp = p + ".tmp"
    os.makedirs(os.path.dirname(os.path.abspath(p)), exist_ok=True)
    with open(tmp, "w") as fh:
        json.dump(_TABLE, fh, indent=1, sort_keys=True)
    os.replace(tmp, p)


def table():
    """A copy of the current choices {key: {"choice": c, "ms": {c: t}}}."""
    with _LOCK:
        _load()
        return json.loads(json.dumps(_TABLE))


def choose(kernel, shape_key, candidates, run, reps=3):
    """Return the fastest of `candidates` for (kernel, shape_key).

    run(c) enqueues one launch of variant c on the current stream and raises
    ValueError (or RuntimeError) if the library rejects c.  Each surviving candidate is run
    once to warm up, then timed over `reps` launches."""
    dev = torch.cuda.get_device_name(torch.cuda.current_device())
    key = f"{kernel}|{dev}|{shape_key}"
    with _LOCK:
        _load()
        hit = _TABLE.get(key)
    if hit is not None and hit["choice"] in [str(c) for c in candidates]:
        return _decode(hit["choice"], candidates)
    times = {}
    for c in candidates:
        try:
            run(c)
        except (ValueError, RuntimeError):
            continue
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            run(c)
        e1.record()
        e1.synchronize()
        times[str(c)] = e0.elapsed_time(e1) / reps
    if not times:
        raise RuntimeError(f"autotune: no valid variant of {kernel} for {shape_key}")
    best = min(times, key=times.get)
    with _LOCK:
        _TABLE[key] = {"choice": best, "ms": times}
        _save()
    return _decode(best, candidates)


def _decode(s, candidates):
    for c in candidates:
        if str(c) == s:
            return c
    raise KeyError(s)
