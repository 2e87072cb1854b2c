"""Per-call device timing with HIP events on the stream the kernels run on.

The library launches on torch's current stream (the Context follows it),
so events recorded on that stream bracket exactly the kernels of one C-ABI
call.  Disabled by default; ``enable()`` turns recording on.
"""
import contextlib
from collections import defaultdict

import torch

_enabled = False
_only = None          # region names that record events (None: all)
_drain = False        # synchronise before each region's start event
_records = defaultdict(list)
_pool = []            # timing events kept for reuse
_pairs = 0


def enable(on=True, regions=None, drain=False):
    """Turn recording on or off; ``regions`` limits the event pairs to the
    named regions (pair counting is on either way), which keeps the
    instrumentation of a timed loop to the kernels it measures.  ``drain``
    synchronises before each region, so that a region's time is its own
    kernels plus their launch latency and never includes work queued before
    it or the GPU's idle time while the host waits inside an earlier call
    (for per-call breakdowns, not for timed loops)."""
    global _enabled, _only, _drain
    _enabled = bool(on)
    _only = set(regions) if regions is not None else None
    _drain = bool(drain)


def _event():
    return _pool.pop() if _pool else torch.cuda.Event(enable_timing=True)


def reset():
    global _pairs
    for evs in _records.values():
        for s, e in evs:
            _pool.append(s)
            _pool.append(e)
    _records.clear()
    _pending_pairs.clear()
    _pairs = 0


_pending_pairs = []


def count_pairs(n):
    """Add n pair interactions (list entries evaluated by an operator): an
    int, or a one-element device tensor read when pairs() is called."""
    global _pairs
    if _enabled:
        if isinstance(n, torch.Tensor):
            _pending_pairs.append(n)
        else:
            _pairs += int(n)


def pairs():
    """Total pair interactions counted (synchronises for device counts)."""
    global _pairs
    if _pending_pairs:
        _pairs += int(torch.stack([t.reshape(()) for t in _pending_pairs]).sum().item())
        _pending_pairs.clear()
    return _pairs


@contextlib.contextmanager
def region(name):
    if not _enabled or (_only is not None and name not in _only):
        yield
        return
    s = _event()
    e = _event()
    if _drain:
        torch.cuda.synchronize()
    s.record()
    try:
        yield
    finally:
        e.record()
        _records[name].append((s, e))


def summary():
    """{name: {"count": n, "mean_ms": m, "total_ms": t}} (synchronises)."""
    torch.cuda.synchronize()
    out = {}
    for name, evs in _records.items():
        ts = [s.elapsed_time(e) for s, e in evs]
        out[name] = {"count": len(ts), "mean_ms": sum(ts) / len(ts), "total_ms": sum(ts)}
    return out
