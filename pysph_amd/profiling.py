"""Per-call device timing with HIP events on the stream the kernels run on.

The library launches on torch's current stream (the Context follows it),
so events recorded on that stream bracket exactly the kernels of one C-ABI
call.  Disabled by default; ``enable()`` turns recording on.
"""
import contextlib
from collections import defaultdict

import torch

_enabled = False
_records = defaultdict(list)
_pairs = 0


def enable(on=True):
    global _enabled
    _enabled = bool(on)


def reset():
    global _pairs
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
    if not _enabled:
        yield
        return
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
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
