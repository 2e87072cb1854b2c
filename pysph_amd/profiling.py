"""Per-call device timing with HIP events on the stream the kernels run on.

The library launches on torch's current stream (the Context follows it),
so events recorded on that stream bracket exactly the kernels of one C-ABI
call.  Disabled by default; ``enable()`` turns recording on.
"""
import contextlib
import ctypes
import os
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


_HIP_EVENT_DISABLE_SYSTEM_FENCE = 0x20000000
_hip = None


def _hip_runtime():
    """The HIP runtime torch loaded (the one every launch here goes through)."""
    global _hip
    if _hip is None:
        path = os.path.join(os.path.dirname(torch.__file__), "lib", "libamdhip64.so")
        hip = ctypes.CDLL(path if os.path.exists(path) else "libamdhip64.so")
        vp = ctypes.c_void_p
        hip.hipEventCreateWithFlags.argtypes = [ctypes.POINTER(vp), ctypes.c_uint]
        hip.hipEventRecord.argtypes = [vp, vp]
        hip.hipEventSynchronize.argtypes = [vp]
        hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        _hip = hip
    return _hip


class _TimingEvent:
    """A HIP timing event recorded without the system-scope fence of a
    default event (hipEventDisableSystemFence): it only timestamps the
    stream, so the writeback and invalidation of the caches that a default
    event's fence makes -- about 6 us of idle GPU per record on gfx950 --
    stay out of the timed work.  Nothing reads memory through it."""

    def __init__(self):
        self._hip = _hip_runtime()
        self.h = ctypes.c_void_p()
        rc = self._hip.hipEventCreateWithFlags(ctypes.byref(self.h),
                                               _HIP_EVENT_DISABLE_SYSTEM_FENCE)
        if rc != 0:
            raise RuntimeError(f"hipEventCreateWithFlags failed ({rc})")

    def record(self):
        stream = torch.cuda.current_stream().cuda_stream
        rc = self._hip.hipEventRecord(self.h, ctypes.c_void_p(stream))
        if rc != 0:
            raise RuntimeError(f"hipEventRecord failed ({rc})")

    def elapsed_time(self, end):
        """Milliseconds from this event to `end` (waits for `end`)."""
        rc = self._hip.hipEventSynchronize(end.h)
        ms = ctypes.c_float(0.0)
        if rc == 0:
            rc = self._hip.hipEventElapsedTime(ctypes.byref(ms), self.h, end.h)
        if rc != 0:
            raise RuntimeError(f"hipEventElapsedTime failed ({rc})")
        return float(ms.value)


def _event():
    return _pool.pop() if _pool else _TimingEvent()


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
