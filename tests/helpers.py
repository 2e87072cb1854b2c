"""Shared test helpers: error metrics and oracle runs of the configs."""
import numpy as np


def rel_err(a, ref, floor_frac=1e-6):
    """max_i |a_i - ref_i| / max(|ref_i|, floor_frac * max|ref|).

    The floor keeps components that cancel to ~0 (interior accelerations of
    a lattice) from dividing by zero; with floor_frac = 1e-6 a component
    passes a 1e-6 bar only if its absolute error is below 1e-12 * max|ref|."""
    a = np.asarray(a, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    if ref.size == 0:
        return 0.0
    scale = np.max(np.abs(ref))
    den = np.maximum(np.abs(ref), floor_frac * scale if scale > 0 else 1e-300)
    return float(np.max(np.abs(a - ref) / den))


def bit_equal_frac(a, ref):
    a = np.asarray(a)
    ref = np.asarray(ref)
    if a.size == 0:
        return 1.0
    return float(np.mean(a.view(np.int64) == ref.view(np.int64)))


def rows_as_sets_equal(off_a, idx_a, off_b, idx_b):
    if len(off_a) != len(off_b):
        return False
    if not np.array_equal(np.diff(off_a), np.diff(off_b)):
        return False
    for i in range(len(off_a) - 1):
        ra = np.sort(idx_a[off_a[i]:off_a[i + 1]])
        rb = np.sort(idx_b[off_b[i]:off_b[i + 1]])
        if not np.array_equal(ra, rb):
            return False
    return True


def sorted_rows(off, idx):
    """Row-wise sorted copy of a CSR index array (vectorised)."""
    n = len(off) - 1
    row = np.repeat(np.arange(n), np.diff(off))
    order = np.lexsort((idx, row))
    return idx[order]
