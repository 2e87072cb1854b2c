"""sph_extrema_noalloc (the reductions of the time-step functions with their
partials in the library's file-scope device array, results in device memory)
against sph_extrema (context scratch, results on the host), called through
the C ABI: two calls in a row, the second with a smaller input, each bit-equal
to sph_extrema on the same arrays; NaN propagation and the empty input."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _extrema(ctx, n, arrays):
    from pysph_amd import _lib
    p = [_lib.ptr(a) if a is not None else None for a in arrays] + [None] * (3 - len(arrays))
    out = (ctypes.c_double * 3)()
    ctx.check(ctx.lib.sph_extrema(ctx.handle, n, len(arrays), p[0], p[1], p[2], out))
    return np.array(out[:])


def _extrema_noalloc(ctx, n, arrays, out_dev):
    from pysph_amd import _lib
    p = [_lib.ptr(a) if a is not None else None for a in arrays] + [None] * (3 - len(arrays))
    out_dev.fill_(12345.0)
    ctx.check(ctx.lib.sph_extrema_noalloc(ctx.handle, n, len(arrays), p[0], p[1], p[2],
                                          _lib.ptr(out_dev)))
    torch.cuda.synchronize()
    return out_dev.cpu().numpy().copy()


def _same(a, b):
    return np.array_equal(np.asarray(a).view(np.int64), np.asarray(b).view(np.int64))


def test_noalloc_extrema_matches_allocating_entry_point(gpu):
    ctx = gpu
    rng = np.random.default_rng(20261203)
    out_dev = torch.empty(3, dtype=torch.float64, device="cuda")
    # a large call (every one of the 512 partial blocks used), then smaller ones
    for n in (3_000_001, 777, 1):
        for ncomp in (1, 2, 3):
            arrs = [torch.from_numpy(rng.normal(size=n) * 10.0).cuda() for _ in range(ncomp)]
            ref = _extrema(ctx, n, arrs)
            got = _extrema_noalloc(ctx, n, arrs, out_dev)
            print(f"n={n} ncomp={ncomp}: {got} (sph_extrema {ref})")
            assert _same(got, ref), (n, ncomp, got, ref)
            # the values themselves, from numpy on the host copies
            h = [a.cpu().numpy() for a in arrs]
            mag = np.abs(h[0]) if ncomp == 1 else np.sqrt(sum(x * x for x in h))
            assert got[0] == mag.max() and got[1] == h[0].min() and got[2] == h[0].max()


def test_noalloc_extrema_nan_and_empty(gpu):
    ctx = gpu
    out_dev = torch.empty(3, dtype=torch.float64, device="cuda")
    a = torch.arange(1000, dtype=torch.float64, device="cuda")
    a[517] = float("nan")
    ref = _extrema(ctx, 1000, [a])
    got = _extrema_noalloc(ctx, 1000, [a], out_dev)
    assert np.all(np.isnan(got)) and np.all(np.isnan(ref))
    ref0 = _extrema(ctx, 0, [a])
    got0 = _extrema_noalloc(ctx, 0, [a], out_dev)
    assert _same(got0, ref0) and got0[0] == -np.inf and got0[1] == np.inf
