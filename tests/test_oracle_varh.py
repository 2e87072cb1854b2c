"""Pinning of the oracle's variable-h symmetric lists (VarHNbrParticleLocator,
base/nnps.pyx:691-786) on the CPU: the reference's golden rows
(base/tests/test_variable_h.py:14-81), brute-force sets, the row order
(own query in stencil order, then the appended sources in ascending index)
and symmetry."""
import json
import os

import numpy as np
import pytest

from oracle import cpu_ref

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "reference_fixtures.json")))


def _varh_1d():
    f = FIX["varh_1d"]
    x = np.linspace(0, 1, f["n"])
    h = np.full(f["n"], (x[1] - x[0]) + f["h_extra"])
    h[f["wide"]] *= 2
    z = np.zeros(f["n"])
    return {"x": x, "y": z, "z": z.copy(), "h": h}, f


def test_golden_rows_1d():
    d, f = _varh_1d()
    cs = cpu_ref.compute_cell_size(d["h"].max(), f["radius_scale"])
    off, idx = cpu_ref.nnps_varh(d, d, f["radius_scale"], cs)
    for i, want in enumerate(f["rows"]):
        assert sorted(idx[off[i]:off[i + 1]].tolist()) == want, i
    # the plain FixedDest rows miss the wide particle's reach
    off1, idx1 = cpu_ref.nnps(d, d, f["radius_scale"], cs, True)
    assert 5 not in idx1[off1[1]:off1[2]].tolist()


def _d2(dst, src):
    """cPoint_distance2(src, dst) for every (dst, src) pair, summed x, y, z."""
    ex = src["x"][None, :] - dst["x"][:, None]
    ey = src["y"][None, :] - dst["y"][:, None]
    ez = src["z"][None, :] - dst["z"][:, None]
    return (ex * ex + ey * ey) + ez * ez


def _cloud(seed, n, dim, spread=(0.5, 1.0)):
    rng = np.random.default_rng(seed)
    pts = rng.uniform(-1.0, 1.0, size=(3, n))
    pts[dim:] = 0.0
    h = 0.12 * rng.uniform(*spread, n)
    return {"x": pts[0], "y": pts[1], "z": pts[2], "h": h}


def _check_rows(dst, src, rs, cs, ns_real, same):
    off, idx = cpu_ref.nnps_varh(dst, src, rs, cs, ns_real=ns_real)
    nd, ns = len(dst["x"]), len(src["x"])
    d2 = _d2(dst, src)
    rd = (dst["h"] * rs) ** 2
    rj = (src["h"] * rs) ** 2
    own = d2 < rd[:, None]
    rev = (d2 < rj[None, :]) & (np.arange(ns) < ns_real)[None, :]
    # the plain lists in reference order, own index not moved
    foff, fidx = cpu_ref.nnps(dst, src, rs, cs, False)
    for i in range(nd):
        row = idx[off[i]:off[i + 1]].tolist()
        assert len(row) == len(set(row)), "duplicate entry"
        assert set(row) == set(np.nonzero(own[i] | rev[i])[0].tolist()), i
        nf = int(own[i].sum())
        assert row[:nf] == fidx[foff[i]:foff[i + 1]].tolist(), "own query not first, in order"
        tail = row[nf:]
        assert tail == sorted(tail), "appended sources not in ascending index"
        assert set(tail) == set(np.nonzero(rev[i] & ~own[i])[0].tolist())
    return off, idx


@pytest.mark.parametrize("seed,n,dim,cell_fac", [(0, 700, 3, 2.0), (1, 600, 2, 2.0),
                                                  (2, 300, 1, 2.0), (3, 500, 3, 0.7)])
def test_rows_vs_brute_force_one_array(seed, n, dim, cell_fac):
    d = _cloud(seed, n, dim)
    cs = cpu_ref.compute_cell_size(d["h"].max(), cell_fac)
    off, idx = _check_rows(d, d, 2.0, cs, n, True)
    # symmetric: j in row(i) <=> i in row(j)
    pairs = {(i, int(j)) for i in range(n) for j in idx[off[i]:off[i + 1]]}
    assert all((j, i) in pairs for i, j in pairs)
    assert any(len(idx[off[i]:off[i + 1]]) > 0 for i in range(n))


def test_rows_two_arrays_and_non_real_sources():
    dst = _cloud(5, 400, 3)
    src = _cloud(6, 500, 3, spread=(0.3, 1.4))
    cs = cpu_ref.compute_cell_size(max(dst["h"].max(), src["h"].max()), 2.0)
    _check_rows(dst, src, 2.0, cs, 500, False)
    # only real sources (the first ns_real) reach back
    off, idx = _check_rows(dst, src, 2.0, cs, 350, False)
    own = _d2(dst, src) < ((dst["h"] * 2.0) ** 2)[:, None]
    for i in range(len(dst["x"])):
        extra = set(idx[off[i]:off[i + 1]].tolist()) - set(np.nonzero(own[i])[0].tolist())
        assert all(j < 350 for j in extra)


def test_uniform_h_equals_fixed_dest_without_own_move():
    d = _cloud(7, 500, 3, spread=(1.0, 1.0))
    cs = cpu_ref.compute_cell_size(d["h"].max(), 2.0)
    off, idx = cpu_ref.nnps_varh(d, d, 2.0, cs)
    foff, fidx = cpu_ref.nnps(d, d, 2.0, cs, False)
    assert np.array_equal(off, foff) and np.array_equal(idx, fidx)
    # the FixedDest lists move the own index last; the variable-h lists do not
    soff, sidx = cpu_ref.nnps(d, d, 2.0, cs, True)
    assert not np.array_equal(sidx, idx)
