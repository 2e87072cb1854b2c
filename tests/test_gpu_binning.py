"""Binning of crowded cells.  One particle with a large h sets the cell size
(radius_scale * max h), so the others crowd into a few cells; the counting
sort must still order every cell's members by ascending index (base/cell.pyx:
1229-1306) at a cost that does not grow with the square of a cell's size.
Cells of a few members up to 20,000, across the sizes where the kernel
switches from counting to the merge sort (more than 1,024 members) and where
the merge sort needs several passes (more than 2,048)."""
import time

import numpy as np
import pytest

from oracle import cpu_ref

pytestmark = pytest.mark.gpu


def _clusters(sizes, seed=3, small_h=1e-3):
    """Clusters of the given sizes, each inside its own cell of side 2 (the
    cell size set by one particle of h = 1), plus a few scattered particles;
    indices interleaved over the clusters."""
    rng = np.random.default_rng(seed)
    pts = []
    for k, m in enumerate(sizes):
        c = np.array([2.0 * k + 0.05, 0.05, 0.05])
        pts.append(c + rng.uniform(0.0, 1.9, (m, 3)))
    pts.append(rng.uniform(-4.0, 4.0 + 2.0 * len(sizes), (50, 3)))
    p = np.concatenate(pts)
    p = p[rng.permutation(len(p))]
    h = np.full(len(p), small_h)
    h[7] = 1.0
    return {"x": p[:, 0].copy(), "y": p[:, 1].copy(), "z": p[:, 2].copy(), "h": h,
            "m": np.ones(len(p)), "rho": np.ones(len(p))}


@pytest.mark.parametrize("sizes", [(20000,), (1024, 1025, 2048, 2049, 300, 5000, 4097)])
def test_crowded_cells_sorted_by_index(gpu, sizes):
    import torch
    from pysph_amd.base import api as base
    d = _clusters(sizes)
    n = len(d["x"])
    pa = base.get_particle_array(name="fluid", **d)
    particles = base.Particles([pa])
    cm = particles.cell_manager
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cm.update()
    torch.cuda.synchronize()
    t_first = time.perf_counter() - t0
    pa.set_dirty(True)   # a second binning (device-side cell count)
    t0 = time.perf_counter()
    cm.update()
    torch.cuda.synchronize()
    t_second = time.perf_counter() - t0
    b = cm.bins["fluid"]
    cell_of = b.cell_of.cpu().numpy()[:n]
    perm = b.perm.cpu().numpy()[:n]
    rank = b.rank.cpu().numpy()[:n]
    want = np.lexsort((np.arange(n), cell_of)).astype(np.int32)
    counts = np.bincount(cell_of)
    print(f"{n} particles, largest cell {counts.max()} members: binning {1e3 * t_first:.1f} ms "
          f"(first, synchronous), {1e3 * t_second:.1f} ms (device-side)")
    assert np.array_equal(perm, want)
    assert np.array_equal(rank[perm], np.arange(n))
    assert t_second < 1.0
    # lists over the crowded cells equal the oracle's, order included
    loc = particles.nnps_manager.get_neighbor_particle_locator(pa, pa, 2.0)
    off, idx = (t.cpu().numpy() for t in loc.get_csr())
    full = {k: pa.get(k) for k in ("x", "y", "z", "h")}
    roff, ridx = cpu_ref.nnps(full, full, 2.0, cm.cell_size, True, nthreads=16)
    assert np.array_equal(off, roff) and np.array_equal(idx, ridx)


@pytest.mark.parametrize("order", ["linear", "morton"])
def test_large_sparse_grid(gpu, order):
    """30,000 particles spread over a grid of 198^3 cells: the linear keys
    (7.8M cells, 1,895 tiles of the two-launch scan of the cell counts) and
    the Morton keys (2^24 cells, past that scan's tile limit: the multi-level
    scans) give the same cell starts, non-empty cell list and order inside
    the cells as a host sort of the cell keys, and lists equal to the
    oracle's."""
    import torch
    from pysph_amd.base import api as base
    rng = np.random.default_rng(17)
    n = 30000
    d = {k: rng.uniform(0.0, 190.0, n) for k in ("x", "y", "z")}
    d.update(h=np.full(n, 0.5), m=np.ones(n), rho=np.ones(n))
    pa = base.get_particle_array(name="fluid", **d)
    particles = base.Particles([pa], cell_order=order)
    cm = particles.cell_manager
    cm.update()
    pa.set_dirty(True)   # a second binning (device-side cell count)
    cm.update()
    torch.cuda.synchronize()
    b = cm.bins["fluid"]
    nc = int(cm.grid.ncells)
    assert nc > (7_000_000 if order == "linear" else 16_000_000), nc
    cell_of = b.cell_of.cpu().numpy()[:n]
    want_perm = np.lexsort((np.arange(n), cell_of)).astype(np.int32)
    assert np.array_equal(b.perm.cpu().numpy()[:n], want_perm)
    keys = np.unique(cell_of)
    start = b.cell_start.cpu().numpy()
    assert start.shape[0] == nc + 1
    want_start = np.searchsorted(np.sort(cell_of), np.arange(nc + 1), side="left")
    assert np.array_equal(start, want_start.astype(start.dtype))
    ncells = int(b.ncells_dev.item())
    assert ncells == len(keys)
    assert np.array_equal(b.cells.cpu().numpy()[:ncells], keys.astype(np.int32))
    loc = particles.nnps_manager.get_neighbor_particle_locator(pa, pa, 2.0)
    off, idx = (t.cpu().numpy() for t in loc.get_csr())
    full = {k: pa.get(k) for k in ("x", "y", "z", "h")}
    roff, ridx = cpu_ref.nnps(full, full, 2.0, cm.cell_size, True, nthreads=16)
    assert np.array_equal(off, roff) and np.array_equal(idx, ridx)
    print(f"{order}: {nc} cells, {ncells} non-empty, {len(idx)} list entries")


def test_two_list_scatter_equals_two_scatters(gpu):
    """sph_scatter_sorted2 (the records and the compact (x, y, z, m) rows in
    one pass) writes what two sph_scatter_sorted calls write, and both equal
    the numpy scatter dst[rank[i]] = src[i]; an empty second list is the
    plain scatter."""
    import ctypes

    import torch

    from pysph_amd import _lib
    ctx = _lib.context(torch.device("cuda", 0))
    rng = np.random.default_rng(17)
    n = 100_003
    rank = torch.from_numpy(rng.permutation(n).astype(np.int32)).cuda()
    src = [torch.from_numpy(rng.standard_normal(n)).cuda() for _ in range(11)]
    p = _lib.ptr

    def lists(rec, pm, nb=4):
        a = _lib.PropList()
        a.nprops, a.dst_stride = 11, 12
        for k in range(11):  # fields 0..10 of the 12-double rows (pairs adjacent)
            a.src[k], a.dst[k] = p(src[k]), rec.data_ptr() + 8 * k
        b = _lib.PropList()
        b.nprops, b.dst_stride = nb, 4
        for k, f in enumerate((0, 1, 2, 3)[:nb]):
            b.src[k], b.dst[k] = p(src[f]), pm.data_ptr() + 8 * k
        return a, b

    rec1 = torch.full((n, 12), -1.0, dtype=torch.float64, device="cuda")
    pm1 = torch.full((n, 4), -1.0, dtype=torch.float64, device="cuda")
    rec2, pm2 = rec1.clone(), pm1.clone()
    a, b = lists(rec1, pm1)
    ctx.check(ctx.lib.sph_scatter_sorted2(ctx.handle, n, p(rank), ctypes.byref(a), ctypes.byref(b)))
    a, b = lists(rec2, pm2)
    ctx.check(ctx.lib.sph_scatter_sorted(ctx.handle, n, p(rank), ctypes.byref(a)))
    ctx.check(ctx.lib.sph_scatter_sorted(ctx.handle, n, p(rank), ctypes.byref(b)))
    torch.cuda.synchronize()
    assert torch.equal(rec1, rec2) and torch.equal(pm1, pm2)
    r = rank.cpu().numpy()
    want = np.full((n, 12), -1.0)
    for k in range(11):
        want[r, k] = src[k].cpu().numpy()
    assert np.array_equal(rec1.cpu().numpy(), want)
    assert np.array_equal(pm1.cpu().numpy(), want[:, :4])
    rec3 = torch.full((n, 12), -1.0, dtype=torch.float64, device="cuda")
    a, b = lists(rec3, pm1, nb=0)
    ctx.check(ctx.lib.sph_scatter_sorted2(ctx.handle, n, p(rank), ctypes.byref(a), ctypes.byref(b)))
    torch.cuda.synchronize()
    assert np.array_equal(rec3.cpu().numpy(), want)
