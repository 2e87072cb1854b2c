"""Cell binning and neighbour lists on the device (pysph.base.cell / nnps).

``CellManager`` (base/cell.pyx:742-1643): one cell grid shared by all binned
arrays, cell size ``radius_scale * max(h)`` clamped by min/max cell size
(:911-941).  ``update`` re-bins lazily when any array is dirty (:943-976)
with a counting sort into Morton-ordered cells (``sph_grid_build`` +
``sph_bin_array``).

``FixedDestNbrParticleLocator`` (base/nnps.pyx:385-686): one neighbour list
per destination particle, cached as a CSR (int64 offsets, int32 indices in
the source's particle order) and rebuilt lazily when the cell manager has
re-binned (``sph_nnps_build``).  Rows hold the reference's order: stencil
cells x -> y -> z, ascending index inside a cell, own index last.

``NNPSManager`` (base/nnps.pyx:791-932): one locator per
(source, destination, radius_scale).
"""
import ctypes

import numpy as np
import torch

from .. import _lib


class NeighborLocatorType:
    """base/nnps.pyx NeighborLocatorType values used by the API."""
    SPHNeighborLocator = 0
    NSquareNeighborLocator = 1


class _Bins:
    """Per-array result of the counting sort."""

    __slots__ = ("cell_start", "perm", "sx", "sy", "sz", "cell_of", "n")

    def __init__(self, n, ncells, device):
        self.n = n
        self.cell_start = torch.empty(ncells + 1, dtype=torch.int32, device=device)
        self.perm = torch.empty(max(n, 1), dtype=torch.int32, device=device)
        self.sx = torch.empty(max(n, 1), dtype=torch.float64, device=device)
        self.sy = torch.empty(max(n, 1), dtype=torch.float64, device=device)
        self.sz = torch.empty(max(n, 1), dtype=torch.float64, device=device)
        self.cell_of = torch.empty(max(n, 1), dtype=torch.int32, device=device)


class CellManager:
    """Device cell grid over a list of particle arrays (base/cell.pyx:742-1643)."""

    ORDERS = {"morton": _lib.ORDER_MORTON, "linear": _lib.ORDER_LINEAR}

    def __init__(self, arrays_to_bin=(), min_cell_size=-1.0, max_cell_size=0.0, initialize=True,
                 radius_scale=2.0, domain=None, order="morton", max_cells=1 << 28):
        self.arrays_to_bin = list(arrays_to_bin)
        self.min_cell_size = float(min_cell_size)
        self.max_cell_size = float(max_cell_size)
        self.radius_scale = float(radius_scale)
        self.domain = domain
        self.order = order
        self.max_cells = int(max_cells)
        self.array_indices = {a.name: i for i, a in enumerate(self.arrays_to_bin)}
        self.cell_size = 0.0
        self.grid = None
        self.bins = {}
        self.version = 0
        self.is_dirty = True
        self.initialized = False
        self.jump_tolerance = 1
        if initialize:
            self.initialize()

    def initialize(self):
        """Mark the arrays for binning.  The first binning happens at the
        first neighbour query (the lists are lazy anyway), so building a
        Particles/Solver touches no device."""
        if self.initialized:
            return
        self.is_dirty = True
        self.initialized = True

    def _ctx(self):
        return _lib.context(self.arrays_to_bin[0].device if self.arrays_to_bin else None)

    def update_status(self):
        for a in self.arrays_to_bin:
            if a.is_dirty:
                self.is_dirty = True

    def update(self):
        """Re-bin if any array moved (base/cell.pyx:943-976)."""
        self.update_status()
        if not self.is_dirty:
            return 0
        arrays = self.arrays_to_bin
        if not arrays:
            self.is_dirty = False
            return 0
        ctx = self._ctx()
        lib = ctx.lib
        n = len(arrays)
        xs = (ctypes.c_void_p * n)(*[_lib.ptr(a.get_carray("x")) for a in arrays])
        ys = (ctypes.c_void_p * n)(*[_lib.ptr(a.get_carray("y")) for a in arrays])
        zs = (ctypes.c_void_p * n)(*[_lib.ptr(a.get_carray("z")) for a in arrays])
        hs = (ctypes.c_void_p * n)(*[_lib.ptr(a.get_carray("h")) for a in arrays])
        ns = (ctypes.c_int64 * n)(*[a.get_number_of_particles() for a in arrays])
        grid = _lib.SphGrid()
        ctx.check(lib.sph_grid_build(ctx.handle, n, xs, ys, zs, hs, ns, self.radius_scale,
                                     self.min_cell_size, self.max_cell_size,
                                     self.ORDERS[self.order], self.max_cells,
                                     ctypes.byref(grid)))
        self.grid = grid
        self.cell_size = grid.cell_size
        for a in arrays:
            np_ = a.get_number_of_particles()
            b = self.bins.get(a.name)
            if b is None or b.n != np_ or b.cell_start.numel() != grid.ncells + 1:
                b = _Bins(np_, grid.ncells, a.device)
                self.bins[a.name] = b
            ctx.check(lib.sph_bin_array(ctx.handle, ctypes.byref(grid),
                                        _lib.ptr(a.get_carray("x")), _lib.ptr(a.get_carray("y")),
                                        _lib.ptr(a.get_carray("z")), np_,
                                        _lib.ptr(b.cell_start), _lib.ptr(b.perm),
                                        _lib.ptr(b.sx), _lib.ptr(b.sy), _lib.ptr(b.sz),
                                        _lib.ptr(b.cell_of)))
        for a in arrays:
            a.set_dirty(False)
        self.version += 1
        self.is_dirty = False
        return 0

    # reference-style accessors
    def get_number_of_cells(self):
        """Number of non-empty cells (len(cells_dict) in the reference)."""
        self.update()
        total = 0
        for b in self.bins.values():
            counts = (b.cell_start[1:] - b.cell_start[:-1])
            total = counts if isinstance(total, int) else total + counts
        if isinstance(total, int):
            return 0
        return int((total > 0).sum().item())

    def cell_counts(self):
        """Host dict cell-key -> particles per array (tests/diagnostics)."""
        self.update()
        out = {}
        for name, b in self.bins.items():
            c = (b.cell_start[1:] - b.cell_start[:-1]).cpu().numpy()
            out[name] = c
        return out


class FixedDestNbrParticleLocator:
    """Cached per-destination neighbour lists (base/nnps.pyx:385-686)."""

    def __init__(self, source, dest, radius_scale, cell_manager=None, h="h"):
        if cell_manager is None:
            raise ValueError("FixedDestNbrParticleLocator needs a CellManager")
        self.source = source
        self.dest = dest
        self.radius_scale = float(radius_scale)
        self.cell_manager = cell_manager
        self.h = h
        self.source_index = cell_manager.array_indices.get(source.name)
        self.dest_index = cell_manager.array_indices.get(dest.name)
        if self.dest_index is None:
            raise ValueError(f"Destination {dest.name} does not exist")
        if self.source_index is None:
            raise ValueError(f"Source {source.name} does not exist")
        self.offsets = None
        self.indices = None
        self.nnz = 0
        self.built_version = -1
        self.is_dirty = True
        self.nbuilds = 0
        self._host = None

    def update_status(self):
        if not self.is_dirty:
            self.is_dirty = self.source.is_dirty or self.dest.is_dirty

    def update(self):
        """Rebuild the cached lists if the cell manager re-binned."""
        cm = self.cell_manager
        cm.update()
        if self.built_version == cm.version and not self.is_dirty:
            return 0
        self._build()
        self.built_version = cm.version
        self.is_dirty = False
        self._host = None
        return 0

    def _build(self):
        cm = self.cell_manager
        ctx = cm._ctx()
        lib = ctx.lib
        dst, src = self.dest, self.source
        nd = dst.num_real_particles
        b = cm.bins[src.name]
        dev = dst.device
        if self.offsets is None or self.offsets.numel() != nd + 1:
            self.offsets = torch.empty(nd + 1, dtype=torch.int64, device=dev)
        if self.indices is None:
            self.indices = torch.empty(max(1, 64 * nd), dtype=torch.int32, device=dev)
        nnz = ctypes.c_int64(0)
        args = lambda: (ctx.handle, ctypes.byref(cm.grid), _lib.ptr(dst.get_carray("x")),
                        _lib.ptr(dst.get_carray("y")), _lib.ptr(dst.get_carray("z")),
                        _lib.ptr(dst.get_carray(self.h)), nd, _lib.ptr(b.cell_start),
                        _lib.ptr(b.perm), _lib.ptr(b.sx), _lib.ptr(b.sy), _lib.ptr(b.sz),
                        b.n, self.radius_scale, int(src is dst), _lib.ptr(self.offsets),
                        _lib.ptr(self.indices), self.indices.numel(), ctypes.byref(nnz))
        rc = ctx.check(lib.sph_nnps_build(*args()), allow=(_lib.SPH_ENEEDCAP,))
        if rc == _lib.SPH_ENEEDCAP:
            self.indices = torch.empty(int(nnz.value * 1.15) + 1024, dtype=torch.int32,
                                       device=dev)
            ctx.check(lib.sph_nnps_build(*args()))
        self.nnz = int(nnz.value)
        self.nbuilds += 1

    def get_csr(self):
        """Device CSR (offsets int64[nd+1], indices int32[nnz]) after update()."""
        self.update()
        return self.offsets, self.indices[:self.nnz]

    def _host_csr(self):
        if self._host is None:
            off, idx = self.get_csr()
            self._host = (off.cpu().numpy(), idx.cpu().numpy())
        return self._host

    def get_nearest_particles(self, dest_p_index, exclude_self=False):
        """Row of one destination as int64 numpy (own index last when src is dest)."""
        self.update()
        off, idx = self._host_csr()
        row = idx[off[dest_p_index]:off[dest_p_index + 1]].astype(np.int64)
        if exclude_self and self.source is self.dest:
            row = row[:-1]
        return row

    py_get_nearest_particles = get_nearest_particles

    def py_update(self):
        return self.update()

    def py_update_status(self):
        return self.update_status()


class NNPSManager:
    """One locator per (source, dest, radius_scale) (base/nnps.pyx:791-932)."""

    def __init__(self, cell_manager=None, variable_h=False,
                 locator_type=NeighborLocatorType.SPHNeighborLocator):
        if variable_h:
            raise NotImplementedError("variable-h locators are not part of this path")
        self.cell_manager = cell_manager
        self.variable_h = variable_h
        self.locator_type = locator_type
        self.particle_locator_cache = {}

    def get_neighbor_particle_locator(self, source, dest=None, radius_scale=1.0):
        if dest is None:
            dest = source
        key = (id(source), id(dest), float(radius_scale))
        loc = self.particle_locator_cache.get(key)
        if loc is None:
            loc = FixedDestNbrParticleLocator(source, dest, radius_scale, self.cell_manager)
            self.particle_locator_cache[key] = loc
        return loc

    def update(self):
        for loc in self.particle_locator_cache.values():
            loc.update_status()
        if self.cell_manager is not None:
            self.cell_manager.update_status()
        return 0

    py_update = update
