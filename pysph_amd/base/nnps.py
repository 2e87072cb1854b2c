"""Cell binning and neighbour lists on the device (pysph.base.cell / nnps).

``CellManager`` (base/cell.pyx:742-1643): one cell grid shared by all binned
arrays, cell size ``radius_scale * max(h)`` clamped by min/max cell size
(:911-941).  ``update`` re-bins lazily when any array is dirty (:943-976)
with a counting sort into Morton-ordered cells (``sph_grid_build`` +
``sph_bin_array``).  The manager also keeps *sorted shadows*: copies of the
particle properties in cell order, which the operator kernels read with
coalesced, cell-coherent loads.  A shadow is refreshed when the binning
changed or the array's property was written since (``ParticleArray.touch``
versions); kernels that produce a property write both the original-order
array and its shadow.

``FixedDestNbrParticleLocator`` (base/nnps.pyx:385-686): one neighbour list
per destination particle, rebuilt lazily after the cell manager re-binned
(``sph_nnps_build``).  Internally the lists live in sorted space (rows per
sorted destination, entries are sorted source positions); ``get_csr`` /
``get_nearest_particles`` export them to original indices in the
reference's order: stencil cells x -> y -> z, ascending index inside a
cell, own index last.

``NNPSManager`` (base/nnps.pyx:791-932): one locator per
(source, destination, radius_scale).
"""
import ctypes
import itertools
import math

import numpy as np
import torch

from .. import _lib, profiling

#: sorted record layout (SPH_R_* in include/sphhip.h): property -> field
REC = 12
REC_FIELDS = {"x": 0, "y": 1, "z": 2, "m": 3, "u": 4, "v": 5, "w": 6, "rho": 8, "cs": 9,
              "h": 10, "p": 11}
PRHO2_FIELD = 7  # P/rho^2, written by the library (SPH_R_PRHO2)
#: fields of the compact position-mass rows, in row order
PM_FIELDS = ("x", "y", "z", "m")
#: stamps of binnings: a _Bins' stamp changes whenever it is re-binned
_BIN_STAMPS = itertools.count(1)


class _SharedEvent:
    """The event that tells the host a device-to-host copy has landed,
    shared by every copy queued on the stream before it is recorded: the
    grid check, the binning and the list builds of one evaluation end in one
    event (one system-scope fence) rather than one each.  Recorded by
    CellManager.fence() after the list builds, or by the first query() or
    synchronize() if nobody has recorded it yet (it then covers everything
    queued so far, which still includes its copies)."""

    __slots__ = ("ev",)

    def __init__(self):
        self.ev = None

    def record(self):
        if self.ev is None:
            self.ev = torch.cuda.Event()
            self.ev.record()

    def query(self):
        self.record()
        return self.ev.query()

    def synchronize(self):
        self.record()
        self.ev.synchronize()


class _PinnedSlots:
    """A few page-locked host buffers for the sizes and status words the
    device-side builds send back, allocated once: steady-state builds make no
    host allocation.  take(ev) returns the next buffer, whose copy the
    caller queues and whose arrival ev (a _SharedEvent) tells; a buffer
    whose last copy has not arrived is waited for first."""

    def __init__(self, n, numel, dtype):
        self.bufs = torch.empty((n, numel), dtype=dtype, pin_memory=True)
        self.events = [None] * n
        self.i = 0

    def take(self, ev):
        k = self.i
        self.i = (k + 1) % len(self.events)
        if self.events[k] is not None:
            self.events[k].synchronize()
        self.events[k] = ev
        return self.bufs[k], ev
SHADOW_PROPS = tuple(REC_FIELDS)


class NeighborLocatorType:
    """base/nnps.pyx NeighborLocatorType values used by the API."""
    SPHNeighborLocator = 0
    NSquareNeighborLocator = 1


class _Bins:
    """Per-array result of the counting sort and its sorted records (one
    row of REC doubles per particle, fields REC_FIELDS)."""

    def __init__(self, n, ncells, device, dtype=torch.float64):
        self.n = n
        self.device = device
        self.cell_start = torch.empty(ncells + 1, dtype=torch.int32, device=device)
        self.perm = torch.empty(max(n, 1), dtype=torch.int32, device=device)
        self.rank = torch.empty(max(n, 1), dtype=torch.int32, device=device)
        self.cell_of = torch.empty(max(n, 1), dtype=torch.int32, device=device)
        self.cells = torch.empty(max(min(n, ncells), 1), dtype=torch.int32, device=device)
        self.ncells_nonempty = 0
        # after the first binning the count of non-empty cells stays on the
        # device (ncells_dev); list builds are launched for cells_bound cells
        # and the host learns the count one binning later (pending)
        self.ncells_dev = None
        self.cells_bound = 0
        self.pending = None
        self.rec = torch.zeros((max(n, 1), REC), dtype=dtype, device=device)
        self.shadow_ver = {}
        # compact sorted (x, y, z, m) rows (FP64 only): the neighbour gathers
        # of density and momentum touch fewer cache lines than in the records
        self.pm = None
        self.pm_ver = {}
        self.stamp = 0
        self.host_slots = None

    def poll(self):
        """Take the device cell count of the last binning if it has arrived,
        and keep the launch bound at twice it."""
        if self.pending is None:
            return
        ev, host = self.pending
        if not ev.query():
            return
        self.pending = None
        self.ncells_nonempty = int(host[0])
        if self.ncells_nonempty > self.cells_bound // 2:
            self.cells_bound = min(max(self.n, 1), 2 * self.ncells_nonempty + 4096)

    def field_ptr(self, prop):
        return self.rec.data_ptr() + self.rec.element_size() * REC_FIELDS[prop]

    def invalidate(self):
        self.shadow_ver = {}
        self.pm_ver = {}


def _grid_signature(g):
    return (g.cell_size, tuple(g.cmin), tuple(g.dims), tuple(g.bits), g.order, g.ncells)


class CellManager:
    """Device cell grid over a list of particle arrays (base/cell.pyx:742-1643)."""

    ORDERS = {"morton": _lib.ORDER_MORTON, "linear": _lib.ORDER_LINEAR}

    def __init__(self, arrays_to_bin=(), min_cell_size=-1.0, max_cell_size=0.0, initialize=True,
                 radius_scale=2.0, domain=None, order="morton", max_cells=1 << 28,
                 precision="double"):
        self.arrays_to_bin = list(arrays_to_bin)
        self.min_cell_size = float(min_cell_size)
        self.max_cell_size = float(max_cell_size)
        self.radius_scale = float(radius_scale)
        self.domain = domain
        self.order = order
        self.max_cells = int(max_cells)
        if precision not in ("double", "single"):
            raise ValueError("precision must be 'double' or 'single'")
        # "single": float sorted records, lists and operator arithmetic in
        # float (the reference's cl_precision="single")
        self.precision = precision
        self.rec_dtype = torch.float32 if precision == "single" else torch.float64
        self.array_indices = {a.name: i for i, a in enumerate(self.arrays_to_bin)}
        self.cell_size = 0.0
        self.grid = None
        self.bins = {}
        self._wanted = {}
        self._h_uniform = False
        self._h_versions = {}
        self.version = 0
        self.is_dirty = True
        self.initialized = False
        self.jump_tolerance = 1
        # a grid built with GRID_PAD empty cells around the particles is kept
        # (no bounding-box readback, so no host synchronisation per binning)
        # while h and the array sizes are unchanged, the moved arrays were
        # last written by the integrator's stage kernels, and the device check
        # of the previous binning found every particle at least GRID_MARGIN
        # cells inside it
        self.grid_pad = self.GRID_PAD
        self.keep_grid = True
        self._grid_key = None
        self._grid_flags = None
        self._grid_pending = None
        self._open_event = None
        self._grid_host = None
        self._made = {}  # array -> record fields kernels write themselves
        self.grid_builds = 0
        if initialize:
            self.initialize()

    GRID_PAD = 4
    GRID_MARGIN = 2

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
        with profiling.region("bin"):
            self._rebin(arrays)
        for a in arrays:
            a.set_dirty(False)
        self.version += 1
        self.is_dirty = False
        return 0

    def _grid_state(self, arrays):
        """What the kept grid depends on besides positions: the arrays, their
        sizes and their h."""
        return tuple((a.name, a.get_number_of_particles(), a.version("h")) for a in arrays)

    def _copy_event(self):
        """The shared event of the device-to-host copies queued now."""
        if self._open_event is None or self._open_event.ev is not None:
            self._open_event = _SharedEvent()
        return self._open_event

    def fence(self):
        """Record the event of the device-to-host copies queued since the
        last one (after an evaluation's list builds)."""
        if self._open_event is not None:
            self._open_event.record()
            self._open_event = None

    def _poll_grid_check(self):
        """Flags of the last device grid check (waits for it: it was queued
        one binning ago).  Raises when that binning was invalid."""
        if self._grid_pending is None:
            return 0
        ev, host = self._grid_pending
        self._grid_pending = None
        ev.synchronize()
        flags = int(host[0])
        if flags & 10:
            self._grid_key = None
            what = "a non-finite coordinate" if flags & 8 else "a particle outside the cell grid"
            raise _lib.SphError(_lib.SPH_EDOMAIN,
                                f"binning found {what}: the evaluation that used it is invalid "
                                f"(particles moved more than {self.GRID_MARGIN} cells in one "
                                f"step, or positions were written without touch())")
        return flags

    def _can_keep_grid(self, arrays):
        if not self.keep_grid or self.grid is None or \
                self._grid_key != self._grid_state(arrays):
            return False
        # every moved array was last written by the integrator's stage
        # kernels (any other write of x, y or z rebuilds the grid)
        for a in arrays:
            if a.is_dirty and getattr(a, "_stepped_xyz", None) != (
                    a.version("x"), a.version("y"), a.version("z")):
                return False
        return not (self._poll_grid_check() & 1)

    def _rebin(self, arrays):
        ctx = self._ctx()
        lib = ctx.lib
        n = len(arrays)
        xs = (ctypes.c_void_p * n)(*[_lib.ptr(a.get_carray("x")) for a in arrays])
        ys = (ctypes.c_void_p * n)(*[_lib.ptr(a.get_carray("y")) for a in arrays])
        zs = (ctypes.c_void_p * n)(*[_lib.ptr(a.get_carray("z")) for a in arrays])
        ns = (ctypes.c_int64 * n)(*[a.get_number_of_particles() for a in arrays])
        if self._can_keep_grid(arrays):
            # the same grid: check on the device that the moved arrays still
            # fit it, read the flags at the next binning
            moved = [k for k, a in enumerate(arrays) if a.is_dirty]
            m = len(moved)
            if self._grid_flags is None:
                self._grid_flags = torch.zeros(1, dtype=torch.int32, device=arrays[0].device)
            ctx.check(lib.sph_grid_check(
                ctx.handle, m, (ctypes.c_void_p * m)(*[xs[k] for k in moved]),
                (ctypes.c_void_p * m)(*[ys[k] for k in moved]),
                (ctypes.c_void_p * m)(*[zs[k] for k in moved]),
                (ctypes.c_int64 * m)(*[ns[k] for k in moved]), ctypes.byref(self.grid),
                self.GRID_MARGIN, _lib.ptr(self._grid_flags)))
            # the last check was read above (_poll_grid_check): its buffer is free
            if self._grid_host is None:
                self._grid_host = torch.empty(1, dtype=torch.int32, pin_memory=True)
            host = self._grid_host
            host.copy_(self._grid_flags, non_blocking=True)
            self._grid_pending = (self._copy_event(), host)
            grid = self.grid
            same_grid = True
        else:
            self._grid_pending = None
            hs = (ctypes.c_void_p * n)(*[_lib.ptr(a.get_carray("h")) for a in arrays])
            grid = _lib.SphGrid()
            ctx.check(lib.sph_grid_build(ctx.handle, n, xs, ys, zs, hs, ns, self.radius_scale,
                                         self.min_cell_size, self.max_cell_size,
                                         self.ORDERS[self.order], self.max_cells,
                                         int(self.grid_pad), ctypes.byref(grid)))
            self.grid_builds += 1
            sig = _grid_signature(grid)
            same_grid = self.grid is not None and _grid_signature(self.grid) == sig
            self.grid = grid
            self._grid_key = self._grid_state(arrays)
            self._h_uniform = grid.min_h == grid.max_h
            self._h_versions = {a.name: a.version("h") for a in arrays}
            self.cell_size = grid.cell_size
        for a in arrays:
            np_ = a.get_number_of_particles()
            b = self.bins.get(a.name)
            if same_grid and b is not None and b.n == np_ and not a.is_dirty:
                continue  # unmoved array on an unchanged grid: its bins still hold
            if b is None or b.n != np_ or b.cell_start.numel() != grid.ncells + 1:
                b = _Bins(np_, grid.ncells, a.device, self.rec_dtype)
                self.bins[a.name] = b
            args = (ctx.handle, ctypes.byref(grid), _lib.ptr(a.get_carray("x")),
                    _lib.ptr(a.get_carray("y")), _lib.ptr(a.get_carray("z")), np_,
                    _lib.ptr(b.cell_start), _lib.ptr(b.perm), _lib.ptr(b.rank),
                    _lib.ptr(b.cell_of), _lib.ptr(b.cells))
            if b.ncells_dev is None:
                # first binning of these bins: read the count (synchronises)
                nne = ctypes.c_int64(0)
                ctx.check(lib.sph_bin_array(*args, ctypes.byref(nne)))
                b.ncells_nonempty = int(nne.value)
                b.cells_bound = min(max(np_, 1), 2 * b.ncells_nonempty + 4096)
                b.ncells_dev = torch.full((1,), b.ncells_nonempty, dtype=torch.int32,
                                          device=a.device)
            else:
                b.poll()
                ctx.check(lib.sph_bin_array_dev(*args, _lib.ptr(b.ncells_dev)))
                if b.host_slots is None:
                    b.host_slots = _PinnedSlots(4, 1, torch.int32)
                host, ev = b.host_slots.take(self._copy_event())
                host.copy_(b.ncells_dev, non_blocking=True)
                b.pending = (ev, host)
            b.stamp = next(_BIN_STAMPS)
            b.invalidate()

    def uniform_h(self):
        """True when every binned particle has the same h (as of the last
        binning, with no h written since): the operator kernels then take
        0.5*(h_a+h_b) == h_a without reading h_b."""
        self.update()
        return bool(self._h_uniform) and all(
            a.version("h") == self._h_versions.get(a.name) for a in self.arrays_to_bin)

    # -- sorted records --------------------------------------------------
    def shadows(self, array, props):
        """The sorted record tensor of `array` with the fields of `props`
        refreshed where stale (binning changed or the property was written)."""
        self.update()
        b = self.bins[array.name]
        # every field asked for so far is refreshed with this one launch, so
        # that one evaluation needs about one gather per array
        wanted = self._wanted.setdefault(array.name, [])
        for p in props:
            if p not in wanted:
                wanted.append(p)
        if any(b.shadow_ver.get(p) != array.version(p) for p in props):
            # ... except the fields a kernel writes into the records itself
            # (written(shadow_written=True): density and the equation of
            # state), unless asked for now: after a binning their copies
            # would be overwritten before anything reads them
            made = self._made.get(array.name, ())
            stale = [p for p in wanted if b.shadow_ver.get(p) != array.version(p)
                     and (p in props or p not in made)]
            # in record order, so that the two fields of a record pair are
            # adjacent in the list and go out as one 16-byte store
            stale.sort(key=lambda q: REC_FIELDS[q])
        else:
            stale = []
        if stale:
            ctx = self._ctx()
            single = self.precision == "single"
            pl = _lib.PropListF32() if single else _lib.PropList()
            pl.nprops = len(stale)
            pl.dst_stride = REC
            for k, p in enumerate(stale):
                src = array.get_carray(p)
                if src is None:
                    raise AttributeError(f"property {p} not present in {array.name}")
                if src.dtype != torch.float64:
                    raise TypeError(f"property {p} of {array.name} must be float64")
                pl.src[k] = _lib.ptr(src)
                pl.dst[k] = b.field_ptr(p)
            # original order -> sorted rows by rank: coalesced reads of every
            # field (the gather by perm reads one cache line per lane and field)
            pm = self._pm_props(array, b) if not single else None
            with profiling.region("gather_sorted"):
                if pm is not None:  # the compact rows in the same pass
                    ctx.check(ctx.lib.sph_scatter_sorted2(ctx.handle, b.n, _lib.ptr(b.rank),
                                                          ctypes.byref(pl), ctypes.byref(pm[0])))
                else:
                    scatter = (ctx.lib.sph_scatter_sorted_f32 if single
                               else ctx.lib.sph_scatter_sorted)
                    ctx.check(scatter(ctx.handle, b.n, _lib.ptr(b.rank), ctypes.byref(pl)))
            for p in stale:
                b.shadow_ver[p] = array.version(p)
            if pm is not None:
                for q in pm[1]:
                    b.pm_ver[q] = array.version(q)
        return b.rec

    def _pm_props(self, array, b):
        """(prop list, fields) of the stale fields of the compact rows of
        `array`, once posmass() has made them, else None."""
        if b.pm is None or b.pm.shape[0] != max(b.n, 1):
            return None
        stale = [q for q in PM_FIELDS if b.pm_ver.get(q) != array.version(q)]
        if not stale:
            return None
        pl = _lib.PropList()
        pl.nprops = len(stale)
        pl.dst_stride = 4
        base = b.pm.data_ptr()
        for k, q in enumerate(stale):
            src = array.get_carray(q)
            if src is None or src.dtype != torch.float64:
                raise TypeError(f"property {q} of {array.name} must be float64")
            pl.src[k] = _lib.ptr(src)
            pl.dst[k] = base + 8 * PM_FIELDS.index(q)
        return pl, stale

    def posmass(self, array):
        """The compact sorted (x, y, z, m) rows of `array` (4 doubles per
        particle, rows in record order) with stale fields refreshed, or None
        in single precision (the float records are read instead)."""
        if self.precision == "single":
            return None
        self.update()
        b = self.bins[array.name]
        if b.pm is None or b.pm.shape[0] != max(b.n, 1):
            b.pm = torch.empty((max(b.n, 1), 4), dtype=torch.float64, device=b.device)
            b.pm_ver = {}
        pm = self._pm_props(array, b)
        if pm is not None:
            ctx = self._ctx()
            pl, stale = pm
            with profiling.region("gather_sorted"):
                ctx.check(ctx.lib.sph_scatter_sorted(ctx.handle, b.n, _lib.ptr(b.rank),
                                                     ctypes.byref(pl)))
            for q in stale:
                b.pm_ver[q] = array.version(q)
        return b.pm

    def shadow_out(self, array, prop):
        """Device pointer of the record field a kernel may write together
        with `prop` (None when `prop` is not a record field)."""
        if prop not in REC_FIELDS:
            return None
        return self.bins[array.name].field_ptr(prop)

    def written(self, array, prop, shadow_written):
        """Record a kernel write of `prop` (and of its record field if it wrote it)."""
        array.touch(prop)
        b = self.bins.get(array.name)
        if b is not None and shadow_written and prop in REC_FIELDS:
            b.shadow_ver[prop] = array.version(prop)
            self._made.setdefault(array.name, set()).add(prop)

    # P/rho^2 of the records (the per-source term of MomentumEquation): fresh
    # while the record's p and rho are the ones it was computed from
    def prho2_fresh(self, array):
        b = self.bins.get(array.name)
        if b is None:
            return False
        vp, vr = array.version("p"), array.version("rho")
        return (b.shadow_ver.get("_prho2") == (vp, vr) and b.shadow_ver.get("p") == vp
                and b.shadow_ver.get("rho") == vr)

    def prho2_written(self, array):
        b = self.bins[array.name]
        b.shadow_ver["_prho2"] = (array.version("p"), array.version("rho"))

    def sorted_field(self, array, prop):
        """Host copy of one record field in sorted order (diagnostics)."""
        rec = self.shadows(array, (prop,))
        return rec[:self.bins[array.name].n, REC_FIELDS[prop]].cpu().numpy()

    # reference-style accessors
    def get_number_of_cells(self):
        """Number of non-empty cells over all binned arrays (len(cells_dict))."""
        self.update()
        total = None
        for b in self.bins.values():
            counts = b.cell_start[1:] - b.cell_start[:-1]
            total = counts if total is None else total + counts
        return 0 if total is None else int((total > 0).sum().item())

    def cell_counts(self):
        """Host dict array name -> particles per cell (diagnostics)."""
        self.update()
        return {name: (b.cell_start[1:] - b.cell_start[:-1]).cpu().numpy()
                for name, b in self.bins.items()}


class FixedDestNbrParticleLocator:
    """Cached per-destination neighbour lists (base/nnps.pyx:385-686)."""

    def __init__(self, source, dest, radius_scale, cell_manager=None, h="h"):
        if cell_manager is None:
            raise ValueError("FixedDestNbrParticleLocator needs a CellManager")
        if h != "h":
            raise NotImplementedError("the smoothing length property must be 'h'")
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
        # SELL-64 lists in sorted space: counts per sorted destination,
        # slice offsets, entries (see sph_nnps_build)
        self.counts = None
        self.slice_off = None
        self.indices = None
        self._nnz = 0
        self.slots = 0
        # builds after the first run without a host synchronisation; their
        # totals and capacity flags arrive later (_pending) and are checked
        # at the next build or when nnz is read
        self._totals = None
        self._pending = []
        self._status = None
        self.built_version = -1
        self.is_dirty = True
        self.nbuilds = 0
        self._export = None
        self._host = None
        # cells near a source cell (sph_cell_near), for a source that was
        # not re-binned between two builds (walls): key, table, reach
        self._near_seen = None
        self._near = None
        self.near_used = False
        self._host_slots = None
        self._pair_slots = {}  # partner locator -> pinned ring of paired builds

    def update_status(self):
        if not self.is_dirty:
            self.is_dirty = self.source.is_dirty or self.dest.is_dirty

    def _near_table(self, cm, bs):
        """(table, reach) for a cross-array build over a source binning that
        also served the previous build, else (None, 0): a source that moves
        every step would rebuild the table at every build.  The kernels skip
        cells on the flag alone, so the table is only handed over while no
        destination h changed since the grid was built (every destination's
        stencil reach is then at most the one of the grid's max_h)."""
        if self.source is self.dest or bs.ncells_dev is None:
            return None, 0
        g = cm.grid
        # the count pass's reach: (int)ceil(h*rs/cell), largest at max_h
        reach = max(int(math.ceil(g.max_h * self.radius_scale / g.cell_size)), 1)
        if cm._h_versions.get(self.dest.name) != self.dest.version("h") or reach > 2:
            self._near_seen = None
            return None, 0
        key = (bs.stamp, _grid_signature(g), reach)
        seen, self._near_seen = self._near_seen, key
        if key != seen:
            return None, 0
        if self._near is not None and self._near[0] == key:
            return self._near[1], self._near[2]
        dev = self.dest.device
        table = torch.empty(max(int(g.ncells), 1), dtype=torch.uint8, device=dev)
        ctx = cm._ctx()
        src = self.source
        p = _lib.ptr
        ctx.check(ctx.lib.sph_cell_near(
            ctx.handle, ctypes.byref(g), p(bs.cell_start), p(bs.cells), p(bs.ncells_dev),
            bs.cells_bound if bs.n > 0 else 0, p(bs.perm), p(src.get_carray("x")),
            p(src.get_carray("y")), p(src.get_carray("z")), reach, p(table)))
        self._near = (key, table, reach)
        return table, reach

    @property
    def nnz(self):
        """Entries of the current lists (waits for a device-side build)."""
        self._poll(block=True)
        return self._nnz

    @nnz.setter
    def nnz(self, v):
        self._nnz = int(v)

    def pair_count(self):
        """Entries of the current lists: an int, or a one-element device
        tensor while a device-side build's totals are still on the way."""
        if self._pending:
            return self._totals[1:2]
        return self._nnz

    def _poll(self, block=False):
        """Check the device-side builds whose totals have arrived: raise if
        one of them ran out of list capacity or cell bound (the evaluation
        that used it is invalid), and grow the list buffer ahead of need."""
        while self._pending:
            ev, host = self._pending[0]
            if block:
                ev.synchronize()
            elif not ev.query():
                return
            self._pending.pop(0)
            slots, nnz, st = int(host[0]), int(host[1]), int(host[2])
            self.slots = slots
            self._nnz = nnz
            if st:
                self._status.zero_()
                why = []
                if st & 1:
                    why.append(f"{slots} list slots needed, {self.indices.numel()} allocated")
                if st & 16:
                    why.append("more non-empty destination cells than the launch bound")
                raise _lib.SphError(_lib.SPH_ENEEDCAP,
                                    f"neighbour lists {self.source.name}->{self.dest.name} "
                                    f"incomplete ({'; '.join(why)}); the evaluation that used "
                                    "them is invalid")

    def update(self):
        """Rebuild the cached lists if the cell manager re-binned."""
        cm = self.cell_manager
        cm.update()
        if self.built_version == cm.version and not self.is_dirty:
            return 0
        with profiling.region("nnps_build"):
            self._build()
        self._built()
        return 0

    def _device_side(self, bd):
        """Builds after the first run without a host synchronisation."""
        return self.nbuilds > 0 and bd.ncells_dev is not None

    def _buffers(self):
        dst = self.dest
        nd_total = dst.get_number_of_particles()
        dev = dst.device
        nslices = (nd_total + 63) // 64
        if self.counts is None or self.counts.numel() != max(nd_total, 1):
            self.counts = torch.empty(max(nd_total, 1), dtype=torch.int32, device=dev)
            self.slice_off = torch.empty(nslices + 1, dtype=torch.int64, device=dev)
        if self.indices is None:
            self.indices = torch.empty(max(1, 96 * nd_total), dtype=torch.int32, device=dev)

    def _dev_args(self, totals=None):
        """The arguments of a device-side build (sph_nnps_dev_args) and its
        totals buffer (`totals`, or a new one).  The list buffer is grown
        ahead of need from the last known size (the lists in it are replaced
        by the build anyway)."""
        cm = self.cell_manager
        dst, src = self.dest, self.source
        bd, bs = cm.bins[dst.name], cm.bins[src.name]
        dev = dst.device
        self._buffers()
        if self.slots > (3 * self.indices.numel()) // 4:
            self.indices = torch.empty(int(self.slots * 1.5) + 4096, dtype=torch.int32,
                                       device=dev)
        if self._status is None:
            self._status = torch.zeros(1, dtype=torch.int32, device=dev)
        if totals is None:
            totals = torch.empty(3, dtype=torch.int64, device=dev)
        near, reach = self._near_table(cm, bs)
        self.near_used = near is not None
        p = _lib.ptr
        a = _lib.NnpsDevArgs()
        a.f32 = int(cm.precision == "single")
        a.same_array = int(src is dst)
        a.d_rec = p(cm.shadows(dst, ("x", "y", "z", "h")))
        if a.f32:
            a.d_x, a.d_y, a.d_z = (p(dst.get_carray(c)) for c in ("x", "y", "z"))
        a.d_perm, a.d_cell_start, a.d_cells = p(bd.perm), p(bd.cell_start), p(bd.cells)
        a.nd_cells_dev, a.max_cells = p(bd.ncells_dev), bd.cells_bound
        a.nd_total, a.nd_real = dst.get_number_of_particles(), dst.num_real_particles
        a.s_cell_start, a.s_rec, a.ns = p(bs.cell_start), p(cm.shadows(src, ("x", "y", "z"))), bs.n
        a.radius_scale = self.radius_scale
        a.counts, a.slice_off, a.indices = p(self.counts), p(self.slice_off), p(self.indices)
        a.cap = self.indices.numel()
        a.totals, a.status = p(totals), p(self._status)
        a.s_near, a.near_reach = p(near), reach
        return a, totals

    def _dev_done(self, totals):
        """Queue the copy of a device-side build's totals to the host."""
        if self._host_slots is None:
            self._host_slots = _PinnedSlots(4, 3, torch.int64)
        self._make_room()
        host, ev = self._host_slots.take(self.cell_manager._copy_event())
        host.copy_(totals, non_blocking=True)   # {slots, nnz, status}
        self._queue(totals, host, ev)

    def _make_room(self):
        """At most two totals pending: a pinned buffer (a ring of four) is
        never reused before the host has read it."""
        if len(self._pending) >= 3:
            self._poll(block=True)

    def _queue(self, totals, host, ev):
        self._totals = totals
        self._pending.append((ev, host))
        self.nbuilds += 1

    def _built(self):
        self.built_version = self.cell_manager.version
        self.is_dirty = False
        self._export = None
        self._host = None

    def _build(self):
        cm = self.cell_manager
        ctx = cm._ctx()
        lib = ctx.lib
        dst, src = self.dest, self.source
        nd_total = dst.get_number_of_particles()
        nd_real = dst.num_real_particles
        bd, bs = cm.bins[dst.name], cm.bins[src.name]
        drec = cm.shadows(dst, ("x", "y", "z", "h"))
        srec = cm.shadows(src, ("x", "y", "z"))
        dev = dst.device
        self._buffers()
        p = _lib.ptr
        self._poll()
        if self._device_side(bd):
            a, totals = self._dev_args()
            if a.f32:
                rc = lib.sph_nnps_build_f32_dev(
                    ctx.handle, ctypes.byref(cm.grid), a.d_rec, a.d_x, a.d_y, a.d_z, a.d_perm,
                    a.d_cell_start, a.d_cells, a.nd_cells_dev, a.max_cells, a.nd_total,
                    a.nd_real, a.s_cell_start, a.s_rec, a.ns, a.radius_scale, a.same_array,
                    a.counts, a.slice_off, a.indices, a.cap, a.totals, a.status, a.s_near,
                    a.near_reach)
            else:
                rc = lib.sph_nnps_build_dev(
                    ctx.handle, ctypes.byref(cm.grid), a.d_rec, a.d_perm, a.d_cell_start,
                    a.d_cells, a.nd_cells_dev, a.max_cells, a.nd_total, a.nd_real,
                    a.s_cell_start, a.s_rec, a.ns, a.radius_scale, a.same_array, a.counts,
                    a.slice_off, a.indices, a.cap, a.totals, a.status, a.s_near, a.near_reach)
            ctx.check(rc)
            self._dev_done(totals)
            cm.fence()
            return
        slots = ctypes.c_int64(0)
        nnz = ctypes.c_int64(0)

        def call():
            if cm.precision == "single":
                return lib.sph_nnps_build_f32(
                    ctx.handle, ctypes.byref(cm.grid), p(drec), p(dst.get_carray("x")),
                    p(dst.get_carray("y")), p(dst.get_carray("z")), p(bd.perm),
                    p(bd.cell_start), p(bd.cells), bd.ncells_nonempty, nd_total, nd_real,
                    p(bs.cell_start), p(srec), bs.n, self.radius_scale, int(src is dst),
                    p(self.counts), p(self.slice_off), p(self.indices), self.indices.numel(),
                    ctypes.byref(slots), ctypes.byref(nnz))
            return lib.sph_nnps_build(ctx.handle, ctypes.byref(cm.grid), p(drec), p(bd.perm),
                                      p(bd.cell_start), p(bd.cells), bd.ncells_nonempty,
                                      nd_total, nd_real, p(bs.cell_start), p(srec), bs.n,
                                      self.radius_scale, int(src is dst), p(self.counts),
                                      p(self.slice_off), p(self.indices), self.indices.numel(),
                                      ctypes.byref(slots), ctypes.byref(nnz))

        rc = ctx.check(call(), allow=(_lib.SPH_ENEEDCAP,))
        if rc == _lib.SPH_ENEEDCAP:
            self.indices = torch.empty(int(slots.value * 1.15) + 4096, dtype=torch.int32,
                                       device=dev)
            ctx.check(call())
        self.nnz = int(nnz.value)
        self.slots = int(slots.value)
        self.nbuilds += 1

    def _stale(self):
        return self.built_version != self.cell_manager.version or self.is_dirty

    def get_sorted_lists(self):
        """SELL-64 lists in sorted space: (counts int32[nd], slice_off
        int64[nd/64+1], indices int32[slots])."""
        self.update()
        return self.counts, self.slice_off, self.indices

    def get_csr(self):
        """Lists in original index space: (offsets int64[nd_real+1],
        indices int32[nnz]); row i = neighbours of destination i."""
        self.update()
        if self._export is None:
            cm = self.cell_manager
            ctx = cm._ctx()
            nd_real = self.dest.num_real_particles
            dev = self.dest.device
            off = torch.empty(nd_real + 1, dtype=torch.int64, device=dev)
            idx = torch.empty(max(1, self.nnz), dtype=torch.int32, device=dev)
            p = _lib.ptr
            ctx.check(ctx.lib.sph_nnps_export(ctx.handle, nd_real, p(cm.bins[self.dest.name].rank),
                                              p(cm.bins[self.source.name].perm), p(self.counts),
                                              p(self.slice_off), p(self.indices), p(off), p(idx)))
            self._export = (off, idx[:self.nnz])
        return self._export

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


def update_together(locators):
    """Bring the lists of several locators up to date, as their update()
    calls in turn would.  Two stale device-side builds of fixed-h lists over
    one cell manager and one destination (a fluid's lists from itself and
    from a wall, say) go out as one paired build (sph_nnps_build_dev_pair):
    the smaller build runs in the tail of the larger one's launches."""
    pair = []
    for loc in locators:
        if type(loc) is not FixedDestNbrParticleLocator or loc in pair:
            continue
        cm = loc.cell_manager
        cm.update()
        if not loc._stale() or not loc._device_side(cm.bins[loc.dest.name]):
            continue
        if pair and (cm is not pair[0].cell_manager or loc.dest is not pair[0].dest):
            continue
        pair.append(loc)
        if len(pair) == 2:
            break
    if len(pair) == 2:
        cm = pair[0].cell_manager
        ctx = cm._ctx()
        with profiling.region("nnps_build"):
            for loc in pair:
                loc._poll()
            tot = torch.empty((2, 3), dtype=torch.int64, device=pair[0].dest.device)
            (a0, _), (a1, _) = (loc._dev_args(tot[k]) for k, loc in enumerate(pair))
            ctx.check(ctx.lib.sph_nnps_build_dev_pair(ctx.handle, ctypes.byref(cm.grid),
                                                      ctypes.byref(a0), ctypes.byref(a1)))
            # both builds' totals in one copy, from a ring of this pair's own
            slots = pair[0]._pair_slots.get(pair[1])
            if slots is None:
                slots = pair[0]._pair_slots[pair[1]] = _PinnedSlots(4, 6, torch.int64)
            for loc in pair:
                loc._make_room()
            host, ev = slots.take(cm._copy_event())
            host.copy_(tot.view(-1), non_blocking=True)
            for k, loc in enumerate(pair):
                loc._queue(tot[k], host[3 * k:3 * k + 3], ev)
                loc._built()
            cm.fence()
    for loc in locators:
        loc.update()


class VarHNbrParticleLocator(FixedDestNbrParticleLocator):
    """Symmetric lists for variable smoothing lengths (base/nnps.pyx:691-786):
    if b is a neighbour of a, a is a neighbour of b.  Row of destination a:
    its own query (radius h_a * radius_scale, stencil order, own index not
    moved last) followed by the real sources b outside that radius whose
    own radius h_b * radius_scale reaches a, in ascending source index --
    the rows the reference builds by appending the reversed locator's rows
    (``sph_nnps_build_varh``).  As in the reference, ``exclude_self`` of
    ``get_nearest_particles`` drops the last entry of the row."""

    def _build(self):
        cm = self.cell_manager
        ctx = cm._ctx()
        lib = ctx.lib
        dst, src = self.dest, self.source
        nd_total = dst.get_number_of_particles()
        bd, bs = cm.bins[dst.name], cm.bins[src.name]
        drec = cm.shadows(dst, ("x", "y", "z", "h"))
        srec = cm.shadows(src, ("x", "y", "z", "h"))
        dev = dst.device
        nslices = (nd_total + 63) // 64
        if self.counts is None or self.counts.numel() != max(nd_total, 1):
            self.counts = torch.empty(max(nd_total, 1), dtype=torch.int32, device=dev)
            self.slice_off = torch.empty(nslices + 1, dtype=torch.int64, device=dev)
        if self.indices is None:
            self.indices = torch.empty(max(1, 96 * nd_total), dtype=torch.int32, device=dev)
        slots = ctypes.c_int64(0)
        nnz = ctypes.c_int64(0)
        p = _lib.ptr
        # a bound of every h now (h may have been written since the binning)
        max_h = max(float(a.get_carray("h")[:a.get_number_of_particles()].max())
                    for a in (dst, src) if a.get_number_of_particles() > 0)

        # the destination's non-empty cells: counted on the device after the
        # first binning (the launch bound then bounds the count)
        if bd.ncells_dev is not None:
            ncells_dev, ncells = p(bd.ncells_dev), bd.cells_bound
        else:
            ncells_dev, ncells = None, bd.ncells_nonempty
        cells = (p(bd.perm), p(bd.cell_start), p(bd.cells), ncells_dev, ncells)

        def call():
            if cm.precision == "single":  # float tests, the binning's double coordinates
                return lib.sph_nnps_build_varh_f32(
                    ctx.handle, ctypes.byref(cm.grid), p(drec), p(dst.get_carray("x")),
                    p(dst.get_carray("y")), p(dst.get_carray("z")), *cells, nd_total,
                    dst.num_real_particles, p(bs.cell_start), p(srec), p(bs.perm), p(bs.rank),
                    bs.n, src.num_real_particles, self.radius_scale, max_h, p(self.counts),
                    p(self.slice_off), p(self.indices), self.indices.numel(),
                    ctypes.byref(slots), ctypes.byref(nnz))
            return lib.sph_nnps_build_varh(
                ctx.handle, ctypes.byref(cm.grid), p(drec), *cells, nd_total,
                dst.num_real_particles, p(bs.cell_start), p(srec), p(bs.perm), p(bs.rank), bs.n,
                src.num_real_particles, self.radius_scale, max_h, p(self.counts),
                p(self.slice_off),
                p(self.indices), self.indices.numel(), ctypes.byref(slots), ctypes.byref(nnz))

        rc = ctx.check(call(), allow=(_lib.SPH_ENEEDCAP,))
        if rc == _lib.SPH_ENEEDCAP:
            self.indices = torch.empty(int(slots.value * 1.15) + 4096, dtype=torch.int32,
                                       device=dev)
            ctx.check(call())
        self.nnz = int(nnz.value)
        self.slots = int(slots.value)
        self.nbuilds += 1


class NNPSManager:
    """One locator per (source, dest, radius_scale) (base/nnps.pyx:791-932);
    ``variable_h`` selects the symmetric variable-h locators (:895-910)."""

    def __init__(self, cell_manager=None, variable_h=False,
                 locator_type=NeighborLocatorType.SPHNeighborLocator):
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
            cls = VarHNbrParticleLocator if self.variable_h else FixedDestNbrParticleLocator
            loc = cls(source, dest, radius_scale, self.cell_manager)
            self.particle_locator_cache[key] = loc
        return loc

    def update(self):
        for loc in self.particle_locator_cache.values():
            loc.update_status()
        if self.cell_manager is not None:
            self.cell_manager.update_status()
        return 0

    py_update = update
