"""Many independent simulations advanced together on one GPU.

``BatchedSimulation`` packs the particle sets of several simulations (e.g.
the 2D elliptical drops of different radii, resolutions and velocity
scales) into one set of device arrays.  Each simulation keeps its own cell
grid -- the one ``CellManager`` would build for it alone -- so its
neighbour lists and every per-particle sum are exactly those of running it
through the single-simulation path (``drop_solver``).  One RK2 step
(``RK2Integrator.integrate``, solver/integrator.py:467-498) of SPHRho,
TaitEquation, MomentumEquation + artificial viscosity and PositionStepping
advances every simulation with one launch per kernel, and needs no host
synchronisation, so ``capture`` records steps in a HIP graph that
``replay`` launches in one call.

Capacities (cell keys, non-empty cells, list slots) bound the device work;
a step that would exceed them sets a status bit and is reported by
``check`` instead of writing out of bounds.
"""
import ctypes

import numpy as np
import torch

from . import _lib

_STATE = ("x", "y", "z", "u", "v", "w", "h", "m", "rho", "p", "cs",
          "x0", "y0", "z0", "u0", "v0", "w0", "ax", "ay", "az", "au", "av", "aw", "dt_fac")

STATUS_BITS = {1: "neighbour-list slots exceed list_cap", 2: "cell keys exceed key_cap",
               4: "non-finite coordinate or h", 8: "grid too large",
               16: "non-empty cells exceed max_cells",
               32: "h differs within a simulation (uniform_h or the reverse mode needs one h)",
               64: "the persistent evaluation's grid barrier timed out"}

# launch shapes of the fused step (sph_batch_desc); 0 = the library default
# (persistent: 1 = one launch per evaluation with a grid barrier, 0 = two;
# staged: 1 = neighbour records staged in LDS)
TUNE_FIELDS = ("op_threads", "op_blocks_per_cu", "nbr_tile", "cells_per_block", "persistent",
               "staged")


class BatchError(RuntimeError):
    pass


def _initial_cells(sims, radius_scale):
    """(cells of every grid, non-empty cells) of the simulations as given."""
    total = nonempty = 0
    for d in sims:
        h = np.asarray(d["h"], dtype=np.float64)
        pts = [np.asarray(d.get(k, np.zeros(len(d["x"]))), dtype=np.float64)
               for k in ("x", "y", "z")]
        ok = np.isfinite(h) & np.isfinite(pts[0]) & np.isfinite(pts[1]) & np.isfinite(pts[2])
        if not ok.any():
            continue
        cs = radius_scale * h[ok].max()
        if cs == 0.0:
            cs = 1.0
        c = [np.floor(a[ok] / cs).astype(np.int64) for a in pts]
        dims = [int(a.max() - a.min() + 1) for a in c]
        total += dims[0] * dims[1] * dims[2]
        keys = ((c[0] - c[0].min()) * dims[1] + (c[1] - c[1].min())) * dims[2] + (c[2] - c[2].min())
        nonempty += len(np.unique(keys))
    return total, nonempty


class BatchedSimulation:
    """Packed simulations sharing one operation list and physics constants.

    sims: list of dicts of numpy arrays x, y, z, u, v, w, h, m, rho, p, cs
    (missing z, w, p are zeros).  Physics: TaitEquation(co, ro, gamma),
    MomentumEquation(alpha, beta, eta), cubic spline, radius scale 2.
    precision "single": float records, lists and operators (the state and
    the stage arithmetic stay FP64).
    """

    def __init__(self, sims, dim=2, co=1400.0, ro=1.0, gamma=7.0, alpha=0.1, beta=0.0, eta=0.1,
                 radius_scale=2.0, device=None, key_cap=None, max_cells=None, list_cap=None,
                 maxch=None, cell_margin=4.0, list_slots_per_particle=None, dt_per_sim=None,
                 precision="double", tune=None):
        if not torch.cuda.is_available():
            raise RuntimeError("pysph_amd needs a HIP device; it has no CPU fallback")
        if dim not in (2, 3):
            raise ValueError("dim must be 2 or 3")
        self.dim = dim
        self.device = torch.device(device) if device is not None else torch.device("cuda", 0)
        self.nsims = len(sims)
        if self.nsims == 0:
            raise ValueError("no simulations")
        counts = [len(d["x"]) for d in sims]
        self.counts = counts
        self.sim_off_host = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        self.n = int(self.sim_off_host[-1])
        if self.n >= 2 ** 31:
            raise ValueError("too many particles for one batch")
        self.stream = torch.cuda.Stream(self.device)
        self.ctx = _lib.Context(self.device, stream=self.stream)
        lib = self.ctx.lib

        def packed(key, default=0.0):
            parts = [np.asarray(d[key], dtype=np.float64) if key in d
                     else np.full(len(d["x"]), default) for d in sims]
            return torch.from_numpy(np.concatenate(parts)).to(self.device)

        with torch.cuda.stream(self.stream):
            self.t = {}
            for k in ("x", "y", "z", "u", "v", "w", "h", "m", "rho", "p", "cs"):
                self.t[k] = packed(k)
            for k in _STATE[11:]:
                self.t[k] = torch.zeros(self.n, dtype=torch.float64, device=self.device)
            self.sim_off = torch.from_numpy(self.sim_off_host).to(self.device)
            self.sim_of = torch.from_numpy(
                np.repeat(np.arange(self.nsims, dtype=np.int32), counts)).to(self.device)
            self.dt_sim = None
            if dt_per_sim is not None:
                dts = np.asarray(dt_per_sim, dtype=np.float64)
                if dts.shape != (self.nsims,):
                    raise ValueError("dt_per_sim needs one time step per simulation")
                self.dt_sim = torch.from_numpy(dts.copy()).to(self.device)
            self.status = torch.zeros(1, dtype=torch.int32, device=self.device)
            self.pairs = torch.zeros(1, dtype=torch.int64, device=self.device)
        total, nonempty = _initial_cells(sims, radius_scale)
        self.key_cap = int(key_cap or min(2 ** 31 - 2, int(cell_margin * total) + 4096))
        self.max_cells = int(max_cells or min(self.key_cap, self.n, 2 * nonempty + 4096))
        per = list_slots_per_particle or (64 if dim == 2 else 192)
        self.list_cap = int(list_cap or per * self.n + 4096)
        self.maxch = int(maxch or (4 if dim == 2 else 16))
        self.params = dict(co=co, ro=ro, gamma=gamma, alpha=alpha, beta=beta, eta=eta,
                           radius_scale=radius_scale)
        d = _lib.BatchDesc()
        d.nsims, d.dim, d.n = self.nsims, dim, self.n
        d.sim_off, d.sim_of = self.sim_off.data_ptr(), self.sim_of.data_ptr()
        for k in _STATE:
            setattr(d, k, self.t[k].data_ptr())
        for k, v in self.params.items():
            setattr(d, k, float(v))
        d.key_cap, d.max_cells, d.list_cap, d.maxch = (self.key_cap, self.max_cells,
                                                       self.list_cap, self.maxch)
        if precision not in ("double", "single"):
            raise ValueError("precision must be 'double' or 'single'")
        self.precision = precision
        d.f32 = 1 if precision == "single" else 0
        # one h per simulation: the operators take h_ab = h_a without reading
        # the source's h (the grid kernel flags a simulation whose h differs)
        self.uniform_h = all(np.ptp(np.asarray(sd["h"], dtype=np.float64)) == 0.0
                             for sd in sims if len(sd["h"]))
        d.uniform_h = 1 if self.uniform_h else 0
        d.status, d.pairs = self.status.data_ptr(), self.pairs.data_ptr()
        d.dt_sim = self.dt_sim.data_ptr() if self.dt_sim is not None else None
        nbytes = ctypes.c_int64(0)
        d.work = None
        self.ctx.check(lib.sph_batch_workspace_size(self.ctx.handle, ctypes.byref(d),
                                                    ctypes.byref(nbytes)))
        with torch.cuda.stream(self.stream):
            self.work = torch.empty(int(nbytes.value), dtype=torch.uint8, device=self.device)
        d.work, d.work_bytes = self.work.data_ptr(), int(nbytes.value)
        self.desc = d
        self.set_tuning(tune or {})
        self.ctx.check(lib.sph_batch_init_workspace(self.ctx.handle, ctypes.byref(d)))
        self.graph = None
        self.graph_steps = 0
        self.steps_done = 0

    # ------------------------------------------------------------------
    def set_tuning(self, tune):
        """Launch shapes of the fused step, e.g. from ``autotune``: a dict
        over TUNE_FIELDS (missing keys keep their value).  Every choice gives
        the same results bit for bit; only the speed differs."""
        for k, v in dict(tune).items():
            if k not in TUNE_FIELDS:
                raise ValueError(f"unknown tuning field {k!r}")
            setattr(self.desc, k, int(v))

    def tuning(self):
        return {k: int(getattr(self.desc, k)) for k in TUNE_FIELDS}

    def snapshot(self):
        """Device copies of the state a step reads (for trial runs)."""
        with torch.cuda.stream(self.stream):
            return {k: v.clone() for k, v in self.t.items()}, self.pairs.clone()

    def restore(self, snap):
        state, pairs = snap
        with torch.cuda.stream(self.stream):
            for k, v in state.items():
                self.t[k].copy_(v)
            self.pairs.copy_(pairs)

    def step(self, dt=1.0):
        """One RK2 step of every simulation (queued on the batch's stream).
        With per-simulation time steps, dt multiplies them (1.0: one step)."""
        self.ctx.check(self.ctx.lib.sph_batch_rk2_step(self.ctx.handle, ctypes.byref(self.desc),
                                                       float(dt)))
        self.steps_done += 1

    def capture(self, dt, nsteps=1):
        """Record `nsteps` RK2 steps in a HIP graph.  Runs one eager step
        first when none ran yet (it sizes the context scratch, which the
        graph then pins)."""
        if self.steps_done == 0:
            self.step(dt)
            self.check()
        self.release_graph()
        g = ctypes.c_void_p()
        self.ctx.check(self.ctx.lib.sph_batch_graph_create(
            self.ctx.handle, ctypes.byref(self.desc), float(dt), int(nsteps), ctypes.byref(g)))
        self.graph = g
        self.graph_steps = int(nsteps)

    def replay(self):
        if self.graph is None:
            raise BatchError("no captured graph")
        self.ctx.check(self.ctx.lib.sph_batch_graph_launch(self.ctx.handle, self.graph))
        self.steps_done += self.graph_steps

    def release_graph(self):
        if self.graph is not None:
            self.ctx.check(self.ctx.lib.sph_batch_graph_destroy(self.ctx.handle, self.graph))
            self.graph = None

    def synchronize(self):
        self.stream.synchronize()

    def check(self):
        """Wait for the queued steps and raise if any exceeded a capacity."""
        self.synchronize()
        st = int(self.status.item())
        if st:
            why = [msg for bit, msg in STATUS_BITS.items() if st & bit]
            raise BatchError(f"batched step void (status {st}): " + "; ".join(why))

    def pair_interactions(self):
        """Neighbour-operator pair evaluations so far (each list build serves
        SPHRho and MomentumEquation)."""
        self.synchronize()
        return 2 * int(self.pairs.item())

    # ------------------------------------------------------------------
    def get(self, key, sim=None):
        """Host copy of a property, packed or of one simulation."""
        self.synchronize()
        a = self.t[key].cpu().numpy()
        if sim is None:
            return a
        return a[self.sim_off_host[sim]:self.sim_off_host[sim + 1]]

    def __del__(self):
        try:
            self.release_graph()
        except Exception:
            pass


def drop_time_step(dx, dt_ref=1e-5, dx_ref=0.025):
    """The elliptical drop's dt = 1e-5 at dx = 0.025, scaled with dx so that
    every resolution keeps that run's CFL number c0*dt/h (at dx = 0.01 the
    unscaled 1e-5 gives c0*dt/h = 1.08 and the drop blows up)."""
    return dt_ref * min(1.0, float(dx) / dx_ref)


def batched_drops(nsims=256, seed=20261202, radius_range=(0.4, 2.0), dxs=(0.01, 0.02, 0.04),
                  vel_range=(50.0, 150.0), rho0=1.0, c0=1400.0, hdx=1.3):
    """The batched elliptical drops: radius ~ U(radius_range), dx drawn from
    dxs, velocity scale ~ U(vel_range) (u = -s x, v = s y), seeded PCG64.
    Returns (list of drop dicts, per-drop parameters incl. the time step
    drop_time_step(dx))."""
    from .solver.generators import elliptical_drop
    rng = np.random.default_rng(seed)
    radii = rng.uniform(radius_range[0], radius_range[1], nsims)
    dx = rng.choice(np.asarray(dxs, dtype=np.float64), nsims)
    vel = rng.uniform(vel_range[0], vel_range[1], nsims)
    sims = [elliptical_drop(dx=float(dx[k]), radius=float(radii[k]), rho0=rho0, c0=c0,
                            vel=float(vel[k]), hdx=hdx) for k in range(nsims)]
    return sims, dict(radius=radii, dx=dx, vel=vel,
                      dt=np.array([drop_time_step(v) for v in dx]))

