"""Benchmark of the PySPH particle hot path on one MI355X.

Default workload: the 3D WCSPH dam break -- 1,000,000 fluid particles
(1.0 x 0.5 x 2.0 column at dx = 0.01) in a 4.0 x 0.5 x 2.5 open tank of
three wall layers (752,208 particles), h = 1.3 dx, rho0 = 1000,
c0 = 10 sqrt(2 g 2), gamma = 7, alpha = 0.1, beta = 0, gravity -9.81 z,
PredictorCorrector, dt = 0.25 h / c0, FP64.  One step = one full
PredictorCorrector step through the public solver API: two evaluations,
each rebuilding the fluid<-fluid and fluid<-wall neighbour lists, then
density, Tait, momentum + AV, gravity, position stepping and the stage
arithmetic.

metric: pair-interactions/s = sum of neighbour-list lengths over every
neighbour-operator evaluation (density and momentum, both sources, both
stages) per second of wall time.

``--workload lattice`` runs the NNPS + SummationDensity micro-workload
(100^3 jittered lattice, h = 1.2 dx; ``--h-jitter 0.2 --variable-h`` gives every
particle its own h and the symmetric variable-h lists) instead;
``--workload dambreak-split``
the same dam break with the operator split of examples/dam-break/
dam_break3d.py (continuity, pressure gradient and Monaghan viscosity as
separate functions, XSPH).

Also reported: the roofline fraction of the dominant kernel (momentum)
measured with HIP events on its stream, and the FP64 CPU oracle timed on
the host cores on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def _args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=("dambreak", "dambreak-split", "lattice", "batch",
                                           "adjoint"),
                    default="dambreak")
    ap.add_argument("--fd-particles", type=int, default=64,
                    help="particles whose u0 and v0 gradients --workload adjoint checks by "
                         "central differences")
    ap.add_argument("--lbfgs-iters", type=int, default=20)
    ap.add_argument("--precision", choices=("double", "single"), default="double",
                    help="arithmetic of the lattice and batch workloads (single: float records, "
                         "lists and operators)")
    ap.add_argument("--nsims", type=int, default=256, help="drops of --workload batch")
    ap.add_argument("--no-autotune", action="store_true",
                    help="--workload batch: keep the default launch shapes")
    ap.add_argument("--launch-shape", default="",
                    help="--workload batch: fixed launch shapes instead of the autotuner, "
                         "e.g. cells_per_block=8,nbr_tile=4")
    ap.add_argument("--graph-steps", type=int, default=10,
                    help="RK2 steps per captured graph (--workload batch)")
    ap.add_argument("--seq-sims", type=int, default=0,
                    help="drops timed one by one on the plain path for the sequential "
                         "comparison of --workload batch (0: all)")
    ap.add_argument("--dx", type=float, default=0.01)
    ap.add_argument("--n", type=int, default=100, help="lattice edge for --workload lattice")
    ap.add_argument("--h-jitter", type=float, default=0.0,
                    help="--workload lattice: h scaled by U(1-f, 1+f) per particle (seeded)")
    ap.add_argument("--variable-h", action="store_true",
                    help="--workload lattice: Particles(variable_h=True), the symmetric lists")
    ap.add_argument("--cell-order", choices=("morton", "linear"), default="morton")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="launch one k_copy of this many MiB before the timed steps (a kernel "
                         "of known traffic that scales hardware byte counters)")
    return ap.parse_args()


def copy_calibration(mib):
    """One sph_copy of `mib` MiB (reads and writes mib MiB each)."""
    import torch
    from pysph_amd import _lib
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    ctx = _lib.context(a.device)
    torch.cuda.synchronize()
    ctx.check(ctx.lib.sph_copy(ctx.handle, n, _lib.ptr(a), _lib.ptr(b)))
    torch.cuda.synchronize()
    del a, b


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_momentum_dambreak_1M_traffic.json")


def measured_traffic(path=None):
    """HBM bytes per launch of the roofline kernel from the committed
    rocprofv3 byte counters (FETCH_SIZE, WRITE_SIZE, scaled by a 1 GiB copy),
    or None."""
    try:
        with open(path or TRAFFIC_FILE) as f:
            return float(json.load(f)["hbm_bytes_per_launch"])
    except (OSError, ValueError, KeyError):
        return None


def momentum_bytes(nd, nnz, ns_other, ns):
    """Algorithmic HBM bytes of MomentumEquation over one source (k_prho2
    over the ns source records if needed, then k_momentum): P/rho^2 of each
    source (p and rho read, P/rho^2 written: 24 B), and per destination the
    int64 row offset, 10 destination properties + dt_fac read, ax ay az
    read-modify-write and dt_fac written (8+11*8+3*16+8 = 152 B), 4 B per
    list entry, and -- when the source is another array -- its 11
    properties once (88 B per source particle).  A launch over several
    sources reads and writes each destination once: pass the entries of all
    lists and the particles of the other arrays."""
    return 24 * ns + nd * 152 + 4 * nnz + 88 * ns_other


def avisc_bytes(nd, nnz, ns_other):
    """Algorithmic HBM bytes of one MonaghanArtificialViscosity launch over
    all its sources: per destination the list count, 9 properties + dt_fac
    read, ax ay az read-modify-write and dt_fac written (4+10*8+3*16+8 =
    140 B), 4 B per list entry, and the other arrays' 10 properties once
    (80 B per source particle)."""
    return nd * 140 + 4 * nnz + 80 * ns_other


def run_dambreak_split(args):
    """The operator split of dam_break3d.py (continuity, Tait, pressure
    gradient, Monaghan viscosity, gravity, XSPH) on the same 1M dam break."""
    import torch
    from pysph_amd import profiling
    from pysph_amd.solver.configs import dambreak_split_solver
    from pysph_amd.solver.generators import dam_break_3d

    data = dam_break_3d(dx=args.dx)
    fl, wl, prm = data
    s, particles, _ = dambreak_split_solver(data=data, cell_order=args.cell_order)
    fluid, wall = particles.arrays
    integ = s.integrator
    dt = prm["dt"]
    avisc_calc = [c for c in integ.calcs if c.id.startswith("avisc")][0]
    for _ in range(args.warmup):
        integ.integrate(dt)
    if args.copy_calibration_mb:
        copy_calibration(args.copy_calibration_mb)
    torch.cuda.synchronize()
    profiling.reset()
    profiling.enable(True)
    for _ in range(min(5, args.steps)):
        integ.integrate(dt)
    profiling.enable(False)
    breakdown = profiling.summary()
    profiling.reset()
    region = avisc_calc.funcs[0].id
    profiling.enable(True, regions=(region,))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        integ.integrate(dt)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    profiling.enable(False)
    pairs = profiling.pairs()
    prof = profiling.summary()
    elapsed = t1 - t0
    nd = fluid.get_number_of_particles()
    nw = wall.get_number_of_particles()
    nnz_ff = avisc_calc.funcs[0].nbr_locator.nnz
    nnz_fw = avisc_calc.funcs[1].nbr_locator.nnz
    av = prof.get(region, {"mean_ms": float("nan"), "count": 0})
    bytes_per_launch = avisc_bytes(nd, nnz_ff + nnz_fw, nw)
    achieved = bytes_per_launch / (av["mean_ms"] * 1e-3) / 1e9
    traffic_file = os.path.join(ROOT, "profiles", "k_pair_avisc_dambreak_1M_traffic.json")
    result = {
        "metric": "pair-interactions/s",
        "value": pairs / elapsed,
        "unit": "pair-interactions/s",
        "ms_per_step": 1e3 * elapsed / args.steps,
        "dtype": "f64",
        "config": {
            "workload": "3D dam break, operator split of dam_break3d.py (continuity, Tait, "
                        "SPHPressureGradient, MonaghanArtificialViscosity, gravity, XSPH), "
                        "PredictorCorrector step, FP64",
            "fluid_particles": nd, "wall_particles": nw, "dx": args.dx, "dt": dt,
            "steps_per_s": args.steps / elapsed, "pairs_per_step": pairs / args.steps,
            "nnz_fluid_fluid": nnz_ff, "nnz_fluid_wall": nnz_fw,
            "kernel_ms": {k: round(v["mean_ms"], 4) for k, v in breakdown.items()},
            "kernel_launches_per_step": {k: v["count"] / min(5, args.steps)
                                         for k, v in breakdown.items()},
        },
        "roofline": {
            "kernel": "k_pair (MonaghanArtificialViscosity, fluid + wall sources)",
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(traffic_file) if args.dx == 0.01 else None,
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "mean_launch_ms": av["mean_ms"],
        },
    }
    return result, data


def run_dambreak(args):
    import torch
    from pysph_amd import profiling
    from pysph_amd.solver.configs import dambreak_solver
    from pysph_amd.solver.generators import dam_break_3d

    data = dam_break_3d(dx=args.dx)
    fl, wl, prm = data
    s, particles, _ = dambreak_solver(data=data, cell_order=args.cell_order)
    fluid, wall = particles.arrays
    integ = s.integrator
    dt = prm["dt"]
    dens_calc = integ.calcs[0]
    mom_calc = [c for c in integ.calcs if c.id.startswith("mom")][0]

    for _ in range(args.warmup):
        integ.integrate(dt)
    if args.copy_calibration_mb:
        copy_calibration(args.copy_calibration_mb)
    # per-region breakdown (config.kernel_ms): a short instrumented run
    # outside the timed loop, events around every library call
    torch.cuda.synchronize()
    profiling.reset()
    profiling.enable(True)
    for _ in range(min(5, args.steps)):
        integ.integrate(dt)
    profiling.enable(False)
    breakdown = profiling.summary()
    # timed loop: events only around the momentum launches (the roofline
    # kernel), pair interactions counted
    profiling.reset()
    profiling.enable(True, regions=("momentum_av",))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        integ.integrate(dt)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    profiling.enable(False)
    pairs = profiling.pairs()
    prof = profiling.summary()
    elapsed = t1 - t0
    nd = fluid.get_number_of_particles()
    nw = wall.get_number_of_particles()
    nnz_ff = mom_calc.funcs[0].nbr_locator.nnz
    nnz_fw = mom_calc.funcs[1].nbr_locator.nnz
    # dominant kernel: momentum, one launch per evaluation over both sources
    # (fluid and wall lists; a PC step evaluates twice).  P/rho^2 of both
    # sources is current when momentum runs (Tait writes the fluid's, the
    # walls' is computed once), so the calls run k_momentum only.
    mom = prof.get("momentum_av", {"mean_ms": float("nan"), "count": 0, "total_ms": float("nan")})
    # the fused launch reads and writes each destination once
    b_eval = momentum_bytes(nd, nnz_ff + nnz_fw, nw, 0)
    evals = 2 * args.steps
    bytes_per_launch = b_eval * evals / max(mom["count"], 1)
    achieved = bytes_per_launch / (mom["mean_ms"] * 1e-3) / 1e9
    result = {
        "metric": "pair-interactions/s",
        "value": pairs / elapsed,
        "unit": "pair-interactions/s",
        "ms_per_step": 1e3 * elapsed / args.steps,
        "dtype": "f64",
        "config": {
            "workload": "3D WCSPH dam break, PredictorCorrector step, FP64",
            "fluid_particles": nd, "wall_particles": nw, "dx": args.dx,
            "h": prm["h"], "c0": prm["c0"], "dt": dt, "cell_order": args.cell_order,
            "steps_per_s": args.steps / elapsed,
            "pairs_per_step": pairs / args.steps,
            "nnz_fluid_fluid": nnz_ff, "nnz_fluid_wall": nnz_fw,
            "mean_neighbours": (nnz_ff + nnz_fw) / nd,
            "kernel_ms": {k: round(v["mean_ms"], 4) for k, v in breakdown.items()},
            "kernel_launches_per_step": {k: v["count"] / min(5, args.steps)
                                         for k, v in breakdown.items()},
            "kernel_ms_note": "per library call, from an instrumented run of "
                              f"{min(5, args.steps)} steps outside the timed loop",
        },
        "roofline": {
            "kernel": "k_momentum",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic() if args.dx == 0.01 else None,
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "mean_launch_ms": mom["mean_ms"],
        },
    }
    return result, data


def run_lattice(args):
    import torch
    from pysph_amd import profiling
    from pysph_amd.base import api as base
    from pysph_amd.sph import api as sph
    from pysph_amd.solver.generators import jittered_lattice_3d

    d = jittered_lattice_3d(n=args.n)
    if args.h_jitter > 0:
        rng = np.random.default_rng(1)
        d["h"] = d["h"] * rng.uniform(1 - args.h_jitter, 1 + args.h_jitter, len(d["x"]))
    pa = base.get_particle_array(name="fluid", cl_precision=args.precision, **d)
    particles = base.Particles([pa], cell_order=args.cell_order, variable_h=args.variable_h)
    kernel = base.CubicSplineKernel(dim=3)
    func = sph.SPHRho.withargs().get_func(pa, pa)
    calc = sph.SPHCalc(particles, [pa], pa, kernel, [func], ['rho'])

    def step():
        pa.set_dirty(True)  # force the rebuild: one step = binning + lists + density
        particles.update()
        calc.sph('rho')

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    profiling.reset()
    profiling.enable(True)
    for _ in range(min(5, args.steps)):
        step()
    profiling.enable(False)
    breakdown = profiling.summary()
    profiling.reset()
    profiling.enable(True, regions=("summation_density",))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    profiling.enable(False)
    prof = profiling.summary()
    n = pa.get_number_of_particles()
    nnz = func.nbr_locator.nnz
    elapsed = t1 - t0
    dens = prof["summation_density"]
    # per destination: row offset, x y z h m read once, rho written; 4 B per entry
    b = n * (8 + 5 * 8 + 8) + 4 * nnz
    achieved = b / (dens["mean_ms"] * 1e-3) / 1e9
    return {
        "metric": "pair-interactions/s",
        "value": nnz * args.steps / elapsed,
        "unit": "pair-interactions/s",
        "ms_per_step": 1e3 * elapsed / args.steps,
        "dtype": "f32" if args.precision == "single" else "f64",
        "config": {"workload": f"3D NNPS + SummationDensity, {args.n}^3 jittered lattice, "
                               f"{'FP32' if args.precision == 'single' else 'FP64'}"
                               + (f", h x U(1-{args.h_jitter}, 1+{args.h_jitter})"
                                  if args.h_jitter > 0 else "")
                               + (", variable-h lists" if args.variable_h else ""),
                   "particles": n, "nnz": nnz, "steps_per_s": args.steps / elapsed,
                   "kernel_ms": {k: round(v["mean_ms"], 4) for k, v in breakdown.items()},
                   "kernel_ms_note": "per library call, from an instrumented run outside the "
                                     "timed loop"},
        "roofline": {"kernel": "k_summation_density", "bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "algorithmic_bytes_per_launch": b,
                     "mean_launch_ms": dens["mean_ms"]},
    }, d


def run_batch(args):
    """256 independent 2D elliptical drops (radius U[0.4, 2], dx from
    {0.01, 0.02, 0.04}, velocity scale U[50, 150], seeded), RK2 with each
    drop's time step 1e-5*min(1, dx/0.025) (the CFL number of the dx = 0.025
    drop), advanced together by captured HIP graphs of --graph-steps steps;
    compared with running the same drops one after another through the plain
    path."""
    import torch
    from pysph_amd.batch import BatchedSimulation, batched_drops
    from pysph_amd.solver.configs import drop_solver
    t_start = time.perf_counter()

    def progress(msg):
        print(f"[batch {time.perf_counter() - t_start:7.2f}s] {msg}", file=sys.stderr, flush=True)

    sims, prm = batched_drops(nsims=args.nsims)
    progress(f"{len(sims)} drops generated, {sum(len(d['x']) for d in sims)} particles")
    b = BatchedSimulation(sims, dt_per_sim=prm["dt"], precision=args.precision)
    dt = 1.0  # multiplies every drop's own time step
    progress(f"packed: key_cap {b.key_cap} max_cells {b.max_cells} list_cap {b.list_cap} "
             f"workspace {b.work.numel() / 2**20:.0f} MiB")
    b.step(dt)
    b.check()
    progress("eager step done")
    tune = None
    if args.launch_shape:
        b.set_tuning({k: int(v) for k, v in
                      (kv.split("=") for kv in args.launch_shape.split(","))})
        progress(f"launch shape {b.tuning()}")
    elif not args.no_autotune:
        from pysph_amd.autotune import autotune
        tune = autotune(b, dt, steps=3, use_cache=False, log=progress)
        progress(f"autotuned: {tune['tuning']} {tune['ms_per_step']:.3f} ms/step "
                 f"(default shape {tune['trials'][0][1]:.3f} ms/step)")
    gs = max(1, args.graph_steps)
    b.capture(dt, nsteps=gs)
    progress(f"graph of {gs} steps captured")
    nrep_w = max(1, args.warmup // gs)
    for _ in range(nrep_w):
        b.replay()
    b.check()
    progress("warm-up replays done")
    pairs0 = b.pair_interactions()
    nrep = max(1, args.steps // gs)
    t0 = time.perf_counter()
    for _ in range(nrep):
        b.replay()
    b.synchronize()
    t1 = time.perf_counter()
    b.check()
    pairs = b.pair_interactions() - pairs0
    steps = nrep * gs
    elapsed = t1 - t0
    batch_sps = steps / elapsed
    progress(f"timed {steps} steps: {1e3 * elapsed / steps:.3f} ms/step")
    # sequential plain-path runs of the same drops (steps/s of the whole set)
    order = np.argsort([len(d["x"]) for d in sims])
    pick = order if not args.seq_sims else order[np.linspace(0, len(order) - 1, args.seq_sims).astype(int)]
    seq_time = 0.0
    seq_steps = 2
    for k in pick:
        solver, particles, _ = drop_solver(data=dict(sims[k]), dx=float(prm["dx"][k]))
        dtk = float(prm["dt"][k])
        solver.integrator.integrate(dtk)          # first build, allocations
        torch.cuda.synchronize()
        ta = time.perf_counter()
        for _ in range(seq_steps):
            solver.integrator.integrate(dtk)
        torch.cuda.synchronize()
        seq_time += (time.perf_counter() - ta) / seq_steps
        progress(f"sequential drop {k}: {len(sims[k]['x'])} particles")
    seq_time *= len(sims) / len(pick)
    seq_sps = 1.0 / seq_time
    return {
        "metric": "steps/s",
        "value": batch_sps,
        "unit": "steps/s (all drops advanced one RK2 step)",
        "ms_per_step": 1e3 * elapsed / steps,
        "dtype": "f32" if args.precision == "single" else "f64",
        "config": {"workload": f"{len(sims)} independent 2D elliptical drops, RK2, "
                               f"{'FP32 operators' if args.precision == 'single' else 'FP64'}, "
                               f"dt 1e-5*min(1, dx/0.025) per drop, captured HIP graphs of "
                               f"{gs} steps",
                   "particles": b.n, "drops": len(sims),
                   "smallest_drop": int(min(b.counts)), "largest_drop": int(max(b.counts)),
                   "pair_interactions_per_s": pairs / elapsed,
                   "sequential_plain_gpu_steps_per_s": seq_sps,
                   "sequential_drops_timed": int(len(pick)),
                   "speedup_vs_sequential": batch_sps / seq_sps,
                   "fused_kernels": "density+EOS, momentum+position stepping+stage",
                   "launch_shape": b.tuning(),
                   "autotune": None if tune is None else {
                       "trials": len(tune["trials"]),
                       "default_shape_ms_per_step": tune["trials"][0][1],
                       "tuned_ms_per_step_eager": tune["ms_per_step"]}},
        "roofline": None,
    }, None


def run_adjoint(args):
    """Gradient of the final-position MSE of the 2D drop (dx = 0.02, 7,860
    particles, RK2, dt = 1e-5, 200 steps, checkpoint every 20) against a
    target run from u = -120 x, v = 120 y, with respect to the per-particle
    initial velocities (around u = -100 x, v = 100 y plus seeded N(0, 5)) and
    alpha; checked against central differences (eps = 1e-6) on 64 random
    particles, then fitted with 20 L-BFGS iterations."""
    import torch
    from pysph_amd.adjoint import Rollout, rollout
    from pysph_amd.batch import BatchedSimulation
    from pysph_amd.solver.generators import elliptical_drop
    t_start = time.perf_counter()

    def progress(msg):
        print(f"[adjoint {time.perf_counter() - t_start:7.2f}s] {msg}", file=sys.stderr,
              flush=True)

    dx, steps, ckpt, dt, eps = 0.02, 200, 20, 1e-5, 1e-6
    d = elliptical_drop(dx=dx, radius=1.0, vel=100.0)
    n = len(d["x"])
    h = float(d["h"][0])
    pos0 = np.stack([d["x"], d["y"]], 1)
    roll = Rollout(pos0, d["m"], h, nsteps=steps, dt=dt, checkpoint=ckpt)
    dev = roll.sim.device
    tx, ty = roll.forward(torch.tensor(-120.0 * d["x"], device=dev),
                          torch.tensor(120.0 * d["y"], device=dev), 0.1)
    tx, ty = tx.clone(), ty.clone()
    rng = np.random.default_rng(20261202)
    u_init = -100.0 * d["x"] + rng.normal(0.0, 5.0, n)
    v_init = 100.0 * d["y"] + rng.normal(0.0, 5.0, n)
    progress(f"{n} particles, target run done")

    def loss_fn(x, y):
        return ((x - tx) ** 2 + (y - ty) ** 2).mean()

    u0 = torch.tensor(u_init, device=dev, requires_grad=True)
    v0 = torch.tensor(v_init, device=dev, requires_grad=True)
    alpha = torch.tensor(0.1, dtype=torch.float64, device=dev, requires_grad=True)
    # warm-up, then timed forward and backward
    x, y = rollout(roll, alpha, u0, v0)
    torch.autograd.grad(loss_fn(x, y), (u0, v0, alpha))
    reps = 3
    t_fwd = t_bwd = 0.0
    for _ in range(reps):
        torch.cuda.synchronize()
        ta = time.perf_counter()
        x, y = rollout(roll, alpha, u0, v0)
        loss = loss_fn(x, y)
        torch.cuda.synchronize()
        tb = time.perf_counter()
        gu, gv, ga = torch.autograd.grad(loss, (u0, v0, alpha))
        torch.cuda.synchronize()
        tc = time.perf_counter()
        t_fwd += (tb - ta) / reps
        t_bwd += (tc - tb) / reps
    progress(f"forward {1e3 * t_fwd:.1f} ms, backward {1e3 * t_bwd:.1f} ms")
    # central differences: every perturbed rollout is one simulation of a batch
    picks = rng.choice(n, args.fd_particles, replace=False)
    which = [(int(i), comp, sgn) for i in picks for comp in ("u", "v") for sgn in (1.0, -1.0)]

    def batch_losses(sim_list, alpha_val):
        b = BatchedSimulation(sim_list, alpha=alpha_val)
        for _ in range(steps):
            b.step(dt)
        b.check()
        xs, ys = b.get("x"), b.get("y")
        txn, tyn = tx.cpu().numpy(), ty.cpu().numpy()
        out = []
        for k in range(len(sim_list)):
            sl = slice(b.sim_off_host[k], b.sim_off_host[k + 1])
            out.append(float((((xs[sl] - txn) ** 2 + (ys[sl] - tyn) ** 2)).mean()))
        return out

    gmap = {"u": gu.cpu().numpy(), "v": gv.cpu().numpy()}

    def fd_errors(e):
        sims_e = []
        for i, comp, sgn in which:
            dd = dict(d)
            dd["u"], dd["v"] = u_init.copy(), v_init.copy()
            dd[comp][i] += sgn * e
            sims_e.append(dd)
        L = batch_losses(sims_e, 0.1)
        g, fd = [], []
        for k in range(0, len(sims_e), 2):
            i, comp, _ = which[k]
            g.append(gmap[comp][i])
            fd.append((L[k] - L[k + 1]) / (2 * e))
        g, fd = np.array(g), np.array(fd)
        rel = np.abs(g - fd) / np.maximum(np.abs(fd), 1e-300)
        return rel, np.abs(g - fd).max() / np.abs(fd).max(), \
            float(np.linalg.norm(g - fd) / np.linalg.norm(fd))

    rel6, norm6, l2_6 = fd_errors(eps)
    rel8, norm8, l2_8 = fd_errors(1e-8)
    d0 = dict(d)
    d0["u"], d0["v"] = u_init.copy(), v_init.copy()
    alpha_sweep = {}
    for e in (1e-5, 1e-6, 1e-7, 1e-8):
        la = batch_losses([d0], 0.1 + e)[0]
        lb = batch_losses([d0], 0.1 - e)[0]
        fda = (la - lb) / (2 * e)
        alpha_sweep[f"{e:g}"] = abs(float(ga) - fda) / abs(fda)
    progress(f"FD eps 1e-6: {len(rel6)} velocity components, |g - FD| / |FD| {l2_6:.3g}, "
             f"max err / max |FD| {norm6:.3g}, componentwise max {rel6.max():.3g} median "
             f"{np.median(rel6):.3g}; eps 1e-8: |g - FD| / |FD| {l2_8:.3g}; alpha rel err by "
             f"eps {alpha_sweep}")
    # fits, loss relative to its initial value l0.  (a) per-particle initial
    # velocities and alpha, as the workload names them.  The final positions
    # move by about T = steps*dt per unit of initial velocity, so the
    # Gauss-Newton curvature of the mean squared error in one velocity
    # component is 2 T^2 / n: velocities in units of s_u = sqrt(l0 n / 2) / T
    # give the velocity block unit curvature; alpha's unit is 1e-3.  (b) the
    # initial velocity field as a correction in the quadratic polynomials
    # {1, x, y, x^2, xy, y^2} per component (the target field -120 x, 120 y
    # is in that space; the seeded noise is not), and alpha.
    l0 = float(loss.item())
    T = steps * dt
    u_base = torch.tensor(u_init, device=dev)
    v_base = torch.tensor(v_init, device=dev)
    xs = torch.tensor(d["x"], device=dev)
    ys = torch.tensor(d["y"], device=dev)
    basis = torch.stack([torch.ones_like(xs), xs, ys, xs * xs, xs * ys, ys * ys], 1)

    def lbfgs(params, inputs, max_eval):
        opt = torch.optim.LBFGS(params, lr=1.0, max_iter=args.lbfgs_iters, max_eval=max_eval,
                                history_size=20, line_search_fn="strong_wolfe",
                                tolerance_grad=1e-14, tolerance_change=1e-18)
        losses = []

        def closure():
            opt.zero_grad()
            a_val, uu, vv = inputs()
            xx, yy = rollout(roll, a_val, uu, vv)
            lo = loss_fn(xx, yy) / l0
            lo.backward()
            losses.append(float(lo.item()) * l0)
            return lo

        t0 = time.perf_counter()
        opt.step(closure)
        return losses, time.perf_counter() - t0

    s_u = (l0 * n / 2.0) ** 0.5 / T
    s_a = 1e-3
    du = torch.zeros(n, dtype=torch.float64, device=dev, requires_grad=True)
    dv = torch.zeros(n, dtype=torch.float64, device=dev, requires_grad=True)
    da = torch.zeros((), dtype=torch.float64, device=dev, requires_grad=True)
    losses, t_fit = lbfgs([du, dv, da], lambda: (0.1 + s_a * da, u_base + s_u * du,
                                                  v_base + s_u * dv), 3 * args.lbfgs_iters)
    progress(f"L-BFGS per-particle velocities + alpha: {len(losses)} evaluations, loss "
             f"{losses[0]:.4g} -> {losses[-1]:.4g}")
    s_c = 10.0
    cu = torch.zeros(6, dtype=torch.float64, device=dev, requires_grad=True)
    cv = torch.zeros(6, dtype=torch.float64, device=dev, requires_grad=True)
    db = torch.zeros((), dtype=torch.float64, device=dev, requires_grad=True)
    losses_f, t_fit_f = lbfgs([cu, cv, db], lambda: (0.1 + s_a * db, u_base + basis @ (s_c * cu),
                                                     v_base + basis @ (s_c * cv)),
                              3 * args.lbfgs_iters)
    progress(f"L-BFGS quadratic velocity field + alpha: {len(losses_f)} evaluations, loss "
             f"{losses_f[0]:.4g} -> {losses_f[-1]:.4g}; x coefficient of u "
             f"{-100.0 + s_c * float(cu[1]):.3f}, y coefficient of v {100.0 + s_c * float(cv[2]):.3f}")
    return {
        "metric": "adjoint/forward cost ratio",
        "value": t_bwd / t_fwd,
        "unit": "backward time / forward time",
        "ms_per_step": 1e3 * t_fwd / steps,
        "higher_is_better": False,
        "dtype": "f64",
        "config": {"workload": f"2D elliptical drop dx {dx} ({n} particles), RK2 dt {dt}, "
                               f"{steps} steps, checkpoint every {ckpt}, FP64",
                   "forward_ms": 1e3 * t_fwd, "backward_ms": 1e3 * t_bwd,
                   "fd_eps": eps, "fd_components": int(len(rel6)),
                   "fd_norm_rel_err": l2_6, "fd_eps_1e-8_norm_rel_err": l2_8,
                   "fd_max_rel_err": float(rel6.max()),
                   "fd_median_rel_err": float(np.median(rel6)),
                   "fd_max_err_over_max_fd": float(norm6),
                   "fd_eps_1e-8_max_rel_err": float(rel8.max()),
                   "fd_eps_1e-8_median_rel_err": float(np.median(rel8)),
                   "fd_alpha_rel_err_by_eps": alpha_sweep,
                   "grad_alpha": float(ga),
                   "lbfgs_iterations": args.lbfgs_iters,
                   "fit_per_particle": {
                       "evaluations": len(losses), "loss_initial": losses[0],
                       "loss_final": losses[-1], "loss_reduction": losses[-1] / losses[0],
                       "alpha_fitted": 0.1 + s_a * float(da.item()), "fit_s": t_fit,
                       "max_velocity_change": float(s_u * max(du.abs().max().item(),
                                                              dv.abs().max().item()))},
                   "fit_quadratic_field": {
                       "evaluations": len(losses_f), "loss_initial": losses_f[0],
                       "loss_final": losses_f[-1],
                       "loss_reduction": losses_f[-1] / losses_f[0],
                       "alpha_fitted": 0.1 + s_a * float(db.item()), "fit_s": t_fit_f,
                       "u_x_coefficient": -100.0 + s_c * float(cu[1]),
                       "v_y_coefficient": 100.0 + s_c * float(cv[2])}},
        "roofline": None,
    }, None


def cpu_baseline_dambreak(args, data):
    """The FP64 oracle on the host cores: one PredictorCorrector step of the
    same 1M-fluid dam break (about 10-30 s of CPU work)."""
    from oracle import cpu_ref
    fl, wl, prm = data
    threads = args.cpu_threads or len(os.sched_getaffinity(0))
    ref = cpu_ref.RefSimulation([cpu_ref.RefArray("fluid", 0, fl), cpu_ref.RefArray("wall", 1, wl)],
                                dim=3, integrator="pc", co=prm["c0"], ro=prm["rho0"], gamma=7.0,
                                alpha=0.1, beta=0.0, gravity=(0.0, 0.0, prm["g"]),
                                density_from=(0, 1), momentum_from=(0, 1), nthreads=threads)
    t0 = time.perf_counter()
    ref.integrate(prm["dt"])
    t1 = time.perf_counter()
    # the reference's Cython loops are serial: one thread on a bounded sample
    # (one evaluation: both lists, SPHRho, Tait, momentum)
    ref1 = cpu_ref.RefSimulation([cpu_ref.RefArray("fluid", 0, fl),
                                  cpu_ref.RefArray("wall", 1, wl)],
                                 dim=3, integrator="pc", co=prm["c0"], ro=prm["rho0"], gamma=7.0,
                                 alpha=0.1, beta=0.0, gravity=(0.0, 0.0, prm["g"]),
                                 density_from=(0, 1), momentum_from=(0, 1), nthreads=1)
    acc = {d.name: {k: np.zeros(d.n) for k in ref1.stepped} for d in ref1.fluids}
    s0 = time.perf_counter()
    ref1.eval(acc)
    s1 = time.perf_counter()
    return {"value": ref.pairs / (t1 - t0), "unit": "pair-interactions/s", "cores": threads,
            "kind": "port",
            "sample": f"1 PredictorCorrector step of the same dam break ({ref.pairs} pair "
                      f"interactions, {t1 - t0:.2f} s, {threads} OpenMP threads, gcc -O2 "
                      f"-ffp-contract=off)",
            "steps_per_s": 1.0 / (t1 - t0),
            "single_thread": {"value": ref1.pairs / (s1 - s0), "cores": 1,
                              "sample": f"1 evaluation (both neighbour lists, SPHRho, Tait, "
                                        f"momentum) of the same dam break, {ref1.pairs} pair "
                                        f"interactions, {s1 - s0:.2f} s"}}


def cpu_baseline_dambreak_split(args, data):
    """The FP64 oracle on the host cores: one PredictorCorrector step of the
    operator-split dam break."""
    from oracle import cpu_ref
    fl, wl, prm = data
    threads = args.cpu_threads or len(os.sched_getaffinity(0))
    ref = cpu_ref.RefSimulation(
        [cpu_ref.RefArray("fluid", 0, fl), cpu_ref.RefArray("wall", 1, wl)], dim=3,
        integrator="pc", co=prm["c0"], ro=prm["rho0"], gamma=prm["gamma"], alpha=prm["alpha"],
        beta=prm["beta"], gravity=(0.0, 0.0, prm["g"]), density_from=(0, 1),
        momentum_from=(0, 1), nthreads=threads, density="continuity", xsph_eps=0.5,
        momentum="split", pgrad_from=(0,))
    t0 = time.perf_counter()
    ref.integrate(prm["dt"])
    t1 = time.perf_counter()
    return {"value": ref.pairs / (t1 - t0), "unit": "pair-interactions/s", "cores": threads,
            "kind": "port",
            "sample": f"1 PredictorCorrector step of the same operator-split dam break "
                      f"({ref.pairs} pair interactions, {t1 - t0:.2f} s, {threads} OpenMP "
                      f"threads)",
            "steps_per_s": 1.0 / (t1 - t0)}


def cpu_baseline_lattice(args, d):
    from oracle import cpu_ref
    import numpy as np
    threads = args.cpu_threads or len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    cs = cpu_ref.compute_cell_size(d["h"].max(), 2.0)
    if args.variable_h:  # the oracle's variable-h lists are serial
        off, idx = cpu_ref.nnps_varh(d, d, 2.0, cs)
    else:
        off, idx = cpu_ref.nnps(d, d, 2.0, cs, True, nthreads=threads)
    rho = np.zeros(len(d["x"]))
    cpu_ref.summation_density(3, d, d, off, idx, rho, nthreads=threads)
    t1 = time.perf_counter()
    return {"value": len(idx) / (t1 - t0), "unit": "pair-interactions/s", "cores": threads,
            "kind": "port", "sample": f"1 NNPS + SummationDensity pass of the same lattice "
                                      f"({t1 - t0:.2f} s, {threads} OpenMP threads"
                                      + (", lists built serially" if args.variable_h else "")
                                      + ")"}


def main():
    args = _args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: pysph_amd runs the hot path on a single MI355X "
              f"and has no multi-GPU path (asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    if args.workload == "dambreak":
        result, data = run_dambreak(args)
    elif args.workload == "dambreak-split":
        result, data = run_dambreak_split(args)
    elif args.workload == "lattice":
        result, data = run_lattice(args)
    elif args.workload == "batch":
        result, data = run_batch(args)
    else:
        result, data = run_adjoint(args)
    out = {"metric": result["metric"], "value": result["value"], "unit": result["unit"],
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": result["ms_per_step"],
           "higher_is_better": result.get("higher_is_better", True), "scaling": "weak",
           "vs_baseline": None, "dtype": result["dtype"], "data": "synthetic",
           "config": result["config"], "roofline": result["roofline"]}
    if not args.no_cpu_baseline and args.workload in ("dambreak", "dambreak-split", "lattice"):
        if args.workload == "dambreak":
            out["cpu_baseline"] = cpu_baseline_dambreak(args, data)
        elif args.workload == "dambreak-split":
            out["cpu_baseline"] = cpu_baseline_dambreak_split(args, data)
        else:
            out["cpu_baseline"] = cpu_baseline_lattice(args, data)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
