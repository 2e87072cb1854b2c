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
measured with HIP events on its stream (algorithmic bytes N(4 + 4 nbar +
15 f) per launch), its FP64 VALU view, the whole step's algorithmic bytes
per second, and the FP64 CPU oracle timed on the host cores on a bounded
sample of the same workload.  The default run also measures the batched
form (the 256-drop batch against sequential plain-path runs) and the
adjoint form (7,860-particle drop, 200 steps: forward, backward, plain-path
forward, central differences) in sub-objects ``batch`` and ``adjoint``.
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
# launch shape of the 256-drop batch picked by pysph_amd.autotune on an MI355X
BATCH_LAUNCH_SHAPE = {"cells_per_block": 8, "nbr_tile": 4}


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
    ap.add_argument("--no-extras", action="store_true",
                    help="default workload: skip the batch and adjoint sub-measurements")
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
VALU_FILE = os.path.join(ROOT, "profiles", "k_momentum_dambreak_1M_valu.json")
FP64_PEAK_TFLOPS = 78.6
# vector issue capacity of an MI355X: 256 CUs x 4 SIMDs, one wave64 VALU
# instruction per SIMD every 4 cycles (16 lanes), 2.4 GHz
SIMDS, CLOCK_HZ, CYCLES_PER_VALU = 1024, 2.4e9, 4


def _profile_json(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def measured_traffic(path=None):
    """HBM bytes per launch of the roofline kernel from the committed
    rocprofv3 byte counters (FETCH_SIZE, WRITE_SIZE, scaled by a 1 GiB copy),
    or None.  A profile of this bench command, not a measurement of this run."""
    d = _profile_json(path or TRAFFIC_FILE)
    try:
        return float(d["hbm_bytes_per_launch"])
    except (TypeError, KeyError, ValueError):
        return None


# Algorithmic bytes: the compulsory HBM traffic of each kernel, every array
# element touched once per launch and neighbour records served on chip.
# N destinations, nbar mean list entries per destination, f bytes per value
# (8: FP64), k permuted SoA properties, P stepped properties.
def alg_bytes(kind, N, nbar=0.0, f=8, k=8, P=6):
    return {"bin": N * (4 * f + 8 + 2 * k * f),          # read x y z h, key + perm, permute k
            "lists": N * (4 * f + 4 + 4 * nbar),         # read x y z h, count, 4 B per entry
            "density": N * (4 + 4 * nbar + 6 * f),       # count, entries, 5 props in, rho out
            "tait": 3 * f * N,                           # rho in, p cs out
            "momentum": N * (4 + 4 * nbar + 15 * f),     # count, entries, 11 props in, 4 out
            "stage": 3 * f * N * P,                      # x0, a in, x out
            "save": 2 * f * N * P,                       # x in, x0 out
            "pc_final": 3 * f * N * P}[kind]


def dambreak_step_bytes(nd, nnz_ff, nnz_fw, dim=3):
    """Algorithmic bytes of one PredictorCorrector step of the dam break:
    two evaluations (the fluid re-binned, fluid<-fluid and fluid<-wall lists,
    density, Tait, momentum), two stage updates, the save of the initial
    values and the final step.  The walls do not move and are binned once."""
    P = 2 * dim
    nb = (nnz_ff + nnz_fw) / nd
    ev = (alg_bytes("bin", nd) + alg_bytes("lists", nd, nnz_ff / nd) +
          alg_bytes("lists", nd, nnz_fw / nd) + alg_bytes("density", nd, nb) +
          alg_bytes("tait", nd) + alg_bytes("momentum", nd, nb))
    return 2 * ev + 2 * alg_bytes("stage", nd, P=P) + alg_bytes("save", nd, P=P) + \
        alg_bytes("pc_final", nd, P=P)


# FP64 operations of one MomentumEquation pair term (mom_term): 6 differences
# (r_ab, v_ab), v.r and |r|^2 (10), the denominator r^2 + eta^2 h^2 (4), mu (2
# with the division), the pressure term (2), the viscosity (cbar, rhobar,
# -alpha c mu + beta mu^2, the division: 10), the accumulation factor (2), the
# kernel gradient (|r| again with the square root 6, q, the branch numerator
# and denominator 4, the quotient and the factor 3) and the three components (9)
MOM_FLOPS_PER_PAIR = 6 + 10 + 4 + 2 + 2 + 10 + 2 + 13 + 9


def valu_block(mean_launch_s, pairs_per_launch):
    """FP64 VALU view of the momentum kernel: algorithmic FP64 operations per
    second against the FP64 vector peak, and the vector-issue floor of the
    instructions the committed rocprofv3 SQ_INSTS_VALU profile counted."""
    achieved = MOM_FLOPS_PER_PAIR * pairs_per_launch / mean_launch_s / 1e12
    out = {"kernel": "k_momentum", "bound": "valu", "achieved": achieved,
           "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
           "flops_per_pair": MOM_FLOPS_PER_PAIR}
    prof = _profile_json(VALU_FILE)
    if prof:
        insts = float(prof["valu_wave_insts_per_launch"])
        floor_s = insts * CYCLES_PER_VALU / SIMDS / CLOCK_HZ
        out.update({"valu_wave_insts_per_launch": insts,
                    "valu_lane_insts_per_pair": insts * 64 / pairs_per_launch,
                    "issue_floor_ms": 1e3 * floor_s,
                    "vector_issue_frac": floor_s / mean_launch_s,
                    "profile": os.path.relpath(VALU_FILE, ROOT)})
    return out


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
    profiling.enable(True, drain=True)
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
    profiling.enable(True, drain=True)
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
    # one launch per evaluation covers both lists and each destination once
    nnz = nnz_ff + nnz_fw
    bytes_per_launch = alg_bytes("momentum", nd, nnz / nd)
    achieved = bytes_per_launch / (mom["mean_ms"] * 1e-3) / 1e9
    step_bytes = dambreak_step_bytes(nd, nnz_ff, nnz_fw)
    step_gbs = step_bytes * args.steps / elapsed / 1e9
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
                              f"{min(5, args.steps)} steps outside the timed loop, the GPU "
                              "drained before each call (its kernels plus their launch "
                              "latency; nested calls count in both)",
        },
        "roofline": {
            "kernel": "k_momentum",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic() if args.dx == 0.01 else None,
            "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT) + " (rocprofv3 FETCH_SIZE and "
                              "WRITE_SIZE passes of this command; not measured in this run)",
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "bytes_model": "N(4 + 4 nbar + 15 f): list count, 4 B per entry, 11 destination "
                           "properties read and 4 written, f = 8; source records on chip",
            "mean_launch_ms": mom["mean_ms"],
            "valu": valu_block(mom["mean_ms"] * 1e-3, nnz),
            "step": {"algorithmic_bytes": step_bytes, "achieved": step_gbs,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": step_gbs / HBM_PEAK_GBS,
                     "note": "all kernels of a PC step: binning, both list builds, density, "
                             "Tait, momentum (2 evaluations) and the stage arithmetic"},
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
    profiling.enable(True, drain=True)
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


def batch_measure(nsims=256, precision="double", tune=None, autotune_steps=0, graph_steps=10,
                  steps=20, warmup=10, seq_sims=0, log=None):
    """The batched drops (radius U[0.4, 2], dx from {0.01, 0.02, 0.04},
    velocity scale U[50, 150], seeded), RK2 with each drop's time step
    1e-5*min(1, dx/0.025) (the CFL number of the dx = 0.025 drop), advanced
    together by captured HIP graphs of graph_steps steps; compared with
    running the same drops one after another through the plain path (all of
    them, or seq_sims spread over the size range, scaled to the whole set)."""
    import torch
    from pysph_amd.batch import BatchedSimulation, batched_drops
    from pysph_amd.solver.configs import drop_solver
    log = log or (lambda msg: None)
    sims, prm = batched_drops(nsims=nsims)
    log(f"{len(sims)} drops generated, {sum(len(d['x']) for d in sims)} particles")
    b = BatchedSimulation(sims, dt_per_sim=prm["dt"], precision=precision)
    dt = 1.0  # multiplies every drop's own time step
    log(f"packed: key_cap {b.key_cap} max_cells {b.max_cells} list_cap {b.list_cap} "
        f"workspace {b.work.numel() / 2**20:.0f} MiB")
    b.step(dt)
    b.check()
    log("eager step done")
    tuned = None
    if tune:
        b.set_tuning(tune)
        log(f"launch shape {b.tuning()}")
    elif autotune_steps:
        from pysph_amd.autotune import autotune
        tuned = autotune(b, dt, steps=autotune_steps, use_cache=False, log=log)
        log(f"autotuned: {tuned['tuning']} {tuned['ms_per_step']:.3f} ms/step "
            f"(default shape {tuned['trials'][0][1]:.3f} ms/step)")
    gs = max(1, graph_steps)
    b.capture(dt, nsteps=gs)
    log(f"graph of {gs} steps captured")
    for _ in range(max(1, warmup // gs)):
        b.replay()
    b.check()
    pairs0 = b.pair_interactions()
    nrep = max(1, steps // gs)
    t0 = time.perf_counter()
    for _ in range(nrep):
        b.replay()
    b.synchronize()
    t1 = time.perf_counter()
    b.check()
    pairs = b.pair_interactions() - pairs0
    nsteps = nrep * gs
    elapsed = t1 - t0
    batch_sps = nsteps / elapsed
    log(f"timed {nsteps} steps: {1e3 * elapsed / nsteps:.3f} ms/step")
    order = np.argsort([len(d["x"]) for d in sims])
    pick = order if not seq_sims else order[np.linspace(0, len(order) - 1,
                                                        seq_sims).astype(int)]
    seq_time = 0.0
    seq_steps = 2
    for k in pick:
        solver, particles, _ = drop_solver(data=dict(sims[k]), dx=float(prm["dx"][k]),
                                           precision=precision)
        dtk = float(prm["dt"][k])
        solver.integrator.integrate(dtk)          # first build, allocations
        torch.cuda.synchronize()
        ta = time.perf_counter()
        for _ in range(seq_steps):
            solver.integrator.integrate(dtk)
        torch.cuda.synchronize()
        seq_time += (time.perf_counter() - ta) / seq_steps
    seq_time *= len(sims) / len(pick)
    seq_sps = 1.0 / seq_time
    log(f"sequential plain path: {len(pick)} drops timed, {seq_sps:.2f} steps/s of the set")
    res = {"steps_per_s": batch_sps, "ms_per_step": 1e3 * elapsed / nsteps,
           "timed_steps": nsteps, "precision": precision,
           "particles": b.n, "drops": len(sims),
           "smallest_drop": int(min(b.counts)), "largest_drop": int(max(b.counts)),
           "pair_interactions_per_s": pairs / elapsed,
           "sequential_plain_gpu_steps_per_s": seq_sps,
           "sequential_drops_timed": int(len(pick)),
           "speedup_vs_sequential": batch_sps / seq_sps,
           "graph_steps": gs, "launch_shape": b.tuning()}
    if tuned is not None:
        res["autotune"] = {"trials": len(tuned["trials"]),
                           "default_shape_ms_per_step": tuned["trials"][0][1],
                           "tuned_ms_per_step_eager": tuned["ms_per_step"]}
    return res


def run_batch(args):
    """The 256-drop batch on its own (--workload batch)."""
    t_start = time.perf_counter()

    def progress(msg):
        print(f"[batch {time.perf_counter() - t_start:7.2f}s] {msg}", file=sys.stderr, flush=True)

    tune = None
    if args.launch_shape:
        tune = {k: int(v) for k, v in (kv.split("=") for kv in args.launch_shape.split(","))}
    r = batch_measure(nsims=args.nsims, precision=args.precision, tune=tune,
                      autotune_steps=0 if (tune or args.no_autotune) else 3,
                      graph_steps=args.graph_steps, steps=args.steps, warmup=args.warmup,
                      seq_sims=args.seq_sims, log=progress)
    cfg = dict(r)
    cfg["workload"] = (f"{r['drops']} independent 2D elliptical drops, RK2, "
                       f"{'FP32 operators' if args.precision == 'single' else 'FP64'}, "
                       f"dt 1e-5*min(1, dx/0.025) per drop, captured HIP graphs of "
                       f"{r['graph_steps']} steps")
    cfg["fused_kernels"] = "density+EOS, momentum+position stepping+stage"
    return {
        "metric": "steps/s",
        "value": r["steps_per_s"],
        "unit": "steps/s (all drops advanced one RK2 step)",
        "ms_per_step": r["ms_per_step"],
        "dtype": "f32" if args.precision == "single" else "f64",
        "config": cfg,
        "roofline": None,
    }, None


def adjoint_measure(fd_particles=64, log=None, reps=3):
    """Gradient of the final-position MSE of the 2D drop (dx = 0.02, 7,860
    particles, RK2, dt = 1e-5, 200 steps, checkpoint every 20) against a
    target run from u = -120 x, v = 120 y, with respect to the per-particle
    initial velocities (around u = -100 x, v = 100 y plus seeded N(0, 5)) and
    alpha: forward and backward time, the same 200 steps through the plain
    per-simulation path, and central differences (eps = 1e-6) on
    fd_particles random particles and alpha.  Returns (summary, state for
    the fits)."""
    import torch
    from pysph_amd.adjoint import Rollout, rollout
    from pysph_amd.batch import BatchedSimulation
    from pysph_amd.solver.configs import drop_solver
    from pysph_amd.solver.generators import elliptical_drop
    log = log or (lambda msg: None)
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
    log(f"{n} particles, target run done")

    def loss_fn(x, y):
        return ((x - tx) ** 2 + (y - ty) ** 2).mean()

    u0 = torch.tensor(u_init, device=dev, requires_grad=True)
    v0 = torch.tensor(v_init, device=dev, requires_grad=True)
    alpha = torch.tensor(0.1, dtype=torch.float64, device=dev, requires_grad=True)
    x, y = rollout(roll, alpha, u0, v0)   # warm-up
    torch.autograd.grad(loss_fn(x, y), (u0, v0, alpha))
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
    log(f"forward {1e3 * t_fwd:.1f} ms, backward {1e3 * t_bwd:.1f} ms")
    # the same 200 steps through the plain per-simulation path (Solver)
    dd = dict(d)
    dd["u"], dd["v"] = u_init.copy(), v_init.copy()
    solver, particles, _ = drop_solver(data=dd, dx=dx)
    solver.integrator.integrate(dt)   # first build, allocations
    torch.cuda.synchronize()
    ta = time.perf_counter()
    for _ in range(steps):
        solver.integrator.integrate(dt)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - ta
    # central differences: every perturbed rollout is one simulation of a batch
    picks = rng.choice(n, fd_particles, replace=False)
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
    d0 = dict(d)
    d0["u"], d0["v"] = u_init.copy(), v_init.copy()

    def central(e):
        """central differences at step e: the velocity components, then alpha"""
        sims_e = []
        for i, comp, sgn in which:
            de = dict(d)
            de["u"], de["v"] = u_init.copy(), v_init.copy()
            de[comp][i] += sgn * e
            sims_e.append(de)
        L = batch_losses(sims_e, 0.1)
        fd = [(L[k] - L[k + 1]) / (2 * e) for k in range(0, len(sims_e), 2)]
        la = batch_losses([d0], 0.1 + e)[0]
        lb = batch_losses([d0], 0.1 - e)[0]
        return np.array(fd + [(la - lb) / (2 * e)])

    g_all = np.array([gmap[comp][i] for i, comp, _ in which[::2]] + [float(ga)])
    fd1 = central(eps)
    rel = np.abs(g_all - fd1) / np.maximum(np.abs(fd1), 1e-300)
    ga_v = float(ga)
    # alpha: the rollout's central difference at smaller steps (over 200
    # steps the loss is strongly curved in alpha at the 1e-6 scale and kinked
    # where viscosity switches flip)
    alpha_sweep = {}
    for e in (2.5e-7, 6.25e-8, 3.125e-8):
        la = batch_losses([d0], 0.1 + e)[0]
        lb = batch_losses([d0], 0.1 - e)[0]
        fda = (la - lb) / (2 * e)
        alpha_sweep[f"{e:g}"] = abs(ga_v - fda) / abs(fda)
    summary = {
        "workload": f"2D elliptical drop dx {dx} ({n} particles), RK2 dt {dt}, {steps} steps, "
                    f"checkpoint every {ckpt}, FP64",
        "forward_ms": 1e3 * t_fwd, "backward_ms": 1e3 * t_bwd,
        "backward_over_forward": t_bwd / t_fwd,
        "gradient_over_forward": (t_fwd + t_bwd) / t_fwd,
        "plain_gpu_forward_ms": 1e3 * t_plain,
        "forward_over_plain_gpu": t_fwd / t_plain,
        "fd": {"eps": eps, "velocity_components": int(len(g_all) - 1), "grad_alpha": ga_v,
               "velocity_norm_rel_err": float(np.linalg.norm(g_all[:-1] - fd1[:-1]) /
                                              np.linalg.norm(fd1[:-1])),
               "velocity_median_rel_err": float(np.median(rel[:-1])),
               "velocity_max_rel_err": float(rel[:-1].max()),
               "alpha_rel_err": float(rel[-1]),
               "alpha_rel_err_smaller_eps": alpha_sweep,
               "note": "adjoint vs central differences of the rollout; the velocity gradient "
                       "agrees as a vector at eps 1e-6; for alpha the 200-step loss is "
                       "strongly curved at that scale and kinked where viscosity switches "
                       "flip, so the difference approaches the adjoint only at smaller eps "
                       "(the tests compare it with the switches held fixed)"},
    }
    log(f"FD eps {eps}: {summary['fd']}")
    state = dict(roll=roll, d=d, n=n, dev=dev, tx=tx, ty=ty, u_init=u_init, v_init=v_init,
                 loss=loss, loss_fn=loss_fn, steps=steps, dt=dt, rollout=rollout)
    return summary, state


def run_adjoint(args):
    """cfg 4 on its own (--workload adjoint): the measurements of
    adjoint_measure, then 20-iteration L-BFGS fits."""
    import torch
    t_start = time.perf_counter()

    def progress(msg):
        print(f"[adjoint {time.perf_counter() - t_start:7.2f}s] {msg}", file=sys.stderr,
              flush=True)

    summary, st = adjoint_measure(args.fd_particles, progress)
    roll, d, n, dev = st["roll"], st["d"], st["n"], st["dev"]
    u_init, v_init, loss, loss_fn = st["u_init"], st["v_init"], st["loss"], st["loss_fn"]
    steps, dt, rollout = st["steps"], st["dt"], st["rollout"]
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
    cfg = dict(summary)
    cfg.update({"lbfgs_iterations": args.lbfgs_iters,
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
                    "v_y_coefficient": 100.0 + s_c * float(cv[2])}})
    return {
        "metric": "adjoint/forward cost ratio",
        "value": summary["backward_over_forward"],
        "unit": "backward time / forward time",
        "ms_per_step": summary["forward_ms"] / steps,
        "higher_is_better": False,
        "dtype": "f64",
        "config": cfg,
        "roofline": None,
    }, None


def cpu_info():
    """Host CPU model and the oracle's compiler flags, for the baseline line."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    from oracle import build as refbuild
    return {"cpu_model": model, "compiler": "gcc " + " ".join(
        c for c in refbuild.CFLAGS if c.startswith(("-O", "-f")))}


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
                      f"interactions, {t1 - t0:.2f} s, {threads} OpenMP threads, gcc -O3 "
                      f"-ffp-contract=off)",
            "steps_per_s": 1.0 / (t1 - t0), **cpu_info(),
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
            "steps_per_s": 1.0 / (t1 - t0), **cpu_info()}


def cpu_baseline_lattice(args, d):
    """The oracle's NNPS + SummationDensity pass over the same lattice.  The
    variable-h lists of the oracle are built by one thread, so with
    --variable-h the list build is timed (and reported) on its own at one
    thread and the density pass at all cores."""
    from oracle import cpu_ref
    threads = args.cpu_threads or len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    cs = cpu_ref.compute_cell_size(d["h"].max(), 2.0)
    if args.variable_h:
        off, idx = cpu_ref.nnps_varh(d, d, 2.0, cs)
    else:
        off, idx = cpu_ref.nnps(d, d, 2.0, cs, True, nthreads=threads)
    t_lists = time.perf_counter() - t0
    rho = np.zeros(len(d["x"]))
    cpu_ref.summation_density(3, d, d, off, idx, rho, nthreads=threads)
    t1 = time.perf_counter()
    out = {"value": len(idx) / (t1 - t0), "unit": "pair-interactions/s", "cores": threads,
           "kind": "port", "sample": f"1 NNPS + SummationDensity pass of the same lattice "
                                     f"({t1 - t0:.2f} s, {threads} OpenMP threads)",
           **cpu_info()}
    if args.variable_h:
        out["cores"] = 1
        out["sample"] = (f"1 NNPS + SummationDensity pass of the same lattice ({t1 - t0:.2f} s): "
                         f"variable-h lists built serially ({t_lists:.2f} s, 1 thread), "
                         f"density on {threads} OpenMP threads ({t1 - t0 - t_lists:.2f} s)")
        out["lists_s_one_thread"] = t_lists
        out["density_s"] = t1 - t0 - t_lists
        out["density_threads"] = threads
    return out


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
    if args.workload == "dambreak" and not args.no_extras:
        # the batched and adjoint forms beside the plain path, bounded (about
        # a minute): the 256-drop batch against sequential plain-path runs of
        # 32 of its drops, and the adjoint configuration's costs and FD check
        t_start = time.perf_counter()

        def log(msg):
            print(f"[extras {time.perf_counter() - t_start:7.2f}s] {msg}", file=sys.stderr,
                  flush=True)

        out["batch"] = batch_measure(nsims=256, tune=BATCH_LAUNCH_SHAPE, graph_steps=10,
                                     steps=20, warmup=10, seq_sims=32, log=log)
        out["batch"]["launch_shape_note"] = ("the autotuner's pick for this batch "
                                             "(--workload batch runs the autotuner)")
        out["adjoint"] = adjoint_measure(64, log)[0]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
