"""GPU parity of the batched path: every packed simulation must follow the
single-simulation path (drop_solver) exactly -- same lists, same sums, so
the bar is bit equality (max relative error printed) -- and graph replays
must reproduce eager steps."""
import numpy as np
import pytest

from tests.helpers import bit_equal_frac, rel_err

pytestmark = pytest.mark.gpu

KEYS = ("x", "y", "u", "v", "rho", "p", "cs", "dt_fac")


def _small_drops(n=5, seed=7):
    from pysph_amd.batch import batched_drops
    return batched_drops(nsims=n, seed=seed, radius_range=(0.25, 0.6), dxs=(0.03, 0.04, 0.05),
                         vel_range=(50.0, 150.0))


def test_batch_matches_single_simulation_runs(gpu):
    from pysph_amd.batch import BatchedSimulation
    from pysph_amd.solver.configs import drop_solver
    sims, prm = _small_drops()
    steps = 4
    dts = np.array([1e-5, 7e-6, 1e-5, 5e-6, 9e-6])[:len(sims)]
    b = BatchedSimulation(sims, dt_per_sim=dts)
    for _ in range(steps):
        b.step()
    b.check()
    worst = 0.0
    for s, d in enumerate(sims):
        solver, particles, _ = drop_solver(data=dict(d), dx=float(prm["dx"][s]))
        for _ in range(steps):
            solver.integrator.integrate(float(dts[s]))
        fl = particles.arrays[0]
        for k in KEYS:
            e = rel_err(b.get(k, s), fl.get(k))
            worst = max(worst, e)
            assert bit_equal_frac(b.get(k, s), fl.get(k)) == 1.0, (s, k, e)
    print(f"batch vs single-simulation runs: {len(sims)} drops, {steps} RK2 steps, "
          f"max rel err {worst:.3g} (bit-equal)")
    assert b.pair_interactions() > 0


@pytest.mark.parametrize("persistent", [0, 1])
def test_batch_graph_replays_match_eager(gpu, persistent):
    """Graph replays (one persistent launch per evaluation, or two) against
    eager steps of the other kind: bit-equal every replay."""
    from pysph_amd.batch import BatchedSimulation
    sims, _ = _small_drops(n=4, seed=9)
    dt = 1e-5
    eager = BatchedSimulation(sims, tune={"persistent": 1 - persistent})
    graph = BatchedSimulation(sims, tune={"persistent": persistent})
    eager.step(dt)
    eager.check()
    graph.capture(dt, nsteps=1)  # runs the first step eagerly, then records one step
    for r in range(4):
        graph.replay()
        graph.check()
        eager.step(dt)
        eager.check()
        for k in KEYS:
            a, b = graph.get(k), eager.get(k)
            assert np.array_equal(a, b), (r, k, rel_err(a, b))
    print(f"graph replays (persistent={persistent}): 4, every replay bit-equal to the eager "
          f"step")
    graph.release_graph()


def test_batch_capacity_overflow_is_reported(gpu):
    from pysph_amd.batch import BatchedSimulation, BatchError
    sims, _ = _small_drops(n=2, seed=3)
    b = BatchedSimulation(sims, list_cap=64)
    b.step(1e-5)
    with pytest.raises(BatchError, match="list_cap"):
        b.check()


def test_batch_void_simulation_is_reported_and_cheap(gpu):
    """A simulation with a non-finite coordinate is flagged and its particles
    get no lists; the others still match their single-simulation runs."""
    import time
    from pysph_amd.batch import BatchedSimulation, BatchError
    from pysph_amd.solver.configs import drop_solver
    sims, prm = _small_drops(n=3, seed=11)
    bad = dict(sims[1])
    bad["x"] = bad["x"].copy()
    bad["x"][5] = np.nan
    b = BatchedSimulation([sims[0], bad, sims[2]])
    t0 = time.perf_counter()
    b.step(1e-5)
    b.synchronize()
    assert time.perf_counter() - t0 < 30.0
    with pytest.raises(BatchError, match="non-finite"):
        b.check()
    for s in (0, 2):
        solver, particles, _ = drop_solver(data=dict(sims[s]), dx=float(prm["dx"][s]))
        solver.integrator.integrate(1e-5)
        for k in ("x", "rho"):
            assert np.array_equal(b.get(k, s), particles.arrays[0].get(k)), (s, k)


# -- fused step, launch shapes, autotuner ------------------------------------
ALL_STATE = ("x", "y", "z", "u", "v", "w", "rho", "p", "cs", "x0", "y0", "z0", "u0", "v0",
             "w0", "ax", "ay", "az", "au", "av", "aw", "dt_fac")


def _cubes(n=3, seed=5):
    """3D blobs: jittered cubic lattices of different sizes and spacings
    with random velocities (Tait fluid at rest density)."""
    rng = np.random.default_rng(seed)
    sims = []
    for k in range(n):
        dx = (0.05, 0.04, 0.06)[k % 3]
        m = 6 + 2 * k
        g = (np.arange(m) + 0.5) * dx
        x, y, z = [a.ravel() for a in np.meshgrid(g, g, g, indexing="ij")]
        x = x + rng.uniform(-0.1, 0.1, x.size) * dx
        N = x.size
        sims.append(dict(x=x, y=y, z=z, u=rng.normal(0, 1.0, N), v=rng.normal(0, 1.0, N),
                         w=rng.normal(0, 1.0, N), h=np.full(N, 1.3 * dx),
                         m=np.full(N, dx ** 3), rho=np.ones(N), p=np.zeros(N),
                         cs=np.full(N, 20.0)))
    return sims


def _pieces_step(b, dt):
    """One RK2 step kernel by kernel (save, eval, stage, eval, stage)."""
    import ctypes
    lib, hd, d = b.ctx.lib, b.ctx.handle, ctypes.byref(b.desc)
    for name, args in (("sph_batch_save_initial", ()), ("sph_batch_eval", ()),
                       ("sph_batch_stage", (dt / 2,)), ("sph_batch_eval", ()),
                       ("sph_batch_stage", (dt,))):
        b.ctx.check(getattr(lib, name)(hd, d, *args))


@pytest.mark.parametrize("dim,precision", [(2, "double"), (2, "single"), (3, "double")])
def test_fused_step_matches_kernel_by_kernel(gpu, dim, precision):
    """sph_batch_rk2_step (two fused kernels per evaluation) against the
    separate kernels, for several launch shapes: bit-equal in every field."""
    from pysph_amd.batch import BatchedSimulation
    if dim == 2:
        sims, _ = _small_drops(n=4, seed=13)
        dts, co = np.array([1e-5, 8e-6, 1e-5, 6e-6]), 1400.0
    else:
        sims, dts, co = _cubes(), None, 20.0
    tunes = [{}, {"op_threads": 64, "op_blocks_per_cu": 4, "nbr_tile": 1, "cells_per_block": 3},
             {"op_threads": 128, "op_blocks_per_cu": 1, "nbr_tile": 4, "cells_per_block": 8},
             {"persistent": 1},
             {"persistent": 1, "op_threads": 64, "op_blocks_per_cu": 2, "nbr_tile": 1},
             {"persistent": 1, "op_threads": 128, "nbr_tile": 4},
             {"staged": 1}, {"staged": 1, "cells_per_block": 4}]
    ref = BatchedSimulation(sims, dim=dim, co=co, dt_per_sim=dts, precision=precision)
    runs = [BatchedSimulation(sims, dim=dim, co=co, dt_per_sim=dts, precision=precision,
                              tune=t) for t in tunes]
    dt = 1.0 if dts is not None else 1e-4
    for _ in range(3):
        _pieces_step(ref, dt)
        for r in runs:
            r.step(dt)
    ref.check()
    for t, r in zip(tunes, runs):
        r.check()
        for k in ALL_STATE:
            if dim == 2 and k in ("z", "w", "z0", "w0", "az", "aw"):
                continue
            a, e = r.get(k), ref.get(k)
            assert np.array_equal(a, e), (t, k, rel_err(a, e))
    print(f"fused step dim {dim} {precision}: {len(tunes)} launch shapes bit-equal to the "
          f"kernel-by-kernel step")


def test_autotune_restores_state_and_keeps_results(gpu, tmp_path, monkeypatch):
    from pysph_amd.autotune import autotune
    from pysph_amd.batch import BatchedSimulation
    monkeypatch.setenv("PYSPH_AMD_CACHE", str(tmp_path))
    sims, _ = _small_drops(n=3, seed=21)
    a = BatchedSimulation(sims)
    b = BatchedSimulation(sims)
    a.step(1e-5)
    b.step(1e-5)
    before = {k: a.get(k) for k in KEYS}
    space = {"op_threads": (64, 256), "nbr_tile": (1, 2), "cells_per_block": (1, 4),
             "op_blocks_per_cu": (0, 8)}
    res = autotune(a, dt=1e-5, steps=1, space=space)
    assert not res["cached"] and len(res["trials"]) >= 5
    for k in KEYS:
        assert np.array_equal(a.get(k), before[k]), k
    for _ in range(2):
        a.step(1e-5)
        b.step(1e-5)
    a.check()
    b.check()
    for k in KEYS:
        assert np.array_equal(a.get(k), b.get(k)), k
    res2 = autotune(BatchedSimulation(sims), dt=1e-5, steps=1, space=space)
    assert res2["cached"] and res2["tuning"] == res["tuning"]
    print(f"autotune: {res['tuning']} {res['ms_per_step']:.3f} ms/step, "
          f"{len(res['trials'])} trials; state restored, results unchanged")


def test_variable_h_batch_matches_single_simulation(gpu):
    """h varying inside a simulation (uniform_h off): the source h is read
    and h_ab = (h_a + h_b)/2 as in the single-simulation path."""
    from pysph_amd.batch import BatchedSimulation
    from pysph_amd.solver.configs import drop_solver
    sims, prm = _small_drops(n=3, seed=17)
    rng = np.random.default_rng(3)
    for d in sims:
        d["h"] = d["h"] * (1.0 + 0.05 * rng.uniform(-1, 1, len(d["h"])))
    b = BatchedSimulation(sims)
    assert not b.uniform_h
    for _ in range(2):
        b.step(5e-6)
    b.check()
    for s, d in enumerate(sims):
        solver, particles, _ = drop_solver(data=dict(d), dx=float(prm["dx"][s]))
        for _ in range(2):
            solver.integrator.integrate(5e-6)
        for k in KEYS:
            a, e = b.get(k, s), particles.arrays[0].get(k)
            assert np.array_equal(a, e), (s, k, rel_err(a, e))
    print("variable-h batch: bit-equal to single-simulation runs")


def test_uniform_h_promise_is_checked(gpu):
    from pysph_amd.batch import BatchedSimulation, BatchError
    sims, _ = _small_drops(n=2, seed=19)
    sims[1]["h"] = sims[1]["h"].copy()
    sims[1]["h"][3] *= 1.01
    b = BatchedSimulation(sims)
    assert not b.uniform_h
    b.desc.uniform_h = 1  # a false promise: the grid kernel flags it
    b.step(1e-5)
    with pytest.raises(BatchError, match="h differs"):
        b.check()
