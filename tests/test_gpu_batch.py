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


def test_batch_graph_replays_match_eager(gpu):
    from pysph_amd.batch import BatchedSimulation
    sims, _ = _small_drops(n=4, seed=9)
    dt = 1e-5
    eager = BatchedSimulation(sims)
    graph = BatchedSimulation(sims)
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
    print("graph replays: 4, every replay bit-equal to the eager step")
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
