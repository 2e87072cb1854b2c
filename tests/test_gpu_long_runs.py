"""Whole trajectories on the GPU against the FP64 oracle, compared after every
step: the 2D elliptical drop for all 1,000 RK2 steps of its configuration,
the 1M-fluid dam break for 20 PredictorCorrector steps at full size, and the
same dam break at dx = 0.04 for the configuration's 500 steps.

Per step the largest relative error (tests.helpers.rel_err) of the density,
the velocity accelerations of the last evaluation and the positions is
recorded and printed as a table.  The kernels reproduce the oracle's
arithmetic operation for operation, the Tait power included (both take x^7
and x^3 correctly rounded), so the trajectories stay equal bit for bit.  With
any one-ulp difference per step (the C library's pow misrounds about one call
in a thousand) the drop's errors grow about tenfold every 20 steps and pass
1e-6 within about 110 steps: the flow amplifies perturbations.
Bar: equal bit for bit at every step (well inside the 1e-6 relative bound)."""
import numpy as np
import pytest

from oracle import cpu_ref
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-6


def _table(label, rows, every):
    print(f"\n{label}: step | rho | accelerations | positions (max relative error)")
    for r in rows:
        if r[0] == 1 or r[0] % every == 0 or r[0] == rows[-1][0]:
            print(f"  {r[0]:5d} | {r[1]:.3g} | {r[2]:.3g} | {r[3]:.3g}")
    worst = np.max(np.array([r[1:] for r in rows]), axis=0)
    print(f"  max   | {worst[0]:.3g} | {worst[1]:.3g} | {worst[2]:.3g}")
    return worst


def _errors(arr, ref_arr, acc, dim, step):
    vel, pos = ("u", "v", "w")[:dim], ("x", "y", "z")[:dim]
    e_rho = rel_err(arr.get("rho"), ref_arr["rho"])
    e_acc = max(rel_err(arr.get(f"_a_{c}_1"), acc[c]) for c in vel)
    e_pos = max(rel_err(arr.get(c), ref_arr[c]) for c in pos)
    return (step, e_rho, e_acc, e_pos)


def test_drop_1000_rk2_steps_vs_oracle(gpu):
    """The 2D elliptical drop (5,025 particles, dx = 0.025, RK2, dt = 1e-5)
    for its 1,000 steps."""
    from pysph_amd.solver.configs import drop_solver
    s, particles, d = drop_solver(dx=0.025, integrator="rk2")
    fl = particles.arrays[0]
    ref = cpu_ref.RefSimulation([cpu_ref.RefArray("fluid", 0, d)], dim=2, integrator="rk2",
                                co=1400.0, ro=1.0, gamma=7.0, alpha=0.1, beta=0.0, nthreads=8)
    rows = []
    for k in range(1, 1001):
        s.integrator.integrate(1e-5)
        acc = ref.integrate(1e-5)["fluid"]
        rows.append(_errors(fl, ref.arrays[0], acc, 2, k))
    worst = _table("2D drop, 1000 RK2 steps", rows, 100)
    assert np.all(worst <= TOL), worst
    assert np.all(worst == 0.0), worst  # the same arithmetic, pow included


def _dambreak_run(dx, steps, every, threads):
    from pysph_amd.solver.configs import dambreak_solver
    from pysph_amd.solver.generators import dam_break_3d
    data = dam_break_3d(dx=dx)
    s, particles, (fl, wl, prm) = dambreak_solver(data=data)
    fluid = particles.arrays[0]
    ref = cpu_ref.RefSimulation([cpu_ref.RefArray("fluid", 0, fl),
                                 cpu_ref.RefArray("wall", 1, wl)],
                                dim=3, integrator="pc", co=prm["c0"], ro=prm["rho0"],
                                gamma=7.0, alpha=0.1, beta=0.0, gravity=(0.0, 0.0, prm["g"]),
                                density_from=(0, 1), momentum_from=(0, 1), nthreads=threads)
    rows = []
    for k in range(1, steps + 1):
        s.integrator.integrate(prm["dt"])
        acc = ref.integrate(prm["dt"])["fluid"]
        rows.append(_errors(fluid, ref.arrays[0], acc, 3, k))
    return _table(f"3D dam break dx {dx} ({len(fl['x'])} fluid, {len(wl['x'])} wall), "
                  f"{steps} PC steps", rows, every)


@pytest.mark.timeout(600)
def test_dambreak_full_size_20_pc_steps_vs_oracle(gpu):
    import os
    worst = _dambreak_run(0.01, 20, 5, min(16, len(os.sched_getaffinity(0))))
    assert np.all(worst <= TOL), worst
    assert np.all(worst == 0.0), worst  # the same arithmetic, pow included


@pytest.mark.timeout(600)
def test_dambreak_500_pc_steps_vs_oracle(gpu):
    import os
    worst = _dambreak_run(0.04, 500, 50, min(16, len(os.sched_getaffinity(0))))
    assert np.all(worst <= TOL), worst
    assert np.all(worst == 0.0), worst  # the same arithmetic, pow included
