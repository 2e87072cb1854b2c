"""The cell manager keeps its padded grid across binnings (no bounding-box
readback) while the arrays are moved only by the integrator; a device check
of every binning sends it back to a full grid build when particles near the
grid's edge, and reports a binning that had particles outside it.  Keeping the
grid changes no result: cell ids are floor(x / cell_size) whatever the grid's
extent."""
import numpy as np
import pytest

from pysph_amd import _lib

pytestmark = pytest.mark.gpu

KEYS = ("x", "y", "z", "u", "v", "w", "rho", "p", "cs")


def _state(pa):
    return {k: pa.get(k) for k in KEYS}


def _bit_equal(a, b):
    return all(np.array_equal(a[k].view(np.int64), b[k].view(np.int64)) for k in a)


def test_kept_grid_dam_break_bit_equal(gpu):
    """The 3D dam break at dx = 0.04 (fluid + walls, PC), 30 steps, with the
    grid kept and with a grid build at every binning: the same bits."""
    from pysph_amd.solver.configs import dambreak_solver
    from pysph_amd.solver.generators import dam_break_3d
    data = dam_break_3d(dx=0.04)
    out, builds = [], []
    for keep in (True, False):
        s, particles, (_, _, prm) = dambreak_solver(data=data)
        cm = particles.cell_manager
        cm.keep_grid = keep
        for _ in range(30):
            s.integrator.integrate(prm["dt"])
        out.append(_state(particles.arrays[0]))
        builds.append(cm.grid_builds)
    print(f"\ndam break dx 0.04, 30 PC steps: grid builds {builds[0]} kept, {builds[1]} without")
    assert _bit_equal(out[0], out[1])
    assert builds[0] <= 2 and builds[1] >= 60


def test_moving_drop_regrids_and_stays_bit_equal(gpu):
    """A drop translating at 500 m/s (0.38 cells per step) crosses the padded
    grid many times over 40 steps: the device check triggers grid builds, and
    the run equals the one that rebuilds the grid at every binning."""
    from pysph_amd.solver.configs import drop_solver
    out, builds = [], []
    for keep in (True, False):
        s, particles, d = drop_solver(dx=0.05, vel=0.0, c0=10.0, dt=1e-4)
        fl = particles.arrays[0]
        fl.set(u=fl.get("u") + 500.0)
        cm = particles.cell_manager
        cm.keep_grid = keep
        for _ in range(40):
            s.integrator.integrate(1e-4)
        out.append(_state(fl))
        builds.append(cm.grid_builds)
    print(f"translating drop, 40 RK2 steps: grid builds {builds[0]} kept, {builds[1]} without; "
          f"x moved {out[0]['x'].mean():.3f}")
    assert _bit_equal(out[0], out[1])
    assert 2 < builds[0] < 40 and builds[1] >= 80


def test_particles_outside_a_kept_grid_are_reported(gpu):
    """Positions moved by a write that skipped touch() (so the manager takes
    them for an integrator step) far outside the grid: the binning's device
    check flags it and the next binning raises instead of going on silently."""
    from pysph_amd.solver.configs import drop_solver
    s, particles, d = drop_solver(dx=0.1)
    fl = particles.arrays[0]
    for _ in range(2):
        s.integrator.integrate(1e-5)
    x = fl.get_carray("x")
    x.add_(50.0)          # no touch(): the versions still match the last stage
    fl.set_dirty(True)
    with pytest.raises(_lib.SphError, match="outside the cell grid"):
        s.integrator.integrate(1e-5)
    # a write through the API rebuilds the grid and runs normally
    fl.set(x=fl.get("x"))
    s.integrator.integrate(1e-5)
    assert np.all(np.isfinite(fl.get("rho")))
