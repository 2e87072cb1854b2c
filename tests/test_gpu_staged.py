"""The LDS-staged density and momentum kernels (linear cell order) against
the gather kernels on the same lists: every output bit-equal, across the cases
the staging has to get right -- several source arrays, variable h, pencils
whose records exceed the LDS capacity (crowded cells), a stencil reach of
more than one cell (cells clamped below the radius), non-LocalReal
destinations, 2D and 3D, whole integrator steps -- and the staged path
against the oracle."""
import numpy as np
import pytest

from oracle import cpu_ref
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


def _both(run):
    """run() twice, staged kernels on and off; returns both results."""
    from pysph_amd.sph import funcs
    old = funcs.STAGED
    try:
        funcs.STAGED = True
        a = run()
        funcs.STAGED = False
        b = run()
    finally:
        funcs.STAGED = old
    return a, b


def _assert_bit_equal(a, b, label):
    for k in a:
        x, y = np.asarray(a[k]), np.asarray(b[k])
        assert x.shape == y.shape, (label, k)
        same = np.array_equal(x.view(np.int64), y.view(np.int64))
        assert same, (label, k, rel_err(x, y))


def _operators(d, dim=3, tag=None, max_cell_size=0.0, alpha=0.3, beta=0.2):
    from pysph_amd.base import api as base
    from pysph_amd.sph import api as sph
    kw = dict(d)
    if tag is not None:
        kw["tag"] = tag
    pa = base.get_particle_array(name="fluid", **kw)
    particles = base.Particles([pa], cell_order="linear", max_cell_size=max_cell_size)
    kernel = base.CubicSplineKernel(dim=dim)
    rho_f = sph.SPHRho.withargs().get_func(pa, pa)
    sph.SPHCalc(particles, [pa], pa, kernel, [rho_f], ['rho']).sph('rho')
    mom = sph.MomentumEquation.withargs(alpha=alpha, beta=beta).get_func(pa, pa)
    calc = sph.SPHCalc(particles, [pa], pa, kernel, [mom], ['_tmpx', '_tmpy', '_tmpz'],
                       integrates=True)
    for k in ('_tmpx', '_tmpy', '_tmpz'):
        pa.get_carray(k).zero_()
    calc.sph('_tmpx', '_tmpy', '_tmpz')
    assert particles.cell_manager.grid.order == 1  # linear
    return {k: pa.get(k) for k in ('rho', '_tmpx', '_tmpy', '_tmpz', 'dt_fac')}


def _state(d, rng, vel=1.0):
    n = len(d["x"])
    d = dict(d)
    d["u"], d["v"], d["w"] = vel * rng.standard_normal((3, n))
    d["p"] = rng.uniform(-1e3, 1e4, n)
    d["cs"] = 50.0 * rng.uniform(0.99, 1.01, n)
    d["rho"] = 1000.0 * rng.uniform(0.98, 1.02, n)
    return d


@pytest.mark.parametrize("case", ["lattice", "variable_h", "crowded", "reach2", "non_real",
                                  "2d"])
def test_staged_equals_gather(gpu, case):
    from pysph_amd.solver.generators import jittered_lattice_3d
    rng = np.random.default_rng(41)
    dim, tag, mcs = 3, None, 0.0
    if case == "lattice":
        d = _state(jittered_lattice_3d(n=24, seed=3, hdx=1.3), rng)
    elif case == "variable_h":
        d = _state(jittered_lattice_3d(n=22, seed=5, hdx=1.3), rng)
        d["h"] = d["h"] * rng.uniform(0.85, 1.15, len(d["h"]))
    elif case == "crowded":
        # h = 3 dx: cells of 6 dx hold ~216 particles, a pencil's three cells
        # ~650 records, beyond the LDS capacity (global reads for those)
        d = _state(jittered_lattice_3d(n=18, seed=7, hdx=3.0), rng)
    elif case == "reach2":
        d = _state(jittered_lattice_3d(n=20, seed=9, hdx=1.3), rng)
        mcs = 1.3 * (1.0 / 20)   # cells of h: a reach of two cells
    elif case == "non_real":
        d = _state(jittered_lattice_3d(n=20, seed=11, hdx=1.3), rng)
        n = len(d["x"])
        tag = np.zeros(n, dtype=np.int64)
        tag[3 * n // 4:] = 1
    else:
        from pysph_amd.solver.generators import elliptical_drop
        d = _state(elliptical_drop(dx=0.02, radius=0.8), rng)
        d["z"] = np.zeros(len(d["x"]))
        d["w"] = np.zeros(len(d["x"]))
        dim = 2
    a, b = _both(lambda: _operators(d, dim=dim, tag=tag, max_cell_size=mcs))
    _assert_bit_equal(a, b, case)
    print(f"{case}: staged density and momentum bit-equal to the gather kernels "
          f"({len(d['x'])} particles)")


def test_staged_dambreak_steps_equal_gather_and_oracle(gpu):
    """Fluid and wall sources in one launch, three PredictorCorrector steps of
    the small dam break: every stepped property bit-equal between the two
    kernel sets, and within 1e-10 of the oracle."""
    from pysph_amd.solver.configs import dambreak_solver
    from pysph_amd.solver.generators import dam_break_3d
    data = dam_break_3d(dx=0.05)

    def run():
        s, particles, _ = dambreak_solver(data=data, cell_order="linear")
        for _ in range(3):
            s.integrator.integrate(data[2]["dt"])
        fl = particles.arrays[0]
        return {k: fl.get(k) for k in ("x", "y", "z", "u", "v", "w", "rho", "p", "dt_fac")}

    a, b = _both(run)
    _assert_bit_equal(a, b, "dambreak")
    fl, wl, prm = data
    ref = cpu_ref.RefSimulation([cpu_ref.RefArray("fluid", 0, fl), cpu_ref.RefArray("wall", 1, wl)],
                                dim=3, integrator="pc", co=prm["c0"], ro=prm["rho0"],
                                gamma=7.0, alpha=0.1, beta=0.0, gravity=(0.0, 0.0, prm["g"]),
                                density_from=(0, 1), momentum_from=(0, 1), nthreads=8)
    for _ in range(3):
        ref.integrate(prm["dt"])
    worst = max(rel_err(a[k], ref.arrays[0][k]) for k in ("x", "y", "z", "u", "v", "w", "rho"))
    print(f"staged dam break (dx 0.05), 3 PC steps: bit-equal to the gather kernels, "
          f"max rel err vs oracle {worst:.3g}")
    assert worst <= 1e-10
