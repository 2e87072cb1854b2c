"""GPU parity of the operators beside the path (SPHDensityRate, XSPHCorrection)
and of the time-step reductions, against the FP64 oracle: bit-level agreement
expected (printed), 1e-6 relative required."""
import numpy as np
import pytest

from oracle import cpu_ref
from tests.helpers import bit_equal_frac, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-6


def _state(n, seed, varh):
    from pysph_amd.solver.generators import jittered_lattice_3d
    rng = np.random.default_rng(seed)
    d = jittered_lattice_3d(n=n, seed=seed, hdx=1.3)
    N = len(d["x"])
    d["u"], d["v"], d["w"] = rng.standard_normal((3, N))
    d["rho"] = 1000.0 * rng.uniform(0.98, 1.02, N)
    d["m"] = d["m"] * rng.uniform(0.9, 1.1, N)
    if varh:
        d["h"] = d["h"] * rng.uniform(0.8, 1.25, N)
    return d


@pytest.mark.parametrize("varh", [False, True])
def test_density_rate_and_xsph_vs_oracle(gpu, varh):
    from pysph_amd.base import api as base
    from pysph_amd.sph import api as sph
    d = _state(20, 7 if varh else 6, varh)
    pa = base.get_particle_array(name="fluid", **d)
    particles = base.Particles([pa])
    kernel = base.CubicSplineKernel(dim=3)
    rate = sph.SPHDensityRate.withargs().get_func(pa, pa)
    calc = sph.SPHCalc(particles, [pa], pa, kernel, [rate], ['_tmpx'], integrates=True)
    pa.get_carray('_tmpx').zero_()
    calc.sph('_tmpx')
    rate_gpu = pa.get('_tmpx')
    xs = sph.XSPHCorrection.withargs(eps=0.3).get_func(pa, pa)
    calc = sph.SPHCalc(particles, [pa], pa, kernel, [xs], ['_tmpx', '_tmpy', '_tmpz'],
                       integrates=True)
    for k in ('_tmpx', '_tmpy', '_tmpz'):
        pa.get_carray(k).zero_()
    calc.sph('_tmpx', '_tmpy', '_tmpz')
    full = {k: pa.get(k) for k in ("x", "y", "z", "h", "m", "rho", "u", "v", "w")}
    off, idx = cpu_ref.nnps(full, full, 2.0, particles.cell_manager.cell_size, True, nthreads=8)
    N = len(d["x"])
    ref_rate = np.zeros(N)
    cpu_ref.density_rate(3, full, full, off, idx, ref_rate, nthreads=8)
    ox, oy, oz = np.zeros(N), np.zeros(N), np.zeros(N)
    cpu_ref.xsph(3, full, full, off, idx, 0.3, ox, oy, oz, nthreads=8)
    errs = [rel_err(rate_gpu, ref_rate)] + [rel_err(pa.get(k), r) for k, r in
                                            (('_tmpx', ox), ('_tmpy', oy), ('_tmpz', oz))]
    print(f"varh {varh}: density rate / xsph rel err", ["%.3g" % e for e in errs],
          "bit-equal rate %.4f xsph-x %.4f" % (bit_equal_frac(rate_gpu, ref_rate),
                                               bit_equal_frac(pa.get('_tmpx'), ox)))
    assert max(errs) <= TOL


def test_drop_continuity_xsph_steps_vs_oracle(gpu):
    """2D drop with the continuity equation and XSPH, RK2, 8 steps."""
    from pysph_amd.solver.configs import drop_solver
    s, particles, d = drop_solver(dx=0.04, density="continuity", xsph_eps=0.5)
    ref = cpu_ref.RefSimulation([cpu_ref.RefArray("fluid", 0, d)], dim=2, integrator="rk2",
                                co=1400.0, ro=1.0, gamma=7.0, alpha=0.1, beta=0.0, nthreads=8,
                                density="continuity", xsph_eps=0.5)
    for _ in range(8):
        s.integrator.integrate(1e-5)
        ref.integrate(1e-5)
    worst = 0.0
    fl = particles.arrays[0]
    for k in ("x", "y", "u", "v", "rho", "p", "cs", "dt_fac"):
        e = rel_err(fl.get(k), ref.arrays[0][k])
        worst = max(worst, e)
        print(f"continuity+xsph {k}: rel err {e:.3g} bit-equal "
              f"{bit_equal_frac(fl.get(k), ref.arrays[0][k]):.4f}")
    assert worst <= TOL
    # the density moved: the continuity equation integrated it
    assert not np.array_equal(fl.get("rho"), d["rho"])


def test_time_step_functions(gpu):
    from pysph_amd.solver import api as solver
    from pysph_amd.solver.configs import drop_solver
    s, particles, d = drop_solver(dx=0.05)
    ts = solver.ViscousAndForceBasedTimeStep(cfl=0.3, co=1400.0, particles=particles)
    s.set_time_step_function(ts)
    s.set_final_time(1.0)
    s.set_time_step(1e-5)
    s.set_output(False)
    s.solve(nsteps=3)   # first step: the force bound uses dt as fmax (reference quirk)
    fl = particles.arrays[0]
    h, dtf = fl.get('h'), fl.get('dt_fac')
    want_visc = min(np.min(0.3 * h / (1400.0 + np.max(dtf))), 1)
    vt = solver.ViscousTimeStep(cfl=0.3, co=1400.0, particles=particles)
    assert vt.compute_time_step(s) == want_visc
    a = [fl.get(f'_a_{c}_1') for c in ('u', 'v')]
    fmax = np.max(np.sqrt(a[0] * a[0] + a[1] * a[1]))
    want = min(want_visc, 0.3 * np.min(np.sqrt(h / fmax)))
    got = ts.compute_time_step(s)
    print(f"viscous dt {want_visc:.6g}, force-based dt {got:.6g}")
    assert got == want


def test_continuity_xsph_single_precision_tracks_double(gpu):
    """cl_precision="single": the float operators (sph_density_rate_f32,
    sph_xsph_f32) over 5 RK2 steps stay within 1e-4 (max error / max
    magnitude) of the FP64 run."""
    from pysph_amd.solver.configs import drop_solver
    runs = {}
    for prec in ("double", "single"):
        s, particles, _ = drop_solver(dx=0.04, density="continuity", xsph_eps=0.5,
                                      precision=prec)
        for _ in range(5):
            s.integrator.integrate(1e-5)
        runs[prec] = particles.arrays[0]
    errs = {}
    for k in ("x", "y", "u", "v", "rho"):
        a, b = runs["single"].get(k), runs["double"].get(k)
        errs[k] = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    print("continuity+xsph FP32 vs FP64 after 5 steps:",
          " ".join(f"{k} {v:.2g}" for k, v in errs.items()))
    assert max(errs.values()) <= 1e-4
