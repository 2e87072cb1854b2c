"""GPU parity of the HIP path against the FP64 oracle, through the public API
(and so through libsphhip's C ABI).

Bars: neighbour lists identical to the oracle (same rows in the same
reference order) and, as sets, to brute force; FP64 densities and
accelerations within 1e-6 relative (tests.helpers.rel_err) of the oracle --
the measured error is printed for every result.
"""
import numpy as np
import pytest
import torch

from oracle import cpu_ref
from tests.helpers import bit_equal_frac, rel_err, rows_as_sets_equal, sorted_rows

pytestmark = pytest.mark.gpu

TOL = 1e-6


def _api():
    from pysph_amd.base import api as base
    from pysph_amd.sph import api as sph
    from pysph_amd.solver import api as solver
    return base, sph, solver


# ---------------------------------------------------------------------------
# neighbour lists
# ---------------------------------------------------------------------------

def _dataset1():
    """base/tests/common_data.py:8-28 (generate_sample_dataset_1)."""
    a = dict(x=np.array([0.25, 0.8, 0.5, 0.8, 0.2, 0.5, 1.5, 1.5]),
             y=np.array([0.25, 0.1, 0.5, 0.8, 0.9, 1.5, 0.5, 1.5]),
             z=np.zeros(8), h=np.ones(8))
    b = dict(x=np.array([0.2, 1.2, 1.5, 0.4]), y=np.zeros(4),
             z=np.array([1.6, 1.5, -0.5, 0.4]), h=np.ones(4))
    return a, b


@pytest.mark.parametrize("order", ["morton", "linear"])
def test_nnps_reference_dataset(gpu, order):
    """base/tests/test_nnps.py:221-270: locator rows of parrs[1] over parrs[0]."""
    base, _, _ = _api()
    a, b = _dataset1()
    pa = base.get_particle_array(name="parr1", **a)
    pb = base.get_particle_array(name="parr2", **b)
    cm = base.CellManager([pa, pb], min_cell_size=1.0, max_cell_size=2.0, radius_scale=3.0,
                          order=order)
    nm = base.NNPSManager(cm)
    expect = {0.5: {3: [0]}, 1.0: {3: [0, 1, 2, 3]}, 4.0: {2: list(range(8))}}
    for rs, rows in expect.items():
        loc = nm.get_neighbor_particle_locator(pa, pb, rs)
        for i, want in rows.items():
            got = sorted(loc.get_nearest_particles(i).tolist())
            assert got == want, (rs, i, got)
    assert cm.cell_size == 2.0
    # whole rows vs oracle (reference order) for every radius scale
    for rs in (0.5, 1.0, 2.0, 4.0):
        loc = nm.get_neighbor_particle_locator(pa, pb, rs)
        off, idx = (t.cpu().numpy() for t in loc.get_csr())
        off_r, idx_r = cpu_ref.nnps(b, a, rs, 2.0, False)
        assert np.array_equal(off, off_r) and np.array_equal(idx, idx_r), rs


@pytest.mark.parametrize("n,seed,varh", [(2000, 1, False), (3000, 2, True), (1, 3, False)])
def test_nnps_random_vs_oracle_and_brute(gpu, n, seed, varh):
    base, _, _ = _api()
    rng = np.random.default_rng(seed)
    x, y, z = rng.uniform(-1, 1, size=(3, n))
    h0 = (8.0 / max(n, 1)) ** (1.0 / 3.0) * 0.6
    h = h0 * (rng.uniform(0.5, 1.0, n) if varh else np.ones(n))
    d = dict(x=x, y=y, z=z, h=h)
    pa = base.get_particle_array(name="fluid", **d)
    cm = base.CellManager([pa], radius_scale=2.0)
    loc = base.NNPSManager(cm).get_neighbor_particle_locator(pa, pa, 2.0)
    off, idx = (t.cpu().numpy() for t in loc.get_csr())
    off_r, idx_r = cpu_ref.nnps(d, d, 2.0, cm.cell_size, True)
    assert np.array_equal(off, off_r)
    assert np.array_equal(idx, idx_r)
    off_b, idx_b = cpu_ref.nnps_brute(d, d, 2.0)
    assert rows_as_sets_equal(off, idx, off_b, idx_b)
    # own index in the last slot of every row
    assert np.all(idx[off[1:] - 1] == np.arange(n))


def _lattice_density(n, order="morton", seed=20261202):
    base, sph, _ = _api()
    from pysph_amd.solver.generators import jittered_lattice_3d
    d = jittered_lattice_3d(n=n, seed=seed)
    pa = base.get_particle_array(name="fluid", **d)
    particles = base.Particles([pa], cell_order=order)
    kernel = base.CubicSplineKernel(dim=3)
    func = sph.SPHRho.withargs().get_func(pa, pa)
    calc = sph.SPHCalc(particles, [pa], pa, kernel, [func], ['rho'])
    calc.sph('rho')
    off, idx = (t.cpu().numpy() for t in func.nbr_locator.get_csr())
    return d, pa.get('rho'), off, idx, particles


@pytest.mark.parametrize("n,order", [(20, "morton"), (40, "morton"), (40, "linear")])
def test_lattice_nnps_density(gpu, n, order):
    d, rho, off, idx, particles = _lattice_density(n, order)
    cs = particles.cell_manager.cell_size
    assert cs == 2.0 * d["h"].max()
    off_r, idx_r = cpu_ref.nnps(d, d, 2.0, cs, True, nthreads=8)
    assert np.array_equal(off, off_r) and np.array_equal(idx, idx_r)
    rho_r = np.zeros(len(rho))
    cpu_ref.summation_density(3, d, d, off_r, idx_r, rho_r, nthreads=8)
    e = rel_err(rho, rho_r)
    print(f"n={n}^3 density rel err {e:.3g}, bit-equal {bit_equal_frac(rho, rho_r):.4f}")
    assert e <= TOL


def test_config1_full_size(gpu):
    """1M-particle jittered lattice: identical lists and densities at full size."""
    d, rho, off, idx, particles = _lattice_density(100)
    cs = particles.cell_manager.cell_size
    off_r, idx_r = cpu_ref.nnps(d, d, 2.0, cs, True, nthreads=16)
    assert np.array_equal(off, off_r)
    assert np.array_equal(idx, idx_r)
    rho_r = np.zeros(len(rho))
    cpu_ref.summation_density(3, d, d, off_r, idx_r, rho_r, nthreads=16)
    e = rel_err(rho, rho_r)
    print(f"1M lattice: nnz={len(idx)} mean row {len(idx)/len(rho):.2f} density rel err "
          f"{e:.3g} bit-equal {bit_equal_frac(rho, rho_r):.5f}")
    assert e <= TOL
    # size-independent invariants: symmetric lists for constant h, self in every row
    srt = sorted_rows(off, idx)
    assert np.all(idx[off[1:] - 1] == np.arange(len(rho)))
    row = np.repeat(np.arange(len(rho)), np.diff(off))
    a = np.sort(row.astype(np.int64) * len(rho) + srt)
    b = np.sort(srt.astype(np.int64) * len(rho) + row)
    assert np.array_equal(a, b), "neighbour relation not symmetric"


# ---------------------------------------------------------------------------
# operators
# ---------------------------------------------------------------------------

def test_tait_vs_oracle(gpu):
    base, sph, _ = _api()
    rng = np.random.default_rng(5)
    rho = 1000.0 * rng.uniform(0.9, 1.1, 10000)
    rho[:3] = [1000.0, 2000.0, 500.0]
    pa = base.get_particle_array(name="f", x=np.zeros(len(rho)), rho=rho)
    f = sph.TaitEquation.withargs(co=62.6418390534633, ro=1000.0, gamma=7.0).get_func(pa, pa)
    f.eval(base.CubicSplineKernel(3), pa.get_carray('p'), pa.get_carray('cs'),
           pa.get_carray('_tmpz'), accumulate=False)
    p_r, cs_r = np.zeros(len(rho)), np.zeros(len(rho))
    cpu_ref.tait(rho, 1000.0, 62.6418390534633, 7.0, p_r, cs_r)
    ep, ec = rel_err(pa.get('p'), p_r), rel_err(pa.get('cs'), cs_r)
    print(f"tait p rel err {ep:.3g} cs rel err {ec:.3g}")
    # both take the correctly rounded power x^7, x^3: bit for bit
    assert np.array_equal(pa.get('p'), p_r) and np.array_equal(pa.get('cs'), cs_r), (ep, ec)
    # closed forms (eos_funcs.pyx:118-129): rho=ro -> p=0, cs=co; rho=2ro -> p=127 B
    B = 62.6418390534633 ** 2 * 1000.0 / 7.0
    assert pa.get('p')[0] == 0.0 and pa.get('cs')[0] == 62.6418390534633
    assert abs(pa.get('p')[1] - 127 * B) <= 1e-12 * 127 * B


def _square_fixture(base):
    """sph/tests/function_test_template.py:45-58: unit square, u=[1,0,0,-1], p=[0,0,1,1]."""
    return dict(x=np.array([0.0, 0.0, 1.0, 1.0]), y=np.array([0.0, 1.0, 1.0, 0.0]),
                z=np.zeros(4), m=np.ones(4), u=np.array([1.0, 0.0, 0.0, -1.0]),
                p=np.array([0.0, 0.0, 1.0, 1.0]))


def test_operators_unit_square(gpu):
    """SPHRho and MomentumEquation(alpha=beta=1, eta=0.1) on the 4-particle
    fixture vs the oracle and vs the all-pairs sums of the reference tests
    (test_density_funcs.py:37-74, test_pressure_funcs.py:110-182)."""
    base, sph, _ = _api()
    d = _square_fixture(base)
    pa = base.get_particle_array(name="test", **d)
    particles = base.Particles([pa])
    kernel = base.CubicSplineKernel(dim=2)
    rho_f = sph.SPHRho.withargs().get_func(pa, pa)
    sph.SPHCalc(particles, [pa], pa, kernel, [rho_f], ['rho']).sph('_tmpx')
    full = {k: pa.get(k) for k in ("x", "y", "z", "h", "m", "rho", "p", "u", "v", "w", "cs")}
    off, idx = cpu_ref.nnps(full, full, 2.0, 2.0, True)
    ref_rho = np.zeros(4)
    cpu_ref.summation_density(2, full, full, off, idx, ref_rho)
    assert rel_err(pa.get('_tmpx'), ref_rho) <= 1e-14
    # all-pairs restatement (h=1: every pair within 2h)
    w, _ = cpu_ref.cubic_spline(2, [np.repeat(full[c], 4) for c in "xyz"],
                                [np.tile(full[c], 4) for c in "xyz"], np.ones(16))
    assert rel_err(pa.get('_tmpx'), w.reshape(4, 4).sum(axis=1)) <= 1e-14

    mom = sph.MomentumEquation.withargs(alpha=1.0, beta=1.0, gamma=1.4, eta=0.1).get_func(pa, pa)
    calc = sph.SPHCalc(particles, [pa], pa, kernel, [mom], ['_tmpx', '_tmpy', '_tmpz'],
                       integrates=True)
    for k in ('_tmpx', '_tmpy', '_tmpz'):
        pa.get_carray(k).zero_()
    calc.sph('_tmpx', '_tmpy', '_tmpz')
    ax, ay, az, dtf = np.zeros(4), np.zeros(4), np.zeros(4), np.zeros(4)
    cpu_ref.momentum(2, full, full, off, idx, 1.0, 1.0, 0.1, ax, ay, az, dtf)
    for got, want in ((pa.get('_tmpx'), ax), (pa.get('_tmpy'), ay), (pa.get('dt_fac'), dtf)):
        assert rel_err(got, want) <= 1e-14


def _random_state(d, rng, rho0=1000.0, c0=50.0, vel=1.0):
    n = len(d["x"])
    d = dict(d)
    d["u"], d["v"], d["w"] = vel * rng.standard_normal((3, n))
    d["rho"] = rho0 * rng.uniform(0.98, 1.02, n)
    d["p"] = rng.uniform(-1e3, 1e4, n)
    d["cs"] = c0 * rng.uniform(0.99, 1.01, n)
    return d


def test_momentum_lattice_vs_oracle(gpu):
    base, sph, _ = _api()
    from pysph_amd.solver.generators import jittered_lattice_3d
    rng = np.random.default_rng(11)
    d = _random_state(jittered_lattice_3d(n=24, seed=3, hdx=1.3), rng)
    pa = base.get_particle_array(name="fluid", **d)
    particles = base.Particles([pa])
    kernel = base.CubicSplineKernel(dim=3)
    mom = sph.MomentumEquation.withargs(alpha=0.1, beta=0.2).get_func(pa, pa)
    calc = sph.SPHCalc(particles, [pa], pa, kernel, [mom], ['_tmpx', '_tmpy', '_tmpz'],
                       integrates=True)
    for k in ('_tmpx', '_tmpy', '_tmpz'):
        pa.get_carray(k).zero_()
    calc.sph('_tmpx', '_tmpy', '_tmpz')
    full = {k: pa.get(k) for k in ("x", "y", "z", "h", "m", "rho", "p", "u", "v", "w", "cs")}
    off, idx = cpu_ref.nnps(full, full, 2.0, particles.cell_manager.cell_size, True, nthreads=8)
    ax, ay, az, dtf = (np.zeros(len(d["x"])) for _ in range(4))
    cpu_ref.momentum(3, full, full, off, idx, 0.1, 0.2, 0.1, ax, ay, az, dtf, nthreads=8)
    errs = [rel_err(pa.get(k), r) for k, r in (('_tmpx', ax), ('_tmpy', ay), ('_tmpz', az),
                                               ('dt_fac', dtf))]
    print("momentum rel err ax ay az dt_fac:", ["%.3g" % e for e in errs],
          "bit-equal ax %.4f" % bit_equal_frac(pa.get('_tmpx'), ax))
    assert max(errs) <= TOL


def test_density_momentum_variable_h_vs_oracle(gpu):
    """Per-particle h (the kernels read h_b and form 0.5*(h_a+h_b))."""
    base, sph, _ = _api()
    from pysph_amd.solver.generators import jittered_lattice_3d
    rng = np.random.default_rng(13)
    d = _random_state(jittered_lattice_3d(n=20, seed=5, hdx=1.3), rng)
    d["h"] = d["h"] * rng.uniform(0.8, 1.25, len(d["x"]))
    pa = base.get_particle_array(name="fluid", **d)
    particles = base.Particles([pa])
    assert not particles.cell_manager.uniform_h()
    kernel = base.CubicSplineKernel(dim=3)
    rho_f = sph.SPHRho.withargs().get_func(pa, pa)
    sph.SPHCalc(particles, [pa], pa, kernel, [rho_f], ['_tmpx']).sph('_tmpx')
    mom = sph.MomentumEquation.withargs(alpha=0.1, beta=0.2).get_func(pa, pa)
    calc = sph.SPHCalc(particles, [pa], pa, kernel, [mom], ['_tmpx', '_tmpy', '_tmpz'],
                       integrates=True)
    rho_gpu = pa.get('_tmpx')
    for k in ('_tmpx', '_tmpy', '_tmpz'):
        pa.get_carray(k).zero_()
    calc.sph('_tmpx', '_tmpy', '_tmpz')
    full = {k: pa.get(k) for k in ("x", "y", "z", "h", "m", "rho", "p", "u", "v", "w", "cs")}
    off, idx = cpu_ref.nnps(full, full, 2.0, particles.cell_manager.cell_size, True, nthreads=8)
    rho = np.zeros(len(d["x"]))
    cpu_ref.summation_density(3, full, full, off, idx, rho, nthreads=8)
    ax, ay, az, dtf = (np.zeros(len(d["x"])) for _ in range(4))
    cpu_ref.momentum(3, full, full, off, idx, 0.1, 0.2, 0.1, ax, ay, az, dtf, nthreads=8)
    errs = [rel_err(rho_gpu, rho)] + [
        rel_err(pa.get(k), r) for k, r in (('_tmpx', ax), ('_tmpy', ay), ('_tmpz', az),
                                           ('dt_fac', dtf))]
    print("variable h rel err rho ax ay az dt_fac:", ["%.3g" % e for e in errs],
          "bit-equal rho %.4f ax %.4f" % (bit_equal_frac(rho_gpu, rho),
                                          bit_equal_frac(pa.get('_tmpx'), ax)))
    assert max(errs) <= TOL


def test_momentum_conservation(gpu):
    """Pairwise antisymmetry: sum_a m_a a_a = 0 for constant h, no body force."""
    base, sph, _ = _api()
    from pysph_amd.solver.generators import jittered_lattice_3d
    rng = np.random.default_rng(12)
    d = _random_state(jittered_lattice_3d(n=16, seed=4, hdx=1.3), rng)
    d["m"] = d["m"] * rng.uniform(0.5, 1.5, len(d["x"]))
    pa = base.get_particle_array(name="fluid", **d)
    particles = base.Particles([pa])
    mom = sph.MomentumEquation.withargs(alpha=0.5, beta=0.5).get_func(pa, pa)
    calc = sph.SPHCalc(particles, [pa], pa, base.CubicSplineKernel(dim=3), [mom],
                       ['_tmpx', '_tmpy', '_tmpz'], integrates=True)
    for k in ('_tmpx', '_tmpy', '_tmpz'):
        pa.get_carray(k).zero_()
    calc.sph('_tmpx', '_tmpy', '_tmpz')
    m = pa.get('m')
    for k in ('_tmpx', '_tmpy', '_tmpz'):
        a = pa.get(k)
        tot = np.sum(m * a)
        assert abs(tot) <= 1e-9 * np.sum(np.abs(m * a)), (k, tot)


# ---------------------------------------------------------------------------
# whole steps
# ---------------------------------------------------------------------------

def _compare_arrays(arr_gpu, ref_arr, keys, label):
    worst = 0.0
    for k in keys:
        e = rel_err(arr_gpu.get(k), ref_arr[k])
        worst = max(worst, e)
        print(f"{label} {k}: rel err {e:.3g} bit-equal {bit_equal_frac(arr_gpu.get(k), ref_arr[k]):.4f}")
    return worst


@pytest.mark.parametrize("integrator,steps", [("rk2", 10), ("pc", 5), ("euler", 5)])
def test_drop_steps_vs_oracle(gpu, integrator, steps):
    from pysph_amd.solver.configs import drop_solver
    s, particles, d = drop_solver(dx=0.025, integrator=integrator)
    ref = cpu_ref.RefSimulation([cpu_ref.RefArray("fluid", 0, d)], dim=2, integrator=integrator,
                                co=1400.0, ro=1.0, gamma=7.0, alpha=0.1, beta=0.0, nthreads=8)
    for _ in range(steps):
        s.integrator.integrate(1e-5)
        ref.integrate(1e-5)
    worst = _compare_arrays(particles.arrays[0], ref.arrays[0],
                            ("x", "y", "u", "v", "rho", "p", "cs", "dt_fac"), integrator)
    assert worst <= TOL


def test_dambreak_small_pc_vs_oracle(gpu):
    from pysph_amd.solver.configs import dambreak_solver
    from pysph_amd.solver.generators import dam_break_3d
    data = dam_break_3d(dx=0.05)
    s, particles, (fl, wl, prm) = dambreak_solver(data=data)
    ref = cpu_ref.RefSimulation([cpu_ref.RefArray("fluid", 0, fl), cpu_ref.RefArray("wall", 1, wl)],
                                dim=3, integrator="pc", co=prm["c0"], ro=prm["rho0"],
                                gamma=7.0, alpha=0.1, beta=0.0, gravity=(0.0, 0.0, prm["g"]),
                                density_from=(0, 1), momentum_from=(0, 1), nthreads=8)
    for _ in range(3):
        s.integrator.integrate(prm["dt"])
        ref.integrate(prm["dt"])
    worst = _compare_arrays(particles.arrays[0], ref.arrays[0],
                            ("x", "y", "z", "u", "v", "w", "rho", "p"), "dambreak")
    assert worst <= TOL
    # walls never move
    assert np.array_equal(particles.arrays[1].get('x'), wl["x"])


def test_dambreak_full_size_pc_step_vs_oracle(gpu):
    """The 1M-fluid dam break (752,208 wall particles) at full size: both
    neighbour lists identical to the oracle's (order included) at the initial
    and at the stepped positions, and one PredictorCorrector step within 1e-6
    of the oracle's."""
    from pysph_amd.solver.configs import dambreak_solver
    from pysph_amd.solver.generators import dam_break_3d
    data = dam_break_3d(dx=0.01)
    s, particles, (fl, wl, prm) = dambreak_solver(data=data)
    fluid, wall = particles.arrays
    mom = [c for c in s.integrator.calcs if c.id.startswith("mom")][0]
    locs = [f.nbr_locator for f in mom.funcs]

    def check_lists(label):
        cur = {k: fluid.get(k) for k in ("x", "y", "z", "h")}
        for loc, src, same in ((locs[0], cur, True), (locs[1], wl, False)):
            off, idx = (t.cpu().numpy() for t in loc.get_csr())
            cs = particles.cell_manager.cell_size
            off_r, idx_r = cpu_ref.nnps(cur, src, 2.0, cs, same, nthreads=16)
            print(f"{label} {loc.source.name}: nnz {len(idx)}")
            assert np.array_equal(off, off_r) and np.array_equal(idx, idx_r), label

    check_lists("initial")
    ref = cpu_ref.RefSimulation([cpu_ref.RefArray("fluid", 0, fl), cpu_ref.RefArray("wall", 1, wl)],
                                dim=3, integrator="pc", co=prm["c0"], ro=prm["rho0"],
                                gamma=7.0, alpha=0.1, beta=0.0, gravity=(0.0, 0.0, prm["g"]),
                                density_from=(0, 1), momentum_from=(0, 1), nthreads=16)
    s.integrator.integrate(prm["dt"])
    ref.integrate(prm["dt"])
    worst = _compare_arrays(fluid, ref.arrays[0], ("x", "y", "z", "u", "v", "w", "rho", "p"),
                            "dambreak 1M")
    assert worst <= TOL
    check_lists("stepped")


def test_device_side_rebuild_matches_and_reports_capacity(gpu):
    """Builds after the first run without a host synchronisation: the lists
    equal the synchronous build's, and a build that runs out of list capacity
    (forced here with a tiny buffer) is reported at the next check instead of
    being used silently."""
    base, sph, _ = _api()
    from pysph_amd import _lib
    from pysph_amd.solver.generators import jittered_lattice_3d
    d = jittered_lattice_3d(n=16, seed=9, hdx=1.3)
    pa = base.get_particle_array(name="fluid", **d)
    particles = base.Particles([pa])
    loc = particles.nnps_manager.get_neighbor_particle_locator(pa, pa, 2.0)
    off0, idx0 = (t.cpu().numpy() for t in loc.get_csr())       # first build: synchronous
    x = pa.get("x")
    pa.set(x=x)                                                  # same positions, dirty
    particles.update()
    loc.update()
    assert loc.nbuilds == 2 and loc._pending                     # device-side build
    off1, idx1 = (t.cpu().numpy() for t in loc.get_csr())
    assert np.array_equal(off0, off1) and np.array_equal(idx0, idx1)
    loc.indices = loc.indices[:64]
    loc.slots = 0                                                # no growth ahead of need
    pa.set(x=x)
    particles.update()
    loc.update()
    with pytest.raises(_lib.SphError, match="incomplete"):
        _ = loc.nnz


def test_many_device_side_builds_in_flight(gpu):
    """Twelve device-side builds queued back to back, their totals coming back
    through a ring of four page-locked buffers: no build reports an error,
    and the entry count read after the last is the oracle's for the last
    positions, as are the exported lists' sizes."""
    base, _, _ = _api()
    from pysph_amd.solver.generators import jittered_lattice_3d
    d = jittered_lattice_3d(n=20, seed=4, hdx=1.3)
    pa = base.get_particle_array(name="fluid", **d)
    particles = base.Particles([pa])
    loc = particles.nnps_manager.get_neighbor_particle_locator(pa, pa, 2.0)
    loc.get_csr()                                   # first build: synchronous
    rng = np.random.default_rng(8)
    x0 = pa.get("x")
    want = []
    for k in range(12):
        x = x0 + rng.uniform(-0.3, 0.3, x0.size) * (d["h"][0] / 1.3)
        pa.set(x=x)
        particles.update()
        loc.update()                                # device-side: totals pending
        full = {kk: pa.get(kk) for kk in ("x", "y", "z", "h")}
        roff, _ = cpu_ref.nnps(full, full, 2.0, particles.cell_manager.cell_size, True,
                               nthreads=8)
        want.append(int(roff[-1]))
    assert len(set(want)) > 1                       # the builds differ
    assert loc.nbuilds == 13
    assert loc.nnz == want[-1]
    off, idx = (t.cpu().numpy() for t in loc.get_csr())
    assert len(idx) == want[-1] and off[-1] == want[-1]


@pytest.mark.parametrize("precision", ["double", "single"])
def test_cross_array_lists_with_near_cells_table(gpu, precision):
    """Fluid <- wall lists of the 3D dam break (dx = 0.04) over builds that
    find the wall binning unchanged: from the third build on, the count pass
    skips a fluid cell with no wall cell in reach by one flag per cell
    (sph_cell_near).  The lists of every such build equal the ones of the
    same build scanning every stencil, and in FP64 the oracle's."""
    base, _, _ = _api()
    from pysph_amd.solver.generators import dam_break_3d
    fl, wl, prm = dam_break_3d(dx=0.04)
    fa = base.get_particle_array(name="fluid", cl_precision=precision, **fl)
    wa = base.get_particle_array(name="wall", cl_precision=precision, **wl)
    particles = base.Particles([fa, wa])
    loc = particles.nnps_manager.get_neighbor_particle_locator(wa, fa, 2.0)
    rng = np.random.default_rng(5)
    x0 = {k: fa.get(k) for k in ("x", "y", "z")}
    used = 0
    for it in range(5):
        off, idx = (t.cpu().numpy() for t in loc.get_csr())
        with_table = loc.near_used
        used += with_table
        # the same build without the table
        near_table = loc._near_table
        loc._near_table = lambda cm, bs: (None, 0)
        loc.is_dirty = True
        off1, idx1 = (t.cpu().numpy() for t in loc.get_csr())
        del loc._near_table
        assert loc._near_table == near_table
        assert np.array_equal(off, off1) and np.array_equal(idx, idx1), it
        if precision == "double":
            d = {k: fa.get(k) for k in ("x", "y", "z", "h")}
            s = {k: wa.get(k) for k in ("x", "y", "z")}
            roff, ridx = cpu_ref.nnps(d, s, 2.0, particles.cell_manager.cell_size, False,
                                      nthreads=8)
            assert np.array_equal(off, roff) and np.array_equal(idx, ridx), it
        empty = int(np.sum(np.diff(off) == 0))
        print(f"build {it}: near table {'used' if with_table else 'not used'}, {len(idx)} "
              f"entries, {empty} of {len(off) - 1} fluid rows empty")
        # move the fluid by up to a third of a spacing (the walls fix the grid)
        fa.set(**{k: x0[k] + rng.uniform(-1.0, 1.0, x0[k].size) * prm["dx"] / 3
                  for k in ("x", "y", "z")})
    assert used >= 2
    # a larger fluid h written after the grid was built: stencils of reach 2,
    # so the table (made for reach 1) must not be handed over
    particles.update()
    loc.get_csr()
    fa.set(h=fa.get("h") * 1.5)
    loc.is_dirty = True
    off, idx = (t.cpu().numpy() for t in loc.get_csr())
    assert not loc.near_used
    if precision == "double":
        d = {k: fa.get(k) for k in ("x", "y", "z", "h")}
        s = {k: wa.get(k) for k in ("x", "y", "z")}
        roff, ridx = cpu_ref.nnps(d, s, 2.0, particles.cell_manager.cell_size, False,
                                  nthreads=8)
        assert np.array_equal(off, roff) and np.array_equal(idx, ridx)
    print(f"fluid h x 1.5 after the grid build: {len(idx)} entries, no near table")


@pytest.mark.parametrize("precision", ["double", "single"])
def test_paired_device_side_builds(gpu, precision):
    """Fluid <- fluid and fluid <- wall lists of the 3D dam break (dx = 0.04)
    built together (nnps.update_together: one count launch and one fill
    launch for both, sph_nnps_build_dev_pair) equal the same builds made one
    at a time, over moves of the fluid, with and without the wall's
    near-cells table; in FP64 also the oracle's.  An empty wall leaves the
    fluid's own build to its single launches."""
    base, _, _ = _api()
    from pysph_amd.base.nnps import update_together
    from pysph_amd.solver.generators import dam_break_3d
    fl, wl, prm = dam_break_3d(dx=0.04)
    fa = base.get_particle_array(name="fluid", cl_precision=precision, **fl)
    wa = base.get_particle_array(name="wall", cl_precision=precision, **wl)
    particles = base.Particles([fa, wa])
    mgr = particles.nnps_manager
    locs = [mgr.get_neighbor_particle_locator(fa, fa, 2.0),
            mgr.get_neighbor_particle_locator(wa, fa, 2.0)]
    ctx = particles.cell_manager._ctx()
    paired = []
    real = ctx.lib.sph_nnps_build_dev_pair

    def counted(*args):
        paired.append(1)
        return real(*args)

    rng = np.random.default_rng(11)
    x0 = {k: fa.get(k) for k in ("x", "y", "z")}
    ctx.lib.sph_nnps_build_dev_pair = counted
    try:
        for it in range(5):
            before = len(paired)
            update_together(locs)
            got = [tuple(t.cpu().numpy() for t in loc.get_csr()) for loc in locs]
            if it > 0:  # the first builds run with a host synchronisation
                assert len(paired) == before + 1, it
            for loc in locs:  # the same builds one at a time
                loc.is_dirty = True
                loc.update()
            assert len(paired) == before + (it > 0)
            for loc, (off, idx) in zip(locs, got):
                off1, idx1 = (t.cpu().numpy() for t in loc.get_csr())
                assert np.array_equal(off, off1) and np.array_equal(idx, idx1), it
            if precision == "double":
                d = {k: fa.get(k) for k in ("x", "y", "z", "h")}
                for src, (off, idx) in zip((fa, wa), got):
                    s = {k: src.get(k) for k in ("x", "y", "z")}
                    roff, ridx = cpu_ref.nnps(d, s, 2.0, particles.cell_manager.cell_size,
                                              src is fa, nthreads=8)
                    assert np.array_equal(off, roff) and np.array_equal(idx, ridx), it
            print(f"build {it}: {len(got[0][1])} + {len(got[1][1])} entries, near table "
                  f"{'used' if locs[1].near_used else 'not used'}")
            fa.set(**{k: x0[k] + rng.uniform(-1.0, 1.0, x0[k].size) * prm["dx"] / 3
                      for k in ("x", "y", "z")})
        assert locs[1].near_used
        # an empty wall: the fluid's own build alone
        wa.remove_particles(np.arange(wa.get_number_of_particles()))
        particles.update()
        n0 = len(paired)
        update_together(locs)
        off, idx = (t.cpu().numpy() for t in locs[1].get_csr())
        assert len(idx) == 0 and np.all(off == 0)
        off, idx = (t.cpu().numpy() for t in locs[0].get_csr())
        locs[0].is_dirty = True
        off1, idx1 = (t.cpu().numpy() for t in locs[0].get_csr())
        assert np.array_equal(off, off1) and np.array_equal(idx, idx1)
        print(f"empty wall: paired entry called {len(paired) - n0} time(s)")
    finally:
        ctx.lib.sph_nnps_build_dev_pair = real


def test_lists_exact_at_the_support_radius(gpu):
    """Pairs exactly at, and within a few FP64 ulps of, the support radius,
    far from the origin: the float distance test of the list kernel leaves
    them to its FP64 fallback, so the lists equal the oracle's."""
    base, _, _ = _api()
    rng = np.random.default_rng(31)
    n = 12
    dx = 0.25
    g = np.arange(n) * dx
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    x, y, z = (1000.0 + X.ravel(), -2000.0 + Y.ravel(), 500.0 + Z.ravel())
    # nudge a third of the particles by a few ulps (both signs)
    k = rng.integers(-4, 5, x.size) * (rng.random(x.size) < 0.33)
    x = x + k * np.spacing(1000.0)
    d = {"x": x, "y": y, "z": z, "h": np.full(x.size, dx), "m": np.ones(x.size),
         "rho": np.ones(x.size)}
    pa = base.get_particle_array(name="fluid", **d)
    particles = base.Particles([pa])
    loc = particles.nnps_manager.get_neighbor_particle_locator(pa, pa, 2.0)
    off, idx = (t.cpu().numpy() for t in loc.get_csr())
    full = {kk: pa.get(kk) for kk in ("x", "y", "z", "h")}
    roff, ridx = cpu_ref.nnps(full, full, 2.0, particles.cell_manager.cell_size, True)
    assert np.array_equal(off, roff) and np.array_equal(idx, ridx)
    # the radius is hit exactly by lattice neighbours two spacings away
    assert (len(idx) - x.size) > 0
    print(f"support-radius lattice far from the origin: {len(idx)} entries, identical to the oracle")
