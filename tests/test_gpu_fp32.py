"""The single-precision variant (float records, lists and operators, the
reference's cl_precision="single"): its lists may differ from the FP64 ones
only for pairs at d ~ 2h, and its densities and accelerations follow the
FP64 results to float accuracy.  Bars (stated, with the measured values
printed): differing pairs below 1e-4 of all pairs, and only pairs whose
distance is within 1e-5 relative of the support radius; densities within
1e-5 relative (tests.helpers.rel_err); batched drops within 1e-4 of the
largest magnitude of each quantity (componentwise relative errors of
quantities that pass through zero, like the velocities at the drop's axes,
measure float rounding of near-zero values, not the variant's accuracy)."""
import numpy as np
import pytest

from oracle import cpu_ref
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


def _pairs(off, idx):
    rows = np.repeat(np.arange(len(off) - 1), np.diff(off))
    return set(zip(rows.tolist(), np.asarray(idx).tolist()))


def test_lattice_nnps_density_single(gpu):
    from pysph_amd.base import api as base
    from pysph_amd.sph import api as sph
    from pysph_amd.solver.generators import jittered_lattice_3d
    d = jittered_lattice_3d(n=40, seed=21)
    pa = base.get_particle_array(name="fluid", cl_precision="single", **d)
    particles = base.Particles([pa])
    assert particles.cell_manager.precision == "single"
    kernel = base.CubicSplineKernel(dim=3)
    func = sph.SPHRho.withargs().get_func(pa, pa)
    calc = sph.SPHCalc(particles, [pa], pa, kernel, [func], ['rho'])
    calc.sph('rho')
    off32, idx32 = (t.cpu().numpy() for t in func.nbr_locator.get_csr())
    cs = particles.cell_manager.cell_size
    off, idx = cpu_ref.nnps(d, d, 2.0, cs, True, nthreads=8)
    rho = np.zeros(len(d["x"]))
    cpu_ref.summation_density(3, d, d, off, idx, rho, nthreads=8)
    a, b = _pairs(off32, idx32), _pairs(off, idx)
    diff = a ^ b
    r = 2.0 * d["h"][0]
    dist = [np.sqrt((d["x"][i] - d["x"][j]) ** 2 + (d["y"][i] - d["y"][j]) ** 2
                    + (d["z"][i] - d["z"][j]) ** 2) for i, j in diff]
    worst_edge = max([abs(x / r - 1.0) for x in dist], default=0.0)
    e = rel_err(pa.get('rho'), rho)
    print(f"FP32 lattice 40^3: {len(diff)} of {len(b)} pairs differ from FP64 "
          f"(all within {worst_edge:.2g} relative of 2h); density rel err {e:.3g}")
    assert len(diff) <= 1e-4 * len(b)
    assert worst_edge <= 1e-5
    assert e <= 1e-5


def test_batched_drops_single_vs_double(gpu):
    from pysph_amd.batch import BatchedSimulation, batched_drops
    sims, prm = batched_drops(nsims=4, seed=5, radius_range=(0.3, 0.6), dxs=(0.04, 0.05))
    b64 = BatchedSimulation(sims, dt_per_sim=prm["dt"])
    b32 = BatchedSimulation(sims, dt_per_sim=prm["dt"], precision="single")
    for _ in range(5):
        b64.step()
        b32.step()
    b64.check()
    b32.check()
    def norm_err(a, b):
        return float(np.abs(a - b).max() / np.abs(b).max())

    errs = {k: norm_err(b32.get(k), b64.get(k)) for k in ("x", "y", "u", "v", "rho", "p")}
    print("FP32 batch vs FP64 after 5 RK2 steps (max error / max magnitude): " +
          " ".join(f"{k} {v:.2g}" for k, v in errs.items()))
    assert max(errs.values()) <= 1e-4


def test_variable_h_lists_single(gpu):
    """Symmetric variable-h lists (VarHNbrParticleLocator) over float records:
    against the FP64 oracle's lists, differing pairs only where the distance
    is within float rounding of one of the two radii 2 h_a, 2 h_b, and SPHRho
    over them within 1e-5 of the FP64 density."""
    from pysph_amd.base import api as base
    from pysph_amd.sph import api as sph
    from pysph_amd.solver.generators import jittered_lattice_3d
    d = jittered_lattice_3d(n=24, seed=23, hdx=1.3)
    rng = np.random.default_rng(2)
    d["h"] = d["h"] * rng.uniform(0.8, 1.2, len(d["h"]))
    pa = base.get_particle_array(name="fluid", cl_precision="single", **d)
    particles = base.Particles([pa], variable_h=True)
    func = sph.SPHRho.withargs().get_func(pa, pa)
    calc = sph.SPHCalc(particles, [pa], pa, base.CubicSplineKernel(dim=3), [func], ['rho'])
    calc.sph('rho')
    off32, idx32 = (t.cpu().numpy() for t in func.nbr_locator.get_csr())
    cs = particles.cell_manager.cell_size
    off, idx = cpu_ref.nnps_varh(d, d, 2.0, cs)
    rho = np.zeros(len(d["x"]))
    cpu_ref.summation_density(3, d, d, off, idx, rho, nthreads=8)
    a, b = _pairs(off32, idx32), _pairs(off, idx)
    diff = a ^ b
    worst_edge = 0.0
    for i, j in diff:
        r = np.sqrt((d["x"][i] - d["x"][j]) ** 2 + (d["y"][i] - d["y"][j]) ** 2
                    + (d["z"][i] - d["z"][j]) ** 2)
        worst_edge = max(worst_edge, min(abs(r / (2 * d["h"][i]) - 1), abs(r / (2 * d["h"][j]) - 1)))
    e = rel_err(pa.get('rho'), rho)
    print(f"FP32 variable-h lists 24^3, h x U(0.8, 1.2): {len(diff)} of {len(b)} pairs differ "
          f"from FP64 (all within {worst_edge:.2g} relative of a radius); density rel err {e:.3g}")
    assert len(diff) <= 1e-4 * len(b)
    assert worst_edge <= 1e-5
    assert e <= 1e-5


@pytest.mark.timeout(600)
def test_lattice_100_single(gpu):
    """The FP32 variant at the configured size of the lattice workload (100^3,
    jitter 0.2 dx, h = 1.2 dx): lists against the FP64 oracle (differing pairs
    only within float rounding of 2h) and densities within 1e-5."""
    from pysph_amd.base import api as base
    from pysph_amd.sph import api as sph
    from pysph_amd.solver.generators import jittered_lattice_3d
    d = jittered_lattice_3d(n=100)
    pa = base.get_particle_array(name="fluid", cl_precision="single", **d)
    particles = base.Particles([pa])
    func = sph.SPHRho.withargs().get_func(pa, pa)
    calc = sph.SPHCalc(particles, [pa], pa, base.CubicSplineKernel(dim=3), [func], ['rho'])
    calc.sph('rho')
    off32, idx32 = (t.cpu().numpy() for t in func.nbr_locator.get_csr())
    cs = particles.cell_manager.cell_size
    off, idx = cpu_ref.nnps(d, d, 2.0, cs, True, nthreads=16)
    rho = np.zeros(len(d["x"]))
    cpu_ref.summation_density(3, d, d, off, idx, rho, nthreads=16)
    n = len(d["x"])

    def codes(o, ix):
        rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(o))
        return rows * n + ix.astype(np.int64)

    diff = np.setxor1d(codes(off32, idx32), codes(off, idx))
    i, j = diff // n, diff % n
    r = np.sqrt((d["x"][i] - d["x"][j]) ** 2 + (d["y"][i] - d["y"][j]) ** 2
                + (d["z"][i] - d["z"][j]) ** 2)
    worst_edge = float(np.max(np.abs(r / (2 * d["h"][i]) - 1))) if len(diff) else 0.0
    e = rel_err(pa.get('rho'), rho)
    print(f"FP32 lattice 100^3: {len(diff)} of {len(idx)} pairs differ from FP64 (all within "
          f"{worst_edge:.2g} relative of 2h); density rel err {e:.3g}")
    assert len(diff) <= 1e-4 * len(idx)
    assert worst_edge <= 1e-5
    assert e <= 1e-5
