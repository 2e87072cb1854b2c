"""SPHRho followed by TaitEquation on the same array runs as one launch
(sph_summation_density_tait, chosen by the integrator when the two calcs
follow each other).  Bar: the same bits as the two calls in turn -- every
state a run leaves (FP64 dam break with walls, FP32 drop), and rho, p, cs and
the record fields (p, cs, P/rho^2) of an array with non-LocalReal particles --
with the fused launch really taken."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("x", "y", "z", "u", "v", "w", "rho", "p", "cs")


def _bits(a):
    return np.ascontiguousarray(a).view(np.int64 if a.dtype == np.float64 else np.int32)


def _equal(a, b):
    return a.shape == b.shape and np.array_equal(_bits(a), _bits(b))


def _state(pa):
    return {k: pa.get(k) for k in KEYS if pa.has_array(k)}


def _count_fused(monkeypatch):
    from pysph_amd.sph import sph_calc
    calls = []
    orig = sph_calc.SPHCalc.sph_with_eos

    def counted(self, nxt):
        calls.append(1)
        return orig(self, nxt)

    monkeypatch.setattr(sph_calc.SPHCalc, "sph_with_eos", counted)
    return calls


def _separate(monkeypatch):
    from pysph_amd.sph import sph_calc
    monkeypatch.setattr(sph_calc.SPHCalc, "fuses_eos", lambda self, nxt: False)


def test_dambreak_fused_equals_separate(gpu, monkeypatch):
    """3D dam break at dx = 0.04 (density from the fluid and the walls), 10 PC
    steps: fused and separate runs leave the same bits."""
    from pysph_amd.solver.configs import dambreak_solver
    from pysph_amd.solver.generators import dam_break_3d
    data = dam_break_3d(dx=0.04)
    out = []
    for fused in (True, False):
        with monkeypatch.context() as m:
            if fused:
                calls = _count_fused(m)
            else:
                _separate(m)
            s, particles, (_, _, prm) = dambreak_solver(data=data)
            for _ in range(10):
                s.integrator.integrate(prm["dt"])
            out.append(_state(particles.arrays[0]))
            if fused:
                assert len(calls) == 20  # two evaluations per PC step
    print(f"\ndam break dx 0.04, 10 PC steps, fused density + Tait: {len(calls)} fused launches")
    for k in out[0]:
        assert _equal(out[0][k], out[1][k]), k


def test_drop_single_fused_equals_separate(gpu, monkeypatch):
    """The 2D drop in single precision (float records), 10 RK2 steps."""
    from pysph_amd.solver.configs import drop_solver
    out = []
    for fused in (True, False):
        with monkeypatch.context() as m:
            if fused:
                calls = _count_fused(m)
            else:
                _separate(m)
            s, particles, d = drop_solver(dx=0.05, precision="single")
            for _ in range(10):
                s.integrator.integrate(1e-5)
            out.append(_state(particles.arrays[0]))
            if fused:
                assert len(calls) == 20
    for k in out[0]:
        assert _equal(out[0][k], out[1][k]), k


def test_non_real_particles_fused_equals_separate(gpu):
    """An array whose last quarter is not LocalReal: rho, p, cs and the record
    fields p, cs, P/rho^2 of the fused launch equal the two calls'."""
    from pysph_amd.base import api as base
    from pysph_amd.sph import api as sph
    rng = np.random.default_rng(11)
    n = 4000
    x, y, z = rng.uniform(0, 1, size=(3, n))
    tag = np.zeros(n, dtype=np.int64)
    tag[3 * n // 4:] = 1
    d = dict(x=x, y=y, z=z, h=np.full(n, 0.06), m=rng.uniform(0.9, 1.1, n),
             rho=1000.0 * rng.uniform(0.99, 1.01, n), cs=np.full(n, 10.0))
    res = []
    for fused in (True, False):
        pa = base.get_particle_array(name="fluid", tag=tag, **d)
        particles = base.Particles([pa])
        kernel = base.CubicSplineKernel(dim=3)
        dcalc = sph.SPHCalc(particles, [pa], pa, kernel, [sph.SPHRho.withargs().get_func(pa, pa)],
                            ['rho'])
        tf = sph.TaitEquation.withargs(co=10.0, ro=1000.0, gamma=7.0).get_func(pa, pa)
        tcalc = sph.SPHCalc(particles, [pa], pa, kernel, [tf], ['p', 'cs'])
        assert dcalc.fuses_eos(tcalc)
        if fused:
            dcalc.sph_with_eos(tcalc)
        else:
            dcalc.sph('rho')
            tcalc.sph('p', 'cs')
        cm = particles.cell_manager
        rec = cm.bins[pa.name].rec[:n].cpu().numpy()
        res.append({k: pa.get(k) for k in ("rho", "p", "cs")} | {"rec": rec})
    for k in res[0]:
        if k == "rec":
            bad = np.nonzero(np.any(_bits(res[0][k]) != _bits(res[1][k]), axis=0))[0]
            assert bad.size == 0, f"record fields {bad.tolist()} differ"
        assert _equal(res[0][k], res[1][k]), k
    assert np.all(res[0]["p"][3 * n // 4:] == 0.0)
