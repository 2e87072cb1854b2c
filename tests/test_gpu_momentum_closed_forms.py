"""MomentumEquation on the GPU against its closed form (tests/closed_forms.py):
the reference's unit-square case (sph/tests/test_pressure_funcs.py:103-182,
alpha = beta = 1, cs = 0, so the beta mu^2 term acts) and two particles of
unequal h, m, rho, P and c on both branches of the spline, approaching and
receding.  Bar: 1e-13 relative (the closed form sums in another order)."""
import numpy as np
import pytest

from tests import closed_forms

pytestmark = pytest.mark.gpu


def _gpu_momentum(c, dim, alpha, beta, eta):
    from pysph_amd.base import api as base
    from pysph_amd.sph import api as sph
    d = closed_forms.as_arrays(c)
    pa = base.get_particle_array(name="fluid", **d)
    particles = base.Particles([pa])
    mom = sph.MomentumEquation.withargs(alpha=alpha, beta=beta, eta=eta).get_func(pa, pa)
    calc = sph.SPHCalc(particles, [pa], pa, base.CubicSplineKernel(dim=dim), [mom],
                       ['_tmpx', '_tmpy', '_tmpz'], integrates=True)
    for k in ('_tmpx', '_tmpy', '_tmpz'):
        pa.get_carray(k).zero_()
    calc.sph('_tmpx', '_tmpy', '_tmpz')
    return np.stack([pa.get('_tmpx'), pa.get('_tmpy'), pa.get('_tmpz')], 1), pa.get('dt_fac')


def test_unit_square_reference_case(gpu):
    c = closed_forms.unit_square()
    want, want_dtf = closed_forms.momentum(2, **c, alpha=1.0, beta=1.0, eta=0.1)
    got, dtf = _gpu_momentum(c, 2, 1.0, 1.0, 0.1)
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    print(f"unit square, alpha = beta = 1, cs = 0: max err / max |a| {err:.3g}")
    assert np.allclose(got, want, rtol=1e-13, atol=1e-15)
    assert np.allclose(dtf, want_dtf, rtol=1e-14, atol=0)


@pytest.mark.parametrize("approach,far,alpha,beta", [
    (True, False, 0.0, 1.0), (True, True, 0.0, 1.0), (True, False, 0.8, 0.5),
    (True, True, 0.8, 0.5), (False, False, 0.8, 0.5)])
def test_two_particles_unequal_states(gpu, approach, far, alpha, beta):
    c = closed_forms.two_particles(approach=approach, far=far)
    want, want_dtf = closed_forms.momentum(3, **c, alpha=alpha, beta=beta)
    got, dtf = _gpu_momentum(c, 3, alpha, beta, 0.1)
    rel = np.max(np.abs(got - want) / np.abs(want))
    print(f"two particles approach={approach} far={far} alpha={alpha} beta={beta}: "
          f"max rel err {rel:.3g}")
    assert np.allclose(got, want, rtol=1e-13, atol=0)
    assert np.allclose(dtf, want_dtf, rtol=1e-14, atol=0)
