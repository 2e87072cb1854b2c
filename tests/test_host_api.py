"""Host-side behaviour of the API mirror (CPU only: no kernel launches)."""
import numpy as np
import pytest

from pysph_amd.base import api as base
from pysph_amd.sph import api as sph
from pysph_amd.solver import api as solver


def test_get_particle_array_defaults():
    """particle_array.pyx:1431-1434 defaults and the always-present tag/group/local/pid."""
    x = np.linspace(0, 1, 5)
    pa = base.get_particle_array(name="f", x=x)
    assert pa.get_number_of_particles() == 5
    assert pa.num_real_particles == 5
    for prop, val in (("y", 0.0), ("m", 1.0), ("h", 1.0), ("rho", 1.0), ("cs", 0.0),
                      ("p", 0.0), ("_tmpx", 0.0)):
        assert np.all(pa.get(prop) == val), prop
    assert np.array_equal(pa.get("idx"), np.arange(5))
    assert np.all(pa.get("tag") == base.LocalReal)
    assert pa.particle_type == base.Fluid
    assert pa.get_type_name("tag") == "long" and pa.get_type_name("local") == "int"


def test_dirty_flag_and_set():
    pa = base.get_particle_array(name="f", x=np.zeros(3))
    pa.set_dirty(False)
    pa.set(rho=np.array([1.0, 2.0, 3.0]))
    assert not pa.is_dirty
    pa.set(y=np.ones(3))
    assert pa.is_dirty
    assert np.array_equal(pa.rho, [1.0, 2.0, 3.0])
    with pytest.raises(AttributeError):
        pa.set(nope=np.ones(3))


def test_remove_particles_swap_with_last():
    """carray remove: descending over sorted indices, swap with the last."""
    pa = base.get_particle_array(name="f", x=np.arange(6, dtype=float))
    pa.remove_particles(np.array([1, 4]))
    # remove 4: [0,1,2,3,5]; remove 1: [0,5,2,3]
    assert np.array_equal(pa.get("x"), [0.0, 5.0, 2.0, 3.0])


def test_circular_patch_counts_and_order():
    from pysph_amd.solver.generators import circular_patch_positions, elliptical_drop
    x, y = circular_patch_positions(0.025)
    assert len(x) == 5025
    assert len(circular_patch_positions(0.02)[0]) == 7860
    assert np.all(np.sqrt(x * x + y * y) - 1 <= 1e-10)
    d = elliptical_drop(dx=0.025)
    assert np.allclose(d["u"], -100 * d["x"]) and np.allclose(d["v"], 100 * d["y"])
    assert np.all(d["h"] == 1.3 * 0.025) and np.all(d["m"] == 0.025 ** 2)


def test_dam_break_counts():
    from pysph_amd.solver.generators import dam_break_3d
    fl, wl, prm = dam_break_3d(dx=0.02)
    assert len(fl["x"]) == 50 * 25 * 100
    # floor, end and side walls: full (nx+6)(ny+6)(nz+3) box minus the tank interior
    assert len(wl["x"]) == (200 + 6) * (25 + 6) * (125 + 3) - 200 * 25 * 125
    assert abs(prm["c0"] - 10 * np.sqrt(2 * 9.81 * 2.0)) < 1e-12
    assert prm["dt"] == 0.25 * prm["h"] / prm["c0"]
    # no wall particle inside the tank
    inside = (wl["x"] > 0) & (wl["x"] < 4.0) & (wl["y"] > 0) & (wl["y"] < 0.5) & (wl["z"] > 0)
    assert not np.any(inside)


def _drop_solver(integrator=solver.RK2Integrator):
    from pysph_amd.solver.generators import elliptical_drop
    pa = base.get_particle_array(name="fluid", type=base.Fluid, **elliptical_drop(dx=0.1))
    particles = base.Particles([pa])
    s = solver.Solver(dim=2, integrator_type=integrator)
    F = base.Fluid
    s.add_operation(solver.SPHOperation(sph.SPHRho.withargs(), on_types=[F], from_types=[F],
                                        updates=['rho'], id='density'))
    s.add_operation(solver.SPHOperation(sph.TaitEquation.withargs(co=1400.0, ro=1.0),
                                        on_types=[F], updates=['p', 'cs'], id='eos'))
    s.add_operation(solver.SPHIntegration(sph.MomentumEquation.withargs(alpha=0.1, beta=0.0),
                                          on_types=[F], from_types=[F], updates=['u', 'v'],
                                          id='mom'))
    s.add_operation_step([F])
    s.setup(particles)
    return s, pa


def test_solver_wiring_without_device():
    """Solver.setup builds calcs and integrator properties (solver.py:358-400,
    integrator.py:180-260) without touching a device."""
    s, pa = _drop_solver()
    ids = [c.id for c in s.integrator.calcs]
    assert ids == ['density_fluid', 'eos_fluid', 'mom_fluid', 'step_fluid']
    integ = s.integrator
    assert integ.initial_properties['fluid'] == {'u': '_u0', 'v': '_v0', 'x': '_x0', 'y': '_y0'}
    assert integ.step_props['fluid'][1]['x'] == ['_x0', '_a_x_1']
    for p in ('_x0', '_a_u_1', '_a_y_1', 'dt_fac'):
        assert pa.get_carray(p) is not None
    mom = integ.calcs[2]
    assert mom.integrates and mom.dst_writes[1] == ['_a_u_1', '_a_v_1']
    assert mom.to_reset == ['dt_fac']
    assert integ.calcs[0].nbr_locators[0].source is pa


def test_operation_order_and_errors():
    s = solver.Solver(dim=2)
    op = solver.SPHOperation(sph.TaitEquation.withargs(), on_types=[base.Fluid], updates=['p'],
                             id='eos')
    s.add_operation(op)
    s.add_operation(solver.SPHOperation(sph.SPHRho.withargs(), on_types=[base.Fluid],
                                        from_types=[base.Fluid], updates=['rho'], id='rho'),
                    before=True, id='eos')
    assert s.order == ['rho', 'eos']
    with pytest.raises(AssertionError):
        s.add_operation(op)
    with pytest.raises(ValueError):
        pa = base.get_particle_array(name="f", x=np.zeros(2))
        bad = solver.SPHOperation(sph.TaitEquation.withargs(), on_types=[base.Fluid],
                                  from_types=[base.Fluid], updates=['p'], id='x')
        bad.get_calcs(base.Particles([pa]), base.CubicSplineKernel(2))


def test_unsupported_features_fail_loudly():
    pa = base.get_particle_array(name="f", x=np.zeros(2))
    with pytest.raises(RuntimeError, match="deltap cannot be negative"):
        sph.MomentumEquation.withargs(deltap=-0.1).get_func(pa, pa)
    with pytest.raises(NotImplementedError):
        sph.SPHRho.withargs(hks=True).get_func(pa, pa)


def test_variable_h_selects_symmetric_locators():
    """NNPSManager(variable_h=True) hands out VarHNbrParticleLocator
    (base/nnps.pyx:895-910), the plain manager FixedDestNbrParticleLocator."""
    pa = base.get_particle_array(name="f", x=np.zeros(2))
    p = base.Particles([pa], variable_h=True)
    loc = p.nnps_manager.get_neighbor_particle_locator(pa, pa, 2.0)
    assert isinstance(loc, base.VarHNbrParticleLocator)
    assert p.nnps_manager.get_neighbor_particle_locator(pa, pa, 2.0) is loc
    p2 = base.Particles([base.get_particle_array(name="g", x=np.zeros(2))])
    assert type(p2.nnps_manager.get_neighbor_particle_locator(p2.arrays[0], p2.arrays[0],
                                                              2.0)) is \
        base.FixedDestNbrParticleLocator


def test_compute_without_device_raises():
    import torch
    if torch.cuda.is_available():
        pytest.skip("device present")
    s, pa = _drop_solver()
    with pytest.raises(RuntimeError, match="HIP device"):
        s.integrator.integrate(1e-5)


def test_kernel_interface():
    k = base.CubicSplineKernel(dim=3)
    assert k.radius() == 2.0 and k.get_type() == 1 and k.dim == 3
    assert abs(k.fac - 1.0 / np.pi) < 1e-16
    assert abs(base.CubicSplineKernel(dim=2).fac - 10 / (7 * np.pi)) < 1e-16
