"""Operation lists of the benchmark workloads, built with the public API.

2D elliptical drop (RK2, no boundaries):
  SPHRho on/from Fluid -> TaitEquation(co, ro, gamma) on Fluid ->
  MomentumEquation(alpha, beta) on/from Fluid -> position stepping.
3D dam break (PredictorCorrector):
  SPHRho on Fluid from Fluid+Solid -> TaitEquation on Fluid ->
  MomentumEquation on Fluid from Fluid+Solid -> GravityForce(gz) ->
  position stepping; walls are a Solid array that is never stepped.
"""
from ..base import api as base
from ..sph import api as sph
from . import api as solver_api
from .generators import dam_break_3d, elliptical_drop


def drop_solver(dx=0.025, radius=1.0, vel=100.0, c0=1400.0, rho0=1.0, alpha=0.1, beta=0.0,
                dt=1e-5, integrator="rk2", data=None, cell_order="morton", density="summation",
                xsph_eps=None, precision="double"):
    """Solver + Particles of the 2D elliptical drop.  density="continuity"
    integrates rho with SPHDensityRate instead of SPHRho; xsph_eps adds the
    XSPH correction to the position stepping."""
    d = data if data is not None else elliptical_drop(dx=dx, radius=radius, rho0=rho0, c0=c0,
                                                      vel=vel)
    fluid = base.get_particle_array(name="fluid", type=base.Fluid, cl_precision=precision, **d)
    particles = base.Particles([fluid], cell_order=cell_order)
    integ = {"rk2": solver_api.RK2Integrator, "pc": solver_api.PredictorCorrectorIntegrator,
             "euler": solver_api.EulerIntegrator}[integrator]
    s = solver_api.Solver(dim=2, integrator_type=integ)
    F = base.Fluid
    if density == "continuity":
        s.add_operation(solver_api.SPHIntegration(sph.SPHDensityRate.withargs(), on_types=[F],
                                                  from_types=[F], updates=['rho'], id='density'))
    else:
        s.add_operation(solver_api.SPHOperation(sph.SPHRho.withargs(), on_types=[F],
                                                from_types=[F], updates=['rho'], id='density'))
    s.add_operation(solver_api.SPHOperation(
        sph.TaitEquation.withargs(co=c0, ro=rho0, gamma=7.0), on_types=[F], updates=['p', 'cs'],
        id='eos'))
    s.add_operation(solver_api.SPHIntegration(
        sph.MomentumEquation.withargs(alpha=alpha, beta=beta), on_types=[F], from_types=[F],
        updates=['u', 'v'], id='mom'))
    s.add_operation_step([F])
    if xsph_eps is not None:
        s.add_operation_xsph(eps=xsph_eps)
    s.set_time_step(dt)
    s.setup(particles)
    return s, particles, d


def dambreak_solver(dx=0.01, data=None, cell_order="morton"):
    """Solver + Particles of the 3D dam break (fluid + three wall layers)."""
    fl, wl, prm = data if data is not None else dam_break_3d(dx=dx)
    fluid = base.get_particle_array(name="fluid", type=base.Fluid, **fl)
    wall = base.get_particle_array(name="wall", type=base.Solid, **wl)
    particles = base.Particles([fluid, wall], cell_order=cell_order)
    s = solver_api.Solver(dim=3, integrator_type=solver_api.PredictorCorrectorIntegrator)
    F, S = base.Fluid, base.Solid
    s.add_operation(solver_api.SPHOperation(sph.SPHRho.withargs(), on_types=[F],
                                            from_types=[F, S], updates=['rho'], id='density'))
    s.add_operation(solver_api.SPHOperation(
        sph.TaitEquation.withargs(co=prm["c0"], ro=prm["rho0"], gamma=prm["gamma"]),
        on_types=[F], updates=['p', 'cs'], id='eos'))
    s.add_operation(solver_api.SPHIntegration(
        sph.MomentumEquation.withargs(alpha=prm["alpha"], beta=prm["beta"]), on_types=[F],
        from_types=[F, S], updates=['u', 'v', 'w'], id='mom'))
    s.add_operation(solver_api.SPHIntegration(
        sph.GravityForce.withargs(gz=prm["g"]), on_types=[F], updates=['u', 'v', 'w'],
        id='gravity'))
    s.add_operation_step([F])
    s.set_time_step(prm["dt"])
    s.setup(particles)
    return s, particles, (fl, wl, prm)


def dambreak_split_solver(dx=0.01, data=None, cell_order="morton", deltap=None, xsph_eps=0.5):
    """The operator split of examples/dam-break/dam_break3d.py on the 3D dam
    break: SPHDensityRate on the fluid from fluid and walls, TaitEquation on
    the fluid, SPHPressureGradient from the fluid (artificial pressure when
    deltap is given), MonaghanArtificialViscosity from fluid and walls,
    gravity, position stepping with XSPH; PredictorCorrector.  The example
    also evolves the walls' density and pressure; here the walls keep their
    initial state, as in the 3D dam break above."""
    fl, wl, prm = data if data is not None else dam_break_3d(dx=dx)
    fluid = base.get_particle_array(name="fluid", type=base.Fluid, **fl)
    wall = base.get_particle_array(name="wall", type=base.Solid, **wl)
    particles = base.Particles([fluid, wall], cell_order=cell_order)
    s = solver_api.Solver(dim=3, integrator_type=solver_api.PredictorCorrectorIntegrator)
    F, S = base.Fluid, base.Solid
    s.add_operation(solver_api.SPHIntegration(sph.SPHDensityRate.withargs(), on_types=[F],
                                              from_types=[F, S], updates=['rho'],
                                              id='density'))
    s.add_operation(solver_api.SPHOperation(
        sph.TaitEquation.withargs(co=prm["c0"], ro=prm["rho0"], gamma=prm["gamma"]),
        on_types=[F], updates=['p', 'cs'], id='eos'))
    s.add_operation(solver_api.SPHIntegration(
        sph.SPHPressureGradient.withargs(deltap=deltap, n=4), on_types=[F], from_types=[F],
        updates=['u', 'v', 'w'], id='pgrad'))
    s.add_operation(solver_api.SPHIntegration(
        sph.MonaghanArtificialViscosity.withargs(alpha=prm["alpha"], beta=prm["beta"]),
        on_types=[F], from_types=[F, S], updates=['u', 'v', 'w'], id='avisc'))
    s.add_operation(solver_api.SPHIntegration(
        sph.GravityForce.withargs(gz=prm["g"]), on_types=[F], updates=['u', 'v', 'w'],
        id='gravity'))
    s.add_operation_step([F])
    if xsph_eps is not None:
        s.add_operation_xsph(eps=xsph_eps)
    s.set_time_step(prm["dt"])
    s.set_output(False)
    s.setup(particles)
    return s, particles, (fl, wl, prm)
