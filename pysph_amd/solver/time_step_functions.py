"""Time-step policies (solver/time_step_functions.py).

The reductions (max dt_fac, min h, max |a|) run on the device
(``sph_extrema``); the rest is the reference's scalar arithmetic.  min over
particles of cfl*h/(c + M) equals cfl*min(h)/(c + M) exactly (rounding is
monotone), and likewise for sqrt(h/fmax).
"""
import ctypes
import math

from .. import _lib


def extrema(array, prop, comps=None):
    """(max |.|, min, max) of a device property (or of the norm of several)."""
    names = comps or (prop,)
    tensors = [array.get_carray(k) for k in names]
    if any(t is None for t in tensors):
        raise AttributeError(f"property missing in {array.name}: {names}")
    n = array.get_number_of_particles()
    ctx = _lib.context(array.device)
    out = (ctypes.c_double * 3)()
    ts = tensors + [None] * (3 - len(tensors))
    ctx.check(ctx.lib.sph_extrema(ctx.handle, n, len(tensors), _lib.ptr(ts[0]), _lib.ptr(ts[1]),
                                  _lib.ptr(ts[2]), out))
    return out[0], out[1], out[2]


class TimeStep:
    """Fixed time step: returns solver.dt (time_step_functions.py:3-6)."""

    def compute_time_step(self, solver):
        return solver.dt


class ViscousTimeStep(TimeStep):
    """dt = min over arrays with dt_fac of min(cfl*h/(co + max(dt_fac)))
    (time_step_functions.py:8-32); 1 when no array has dt_fac."""

    def __init__(self, cfl, co, particles):
        self.cfl = cfl
        self.co = co
        self.particles = particles

    def compute_time_step(self, solver):
        cfl, co = self.cfl, self.co
        dt = 1
        for array in self.particles.arrays:
            if array.get_carray('dt_fac') is not None:
                if array.get_number_of_particles() == 0:
                    raise ValueError("zero-size array in the time step reduction")
                _, _, max_fac = extrema(array, 'dt_fac')
                _, h_min, _ = extrema(array, 'h')
                _dt = cfl * h_min / (co + max_fac)
                dt = min(_dt, dt)
        return dt


class ViscousAndForceBasedTimeStep(ViscousTimeStep):
    """The viscous step, bounded by cfl*min(sqrt(h/fmax)) with fmax the
    largest stage-1 acceleration (time_step_functions.py:34-57)."""

    def compute_time_step(self, solver):
        dt = ViscousTimeStep.compute_time_step(self, solver)
        integrator = solver.integrator
        for array in self.particles.arrays:
            if array.get_carray('_a_u_1') is not None:
                fmax = integrator.get_max_acceleration(array, solver)
                _, h_min, _ = extrema(array, 'h')
                _dt = self.cfl * math.sqrt(h_min / fmax)
                dt = min(dt, _dt)
        return dt


class VelocityBasedTimeStep:
    """dt = cfl * min over arrays of min(h/(cs + |v|)) (time_step_functions.py:59-72)."""

    def __init__(self, particles, cfl=0.3):
        self.cfl = cfl
        self.particles = particles

    def compute_time_step(self, solver):
        raise NotImplementedError("VelocityBasedTimeStep is not part of this path")
