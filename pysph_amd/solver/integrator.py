"""Integrators over SPH calcs (solver/integrator.py:11-683).

``setup_integrator`` creates the ``_<p>0`` initial and ``_a_<p>_<k>``
acceleration properties (integrator.py:180-260); ``eval`` runs the calcs
in solver order (:287-338); ``step`` computes ``x = x0 + a*dt`` for every
stepped property of an array in one HIP launch (:340-367).  Euler (:420-445),
RK2 midpoint (:450-498) and PredictorCorrector (:614-683) keep the
reference's stage sequence, including the neighbour-list rebuild after
every position change.
"""
import ctypes

from .. import _lib


def _mark_stepped(array, stepped):
    """Record the position versions this integrator's stage kernels left: the
    cell manager keeps its grid only for positions written this way (a small
    move per step; any other write of x, y or z rebuilds the grid)."""
    if any(p in ('x', 'y', 'z') for p in stepped):
        array._stepped_xyz = (array.version('x'), array.version('y'), array.version('z'))


class Integrator:

    def __init__(self, particles=None, calcs=()):
        self.particles = particles
        self.calcs = list(calcs)
        self.nsteps = 1
        self.cstep = 1
        self.time = 0.0
        self.local_time = 0.0
        self.rupdate_list = []
        self.step_props = {}
        self.initial_properties = {}
        self.velocity_accelerations = {}

    def set_rupdate_list(self):
        self.rupdate_list = [[] for _ in self.particles.arrays]

    def setup_integrator(self):
        self.arrays = self.particles.arrays
        for array in self.arrays:
            self.step_props[array.name] = {k + 1: {} for k in range(self.nsteps)}
            self.initial_properties[array.name] = {}
            self.velocity_accelerations[array.name] = {k + 1: {} for k in range(self.nsteps)}
        for calc in self.calcs:
            dest = calc.dest
            if not calc.integrates:
                continue
            for update_prop in calc.updates:
                initial_prop = '_' + update_prop + '0'
                dest.add_property({"name": initial_prop})
                self.initial_properties[dest.name][update_prop] = initial_prop
                for k in range(self.nsteps):
                    k_num = k + 1
                    step_prop = '_a_' + update_prop + '_' + str(k_num)
                    dest.add_property({"name": step_prop})
                    self.step_props[dest.name][k_num][update_prop] = [initial_prop, step_prop]
                    calc.dst_writes.setdefault(k_num, []).append(step_prop)
        self.set_rupdate_list()

    def _ctx(self):
        return _lib.context(self.arrays[0].device if self.arrays else None)

    def reset_accelerations(self, step):
        """Zero the acceleration arrays of stage `step`, one launch per array
        (integrator.py:287-308)."""
        ctx = self._ctx()
        for array in self.arrays:
            n = array.get_number_of_particles()
            names = [self.step_props[array.name][step][prop][1]
                     for prop in self.step_props[array.name][step]]
            for c in range(0, len(names), _lib.MAX_PROPS):
                pl = _lib.PropList()
                part = names[c:c + _lib.MAX_PROPS]
                pl.nprops = len(part)
                pl.dst_stride = 1
                for k, name in enumerate(part):
                    pl.dst[k] = _lib.ptr(array.get_carray(name))
                ctx.check(ctx.lib.sph_fill_many(ctx.handle, n, 0.0, ctypes.byref(pl)))
            if names:
                array.touch(*names)

    def save_initial_arrays(self):
        """x0 <- x for the stepped properties, one launch per array
        (integrator.py:262-285)."""
        ctx = self._ctx()
        for array in self.arrays:
            n = array.get_number_of_particles()
            pairs = list(self.initial_properties[array.name].items())
            for c in range(0, len(pairs), _lib.MAX_PROPS):
                pl = _lib.PropList()
                part = pairs[c:c + _lib.MAX_PROPS]
                pl.nprops = len(part)
                pl.dst_stride = 1
                for k, (src, dst) in enumerate(part):
                    pl.src[k] = _lib.ptr(array.get_carray(src))
                    pl.dst[k] = _lib.ptr(array.get_carray(dst))
                ctx.check(ctx.lib.sph_copy_many(ctx.handle, n, ctypes.byref(pl)))
            if pairs:
                array.touch(*[dst for _, dst in pairs])

    def eval(self):
        """Evaluate the calcs in order (solver/integrator.py:287-334).  A
        density calc followed by the equation-of-state calc of the same
        array runs as one launch (SPHCalc.fuses_eos) with the results of
        the two sph() calls."""
        k_num = self.cstep
        calcs = self.calcs
        c = 0
        while c < len(calcs):
            calc = calcs[c]
            calc.dest.set_time(self.local_time)
            if calc.integrates:
                calc.sph(*calc.dst_writes[k_num])
                c += 1
                continue
            nxt = calcs[c + 1] if c + 1 < len(calcs) else None
            if nxt is not None and calc.fuses_eos(nxt):
                # SPHRho then TaitEquation on the same array: one launch with
                # the results of the two calls in turn
                nxt.dest.set_time(self.local_time)
                calc.sph_with_eos(nxt)
                done = (calc, nxt)
            else:
                calc.sph(*calc.updates)
                done = (calc,)
            for d in done:
                self.particles.barrier()
                self.rupdate_list[d.dnum] = [d.updates]
                self.particles.update_remote_particle_properties(self.rupdate_list)
            c += len(done)
        self.particles.barrier()

    def _props_struct(self, array, triples):
        sp = _lib.StepProps()
        sp.nprops = len(triples)
        for k, (x, x0, a) in enumerate(triples):
            sp.x[k] = _lib.ptr(array.get_carray(x))
            sp.x0[k] = _lib.ptr(array.get_carray(x0))
            sp.a[k] = _lib.ptr(array.get_carray(a)) if a is not None else None
        return sp

    def step(self, dt):
        """x = x0 + a*dt for every stepped property (integrator.py:340-367)."""
        k_num = self.cstep
        ctx = self._ctx()
        for array in self.arrays:
            to_step = self.step_props[array.name][k_num]
            if not to_step:
                continue
            triples = [(prop, v[0], v[1]) for prop, v in to_step.items()]
            for c in range(0, len(triples), _lib.MAX_PROPS):
                sp = self._props_struct(array, triples[c:c + _lib.MAX_PROPS])
                ctx.check(ctx.lib.sph_stage_step(ctx.handle, array.get_number_of_particles(),
                                                 ctypes.byref(sp), float(dt)))
            array.touch(*to_step)
            for prop, v in to_step.items():
                if prop in ('u', 'v', 'w'):
                    self.velocity_accelerations[array.name][k_num][v[1]] = array.get_carray(v[1])
                if prop in ('x', 'y', 'z'):
                    array.set_dirty(True)
            _mark_stepped(array, to_step)
        self.cstep += 1

    def get_max_acceleration(self, array, solver):
        """Largest |a| of the stage-1 velocity accelerations (integrator.py:369-412;
        on the first step the reference returns solver.dt)."""
        from .time_step_functions import extrema
        if solver.count == 1:
            return solver.dt
        if array not in self.arrays:
            raise RuntimeError(f"Array {array.name} does not belong to me ")
        acc = float('-inf')
        if array.get_carray('_a_u_1') is not None:
            comps = tuple(f'_a_{c}_1' for c in ('u', 'v', 'w')[:solver.dim])
            mag, _, _ = extrema(array, None, comps)
            acc = max(acc, mag)
        return acc

    def integrate(self, dt):
        raise NotImplementedError


class EulerIntegrator(Integrator):
    """X(t+h) = X + h F(X) (integrator.py:420-445)."""

    def __init__(self, particles, calcs):
        Integrator.__init__(self, particles, calcs)
        self.nsteps = 1

    def integrate(self, dt):
        self.save_initial_arrays()
        self.reset_accelerations(step=1)
        self.local_time = self.time
        self.eval()
        self.step(dt)
        self.particles.update()
        self.cstep = 1


class RK2Integrator(Integrator):
    """Midpoint RK2 (integrator.py:450-498)."""

    def __init__(self, particles, calcs):
        Integrator.__init__(self, particles, calcs)
        self.nsteps = 1

    def integrate(self, dt):
        self.save_initial_arrays()
        self.reset_accelerations(step=1)
        self.local_time = self.time
        self.eval()
        self.step(dt / 2)
        self.particles.update()
        self.cstep = 1
        self.reset_accelerations(step=1)
        self.local_time = self.time + dt / 2
        self.eval()
        self.step(dt)
        self.particles.update()
        self.cstep = 1


class PredictorCorrectorIntegrator(Integrator):
    """Predict / correct / X = 2X - X0 (integrator.py:614-683)."""

    def __init__(self, particles, calcs):
        Integrator.__init__(self, particles, calcs)
        self.nsteps = 1

    def final_step(self):
        ctx = self._ctx()
        for array in self.arrays:
            to_step = self.step_props[array.name][1]
            if not to_step:
                continue
            triples = [(prop, v[0], None) for prop, v in to_step.items()]
            for c in range(0, len(triples), _lib.MAX_PROPS):
                sp = self._props_struct(array, triples[c:c + _lib.MAX_PROPS])
                ctx.check(ctx.lib.sph_pc_final(ctx.handle, array.get_number_of_particles(),
                                               ctypes.byref(sp)))
            array.touch(*to_step)
            if any(p in ('x', 'y', 'z') for p in to_step):
                array.set_dirty(True)
            _mark_stepped(array, to_step)

    def integrate(self, dt):
        self.save_initial_arrays()
        self.reset_accelerations(step=1)
        self.local_time = self.time
        self.eval()
        self.step(dt / 2)
        self.particles.update()
        self.cstep = 1
        self.reset_accelerations(step=1)
        self.local_time = self.time + dt / 2
        self.eval()
        self.step(dt / 2)
        self.particles.update()
        self.final_step()
        self.particles.update()
        self.cstep = 1


integration_methods = [('Euler', EulerIntegrator),
                       ('RK2', RK2Integrator),
                       ('PredictorCorrector', PredictorCorrectorIntegrator)]
