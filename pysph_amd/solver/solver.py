"""Solver: operation registry, setup and the time loop (solver/solver.py:33-775)."""
import logging
import os

from ..base.kernels import CubicSplineKernel
from ..sph.funcs import PositionStepping, XSPHCorrection
from .integrator import EulerIntegrator
from .sph_equation import SPHIntegration
from .time_step_functions import TimeStep

logger = logging.getLogger(__name__)


class Solver:

    def __init__(self, dim, integrator_type=EulerIntegrator):
        self.integrator_type = integrator_type
        self.dim = dim
        self.eps = -1
        self.initialize()

    def initialize(self):
        """solver.py:84-146"""
        self.particles = None
        self.default_kernel = CubicSplineKernel(dim=self.dim)
        self.operation_dict = {}
        self.order = []
        self.t = 0
        self.count = 0
        self.execute_commands = None
        self.pre_step_functions = []
        self.post_step_functions = []
        self.pfreq = 100
        self.kernel_correction = -1
        self.pid = None
        self.rank = 0
        self.in_parallel = False
        self.time_step_function = TimeStep()
        self.dt = 0.0
        self.tf = 0.0
        self.print_properties = ['x', 'u', 'm', 'h', 'p', 'e', 'rho']
        if self.dim > 1:
            self.print_properties.extend(['y', 'v'])
        if self.dim > 2:
            self.print_properties.extend(['z', 'w'])
        self.detailed_output = False
        self.fname = self.__class__.__name__
        self.output_directory = self.fname + '_output'
        # the reference always writes; False keeps solve() from writing files
        self.output_enabled = True
        self._writer = None

    def switch_integrator(self, integrator_type):
        if self.particles is None:
            raise RuntimeError("There are no particles!")
        self.integrator_type = integrator_type
        self.setup(self.particles)

    def add_operation_step(self, types, xsph=False, eps=0.5):
        """Position stepping for the given types (solver.py:164-184)."""
        updates = ['x', 'y', 'z'][:self.dim]
        self.add_operation(SPHIntegration(PositionStepping, on_types=types, updates=updates,
                                          id='step', kernel=None))

    def add_operation_xsph(self, eps, hks=False):
        """XSPH stepping on the arrays the position stepping moves, added to
        the same accelerations (solver.py:186-220)."""
        assert eps > 0, 'Invalid value for XSPH epsilon: %f' % (eps)
        self.eps = eps
        err = "position stepping function does not exist!"
        assert 'step' in self.operation_dict, err
        types = self.operation_dict['step'].on_types
        updates = self.operation_dict['step'].updates
        self.add_operation(SPHIntegration(
            XSPHCorrection.withargs(eps=eps, hks=hks), from_types=types, on_types=types,
            updates=updates, id='xsph', kernel=self.default_kernel))

    def add_operation(self, operation, before=False, id=None):
        """solver.py:222-282"""
        assert operation.id not in self.order, f'Operation {operation.id} exists!'
        assert operation.id not in self.operation_dict
        self.operation_dict[operation.id] = operation
        if id:
            msg = 'The specified operation doesnt exist'
            assert id in self.operation_dict, msg + ' in the calcs dict!'
            assert id in self.order, msg + ' in the order list!'
            if before:
                self.order.insert(self.order.index(id), operation.id)
            else:
                self.order.insert(self.order.index(id) + 1, operation.id)
        else:
            self.order.append(operation.id)

    def replace_operation(self, id, operation):
        assert id in self.operation_dict and id in self.order
        self.operation_dict.pop(id)
        self.operation_dict[operation.id] = operation
        idx = self.order.index(id)
        self.order.insert(idx, operation.id)
        self.order.remove(id)

    def remove_operation(self, id_or_operation):
        id = id_or_operation if isinstance(id_or_operation, str) else id_or_operation.id
        assert id in self.operation_dict and id in self.order, 'id doesnt exist!'
        self.order.remove(id)
        self.operation_dict.pop(id)

    def set_order(self, order):
        for equation_id in order:
            assert equation_id in self.order and equation_id in self.operation_dict
        self.order = list(order)

    def setup(self, particles=None):
        """Build the calcs and the integrator (solver.py:358-400)."""
        if particles:
            self.particles = particles
            particles.kernel = self.default_kernel
            self.integrator = self.integrator_type(particles, calcs=[])
            for equation_id in self.order:
                operation = self.operation_dict[equation_id]
                if operation.kernel is None:
                    operation.kernel = self.default_kernel
                calcs = operation.get_calcs(particles, operation.kernel)
                self.integrator.calcs.extend(calcs)
            self.integrator.setup_integrator()

    def set_final_time(self, tf):
        self.tf = tf

    def set_time_step(self, dt):
        self.dt = dt

    def set_print_freq(self, n):
        self.pfreq = n

    def set_time_step_function(self, time_step_function):
        self.time_step_function = time_step_function

    def set_output_fname(self, fname):
        """solver.py:450-452"""
        self.fname = fname

    def set_output_printing_level(self, detailed_output):
        """solver.py:454-456"""
        self.detailed_output = detailed_output

    def set_output_directory(self, path):
        """solver.py:458-460"""
        self.output_directory = path

    def set_output(self, enabled):
        """Switch the files solve() writes on or off (no reference counterpart)."""
        self.output_enabled = bool(enabled)

    def dump_output(self, dt, *print_properties):
        """Write ``<output_directory>/<fname>_<count>.npz`` in the version-1
        layout (solver.py:605-699): every array's properties not starting
        with '_' plus _tmpx/_tmpy/_tmpz and its constants (all properties with
        detailed output), and ``solver_data = {dt, t, count}``.

        The file is written by a worker thread from host copies taken on the
        operators' stream; ``wait_output`` returns once every file is
        complete, and solve() calls it before returning. ``print_properties``
        is accepted and unused, as in the reference."""
        from . import utils
        if self.particles is None:
            raise RuntimeError("There are no particles!")
        utils.mkdir(self.output_directory)
        arrays = {}
        for array in self.particles.arrays:
            props = {p: array.properties[p]
                     for p in array.output_property_names(all=self.detailed_output)}
            props.update(array.constants)
            arrays[array.name] = props
        path = os.path.join(self.output_directory, self.fname + '_' + str(self.count) + '.npz')
        if self._writer is None:
            self._writer = utils.OutputWriter()
        self._writer.submit(path, arrays, {"dt": dt, "t": self.t, "count": self.count})
        return path

    def wait_output(self):
        """Block until every file dump_output started is written."""
        if self._writer is not None:
            self._writer.wait()

    def load_output(self, count):
        """Resume from ``<fname>_<count>.npz`` (solver.py:701-743).

        ``count == '?'`` returns the available counts, ascending. Otherwise
        every particle array of the solver takes the properties and constants
        stored under its name (its particle type, precision and device stay),
        the particle manager is initialised again, ``t`` and ``count`` are
        restored (``dt`` is not, as in the reference), and an integrator that
        was already set up is rebuilt on the loaded arrays."""
        from . import utils
        self.wait_output()
        available = utils.output_counts(self.output_directory, self.fname)
        if count == '?':
            return available
        if str(count) not in available:
            msg = "File with iteration count `%s` does not exist" % (count)
            msg += "\nValid iteration counts are %s" % (available)
            raise IOError(msg)
        if self.particles is None:
            raise RuntimeError("There are no particles!")
        data = utils.read_output(os.path.join(self.output_directory,
                                              self.fname + '_' + str(count) + '.npz'))
        stored = data["arrays"]
        missing = [a.name for a in self.particles.arrays if a.name not in stored]
        if missing:
            raise KeyError(f"output file has no array named {missing[0]!r}")
        for array in self.particles.arrays:
            props = stored[array.name]
            array.initialize(**utils.property_infos(props))
            array.constants.update(utils.constants_of(props))
        self.particles.initialize()
        self.t = float(data["solver_data"]['t'])
        self.count = int(data["solver_data"]['count'])
        if getattr(self, 'integrator', None) is not None:
            self.setup(self.particles)
        return None

    def solve(self, show_progress=False, nsteps=None):
        """The time loop (solver.py:523-590); ``nsteps`` bounds the step count."""
        dt = self.dt
        if self.output_enabled:
            self.dump_output(dt, *self.print_properties)
        self.integrator.time = self.t
        steps = 0
        while self.t < self.tf if nsteps is None else steps < nsteps:
            self.t += dt
            self.count += 1
            self.particles.update()
            for func in self.pre_step_functions:
                func.eval(self)
            dt = self.time_step_function.compute_time_step(self)
            self.integrator.integrate(dt)
            self.integrator.time += dt
            for array in self.particles.arrays:
                array.set_time(self.t)
            for func in self.post_step_functions:
                func.eval(self)
            if self.output_enabled and self.pfreq and self.count % self.pfreq == 0:
                self.dump_output(dt, *self.print_properties)
            steps += 1
        self.wait_output()
        return steps
