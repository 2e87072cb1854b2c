"""Solver: operation registry, setup and the time loop (solver/solver.py:33-775)."""
import logging

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

    def solve(self, show_progress=False, nsteps=None):
        """The time loop (solver.py:523-590); ``nsteps`` bounds the step count."""
        dt = self.dt
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
            steps += 1
        return steps
