"""The SPH functions of the hot path, evaluated by HIP kernels.

SPHRho            sph/funcs/density_funcs.pyx:14-73
TaitEquation      sph/funcs/eos_funcs.pyx:62-129
MomentumEquation  sph/funcs/pressure_funcs.pyx:174-391 (hks=False, deltap=None)
GravityForce      sph/funcs/external_force.pyx:13-44
PositionStepping  sph/funcs/position_funcs.pyx:10-43
"""
import ctypes

from .. import _lib
from .sph_func import CSPHFunctionParticle, SPHFunction, SPHFunctionParticle


class SPHRho(CSPHFunctionParticle):
    """Summation density rho_a = sum_b m_b W_ab((h_a+h_b)/2)."""

    writes = (True, False, False)

    def __init__(self, source, dest, setup_arrays=True, **kwargs):
        SPHFunctionParticle.__init__(self, source, dest, setup_arrays=True, **kwargs)
        self.id = 'sphrho'
        self.tag = "density"
        self.num_outputs = 1

    def set_src_dst_reads(self):
        self.src_reads = ['x', 'y', 'z', 'h', 'm']
        self.dst_reads = ['x', 'y', 'z', 'h', 'tag']

    def eval(self, kernel, output1, output2, output3, accumulate=True):
        ctx = self._ctx()
        off, idx = self.csr()
        nd_real = self.dest.num_real_particles
        nd = self.dest.get_number_of_particles()
        p = _lib.ptr
        ctx.check(ctx.lib.sph_summation_density(
            ctx.handle, kernel.dim, nd_real, nd, p(off), p(idx), p(self.d('x')), p(self.d('y')),
            p(self.d('z')), p(self.d('h')), p(self.s('x')), p(self.s('y')), p(self.s('z')),
            p(self.s('h')), p(self.s('m')), int(bool(accumulate)), p(output1)))


class TaitEquation(SPHFunction):
    """p = B((rho/ro)^gamma - 1), cs = co (rho/ro)^((gamma-1)/2), B = co^2 ro/gamma."""

    writes = (True, True, False)

    def __init__(self, source, dest, setup_arrays=True, co=1.0, ro=1000.0, gamma=7.0, **kwargs):
        SPHFunction.__init__(self, source, dest, setup_arrays, **kwargs)
        self.co = float(co)
        self.ro = float(ro)
        self.gamma = float(gamma)
        self.B = self.co * self.co * self.ro / self.gamma
        self.id = 'tait'
        self.tag = "state"
        self.num_outputs = 2

    def set_src_dst_reads(self):
        self.src_reads = []
        self.dst_reads = ['rho']

    def eval(self, kernel, output1, output2, output3, accumulate=True):
        ctx = self._ctx()
        p = _lib.ptr
        ctx.check(ctx.lib.sph_tait(ctx.handle, self.dest.num_real_particles,
                                   self.dest.get_number_of_particles(), p(self.d('rho')),
                                   self.ro, self.co, self.gamma, int(bool(accumulate)),
                                   p(output1), p(output2)))


class MomentumEquation(SPHFunctionParticle):
    """Pressure gradient plus Monaghan artificial viscosity, with the
    viscous time-step factor dt_fac (pressure_funcs.pyx:246-391)."""

    writes = (True, True, True)

    def __init__(self, source, dest, setup_arrays=True, alpha=1, beta=1, gamma=1.4, eta=0.1,
                 deltap=None, n=1, eps=0.2, **kwargs):
        SPHFunctionParticle.__init__(self, source, dest, setup_arrays, **kwargs)
        self.alpha = float(alpha)
        self.beta = float(beta)
        self.gamma = float(gamma)
        self.eta = float(eta)
        if deltap is not None:
            raise NotImplementedError("artificial pressure (deltap) is not part of this path")
        self.deltap = deltap
        self.with_correction = False
        self.n = n
        self.eps = eps
        self.id = 'momentumequation'
        self.tag = "velocity"
        self.to_reset = ['dt_fac']

    def set_src_dst_reads(self):
        self.src_reads = ['x', 'y', 'z', 'h', 'm', 'rho', 'p', 'u', 'v', 'w', 'cs']
        self.dst_reads = ['x', 'y', 'z', 'h', 'rho', 'p', 'u', 'v', 'w', 'cs', 'tag']

    def setup_arrays(self):
        if self.dest.get_carray('dt_fac') is None:
            self.dest.add_property({'name': 'dt_fac'})
        if 'dt_fac' not in self.dst_reads:
            self.dst_reads.append('dt_fac')

    def eval(self, kernel, output1, output2, output3, accumulate=True):
        ctx = self._ctx()
        off, idx = self.csr()
        p = _lib.ptr
        dst = _lib.MomDst(*[p(self.d(k)) for k in ('x', 'y', 'z', 'h', 'rho', 'p', 'u', 'v', 'w',
                                                  'cs')])
        src = _lib.MomSrc(*[p(self.s(k)) for k in ('x', 'y', 'z', 'h', 'm', 'rho', 'p', 'u', 'v',
                                                  'w', 'cs')])
        ctx.check(ctx.lib.sph_momentum_av(
            ctx.handle, kernel.dim, self.dest.num_real_particles,
            self.dest.get_number_of_particles(), p(off), p(idx), ctypes.byref(dst),
            ctypes.byref(src), self.alpha, self.beta, self.eta, int(bool(accumulate)),
            p(output1), p(output2), p(output3), p(self.d('dt_fac'))))


class GravityForce(SPHFunction):
    """Constant body force (external_force.pyx:38-44)."""

    writes = (True, True, True)

    def __init__(self, source, dest, setup_arrays=True, gx=0.0, gy=0.0, gz=0.0, hks=False):
        SPHFunction.__init__(self, source, dest, setup_arrays)
        self.id = 'gravityforce'
        self.tag = "velocity"
        self.gx = float(gx)
        self.gy = float(gy)
        self.gz = float(gz)

    def eval(self, kernel, output1, output2, output3, accumulate=True):
        ctx = self._ctx()
        p = _lib.ptr
        ctx.check(ctx.lib.sph_gravity(ctx.handle, self.dest.num_real_particles,
                                      self.dest.get_number_of_particles(), self.gx, self.gy,
                                      self.gz, int(bool(accumulate)), p(output1), p(output2),
                                      p(output3)))


class PositionStepping(SPHFunction):
    """dx/dt = v as an acceleration assignment (position_funcs.pyx:29-43)."""

    writes = (True, True, True)

    def __init__(self, source, dest, setup_arrays=True):
        SPHFunction.__init__(self, source, dest, setup_arrays)
        self.id = 'positionstepper'
        self.tag = "position"

    def set_src_dst_reads(self):
        self.dst_reads = ['u', 'v', 'w', 'tag']
        self.src_reads = []

    def eval(self, kernel, output1, output2, output3, accumulate=True):
        ctx = self._ctx()
        p = _lib.ptr
        ctx.check(ctx.lib.sph_position_stepping(
            ctx.handle, self.dest.num_real_particles, self.dest.get_number_of_particles(),
            p(self.d('u')), p(self.d('v')), p(self.d('w')), p(output1), p(output2), p(output3)))
