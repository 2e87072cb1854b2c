"""The SPH functions of the hot path, evaluated by HIP kernels.

SPHRho            sph/funcs/density_funcs.pyx:14-73
TaitEquation      sph/funcs/eos_funcs.pyx:62-129
MomentumEquation  sph/funcs/pressure_funcs.pyx:174-391 (hks=False; deltap optional)
GravityForce      sph/funcs/external_force.pyx:13-44
PositionStepping  sph/funcs/position_funcs.pyx:10-43
SPHDensityRate    sph/funcs/density_funcs.pyx:85-164 (hks=False)
XSPHCorrection    sph/funcs/xsph_funcs.pyx:9-91 (hks=False)
SPHPressureGradient          sph/funcs/pressure_funcs.pyx:19-168 (hks=False)
MonaghanArtificialViscosity  sph/funcs/viscosity_funcs.pyx:19-157 (hks=False)
MonaghanBoundaryForce        sph/funcs/boundary_funcs.pyx:9-107

The neighbour functions read the sorted shadows of their arrays (see
base/nnps.py) and the sorted-space lists of their locator; results land in
the caller's particle order and, for shadowed properties, in the shadow.
"""
import ctypes
import os

import torch

from .. import _lib, profiling
from ..base.nnps import PRHO2_FIELD, REC, REC_FIELDS, update_together
from .sph_func import CSPHFunctionParticle, SPHFunction, SPHFunctionParticle

_identity = {}

# LDS-staged density and momentum kernels for grids in the linear cell order
# (sph_*_staged); PYSPH_AMD_STAGED=0 keeps the gather kernels.  Both give the
# same results bit for bit.
STAGED = os.environ.get("PYSPH_AMD_STAGED", "1") != "0"


def _staged(cm, funcs):
    """Whether the staged kernels apply: FP64, the linear cell order, and
    FixedDest lists (the variable-h rows append entries out of stencil order)."""
    from ..base.nnps import VarHNbrParticleLocator
    return (STAGED and cm.order == "linear" and cm.precision == "double" and
            not any(isinstance(f.nbr_locator, VarHNbrParticleLocator) for f in funcs))


def _stage_sources(src, k, f, cm):
    """The staged kernels' extra per-source fields: cell starts and whether
    the source is the destination array."""
    src.s_cell_start[k] = _lib.ptr(cm.bins[f.source.name].cell_start)
    src.same[k] = 1 if f.source is f.dest else 0


def _identity_perm(n, device):
    """Sorted order == original order (functions used without a CellManager)."""
    key = (n, str(device))
    t = _identity.get(key)
    if t is None:
        t = torch.arange(max(n, 1), dtype=torch.int32, device=device)
        _identity[key] = t
    return t


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

    def eval(self, kernel, output1, output2, output3, accumulate=True, names=None):
        ctx = self._ctx()
        loc = self.nbr_locator
        cm = loc.cell_manager
        cnt, soff, idx = loc.get_sorted_lists()
        name = names[0] if names else None
        read = ('x', 'y', 'z', 'h') + ((name,) if name in REC_FIELDS and accumulate else ())
        drec = cm.shadows(self.dest, read)
        srec = cm.shadows(self.source, ('x', 'y', 'z', 'h', 'm'))
        out_s = cm.shadow_out(self.dest, name) if name else None
        p = _lib.ptr
        profiling.count_pairs(loc.pair_count())
        fn = (ctx.lib.sph_summation_density_f32 if cm.precision == "single"
              else ctx.lib.sph_summation_density)
        with profiling.region("summation_density"):
            ctx.check(fn(
                ctx.handle, kernel.dim, self.dest.get_number_of_particles(),
                self.dest.num_real_particles, p(cm.bins[self.dest.name].perm), p(cnt), p(soff),
                p(idx),
                p(drec), p(srec), int(cm.uniform_h()), int(bool(accumulate)), p(output1), out_s,
                REC))
        if name:
            cm.written(self.dest, name, out_s is not None)

    @staticmethod
    def fusable(funcs):
        return all(type(f) is SPHRho and f.dest is funcs[0].dest for f in funcs)

    @staticmethod
    def eval_multi(funcs, kernel, output1, output2, output3, accumulate=True, names=None,
                   eos=None):
        """All functions of one calc (one per source array) in one launch:
        the same sums, in the same order, as their eval calls in turn.  With
        eos = (TaitEquation function, its outputs, their names) the launch
        also evaluates that function on the densities it wrote (accumulate
        False for both)."""
        f0 = funcs[0]
        ctx = f0._ctx()
        cm = f0.nbr_locator.cell_manager
        p = _lib.ptr
        src = _lib.Sources()
        src.nsrc = len(funcs)
        update_together([f.nbr_locator for f in funcs])
        for k, f in enumerate(funcs):
            loc = f.nbr_locator
            cnt, soff, idx = loc.get_sorted_lists()
            srec = cm.shadows(f.source, ('x', 'y', 'z', 'h', 'm'))
            pm = cm.posmass(f.source)
            src.counts[k], src.slice_off[k], src.indices[k] = p(cnt), p(soff), p(idx)
            src.s_rec[k], src.ns[k] = p(srec), 0
            src.s_pm[k] = p(pm) if pm is not None else None
            profiling.count_pairs(loc.pair_count())
        name = names[0] if names else None
        read = ('x', 'y', 'z', 'h') + ((name,) if name in REC_FIELDS and accumulate else ())
        drec = cm.shadows(f0.dest, read)
        out_s = cm.shadow_out(f0.dest, name) if name else None
        if _staged(cm, funcs):
            for k, f in enumerate(funcs):
                _stage_sources(src, k, f, cm)
            with profiling.region("summation_density"):
                ctx.check(ctx.lib.sph_summation_density_staged(
                    ctx.handle, ctypes.byref(cm.grid), f0.nbr_locator.radius_scale, kernel.dim,
                    f0.dest.get_number_of_particles(), f0.dest.num_real_particles,
                    p(cm.bins[f0.dest.name].perm), ctypes.byref(src), p(drec),
                    int(cm.uniform_h()), int(bool(accumulate)), p(output1), out_s, REC))
            if name:
                cm.written(f0.dest, name, out_s is not None)
            return
        single = cm.precision == "single"
        if eos is not None:
            tf, touts, tnames = eos
            n0, n1 = tnames[0], tnames[1]
            p_s, cs_s = cm.shadow_out(f0.dest, n0), cm.shadow_out(f0.dest, n1)
            prho2_s = drec.data_ptr() + drec.element_size() * PRHO2_FIELD if n0 == 'p' else None
            fn = (ctx.lib.sph_summation_density_tait_f32 if single
                  else ctx.lib.sph_summation_density_tait)
            with profiling.region("summation_density"):
                ctx.check(fn(ctx.handle, kernel.dim, f0.dest.get_number_of_particles(),
                             f0.dest.num_real_particles, p(cm.bins[f0.dest.name].perm),
                             ctypes.byref(src), p(drec), int(cm.uniform_h()), 0, p(output1),
                             out_s, REC, tf.ro, tf.co, tf.gamma, 0, p(touts[0]), p(touts[1]),
                             p_s, cs_s, prho2_s))
            cm.written(f0.dest, name, True)
            cm.written(f0.dest, n0, p_s is not None)
            cm.written(f0.dest, n1, cs_s is not None)
            if prho2_s is not None:
                cm.prho2_written(f0.dest)
            return
        fn = (ctx.lib.sph_summation_density_multi_f32 if single
              else ctx.lib.sph_summation_density_multi)
        with profiling.region("summation_density"):
            ctx.check(fn(ctx.handle, kernel.dim, f0.dest.get_number_of_particles(),
                         f0.dest.num_real_particles, p(cm.bins[f0.dest.name].perm),
                         ctypes.byref(src), p(drec), int(cm.uniform_h()),
                         int(bool(accumulate)), p(output1), out_s, REC))
        if name:
            cm.written(f0.dest, name, out_s is not None)


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

    def eval(self, kernel, output1, output2, output3, accumulate=True, names=None):
        ctx = self._ctx()
        p = _lib.ptr
        n = self.dest.get_number_of_particles()
        cm = getattr(self, "cell_manager", None)
        n0 = names[0] if names else None
        n1 = names[1] if names else None
        p_s = cs_s = prho2_s = None
        fn = ctx.lib.sph_tait
        if cm is not None and self.dest.name in cm.array_indices:
            read = ('rho',) + tuple(k for k in (n0, n1) if accumulate and k in REC_FIELDS)
            rec = cm.shadows(self.dest, read)
            rho_s, stride = rec.data_ptr() + rec.element_size() * REC_FIELDS['rho'], REC
            perm = cm.bins[self.dest.name].perm
            p_s = cm.shadow_out(self.dest, n0) if n0 else None
            cs_s = cm.shadow_out(self.dest, n1) if n1 else None
            if n0 == 'p':
                # P/rho^2 of the new pressure, for MomentumEquation's sources
                prho2_s = rec.data_ptr() + rec.element_size() * PRHO2_FIELD
            if cm.precision == "single":
                fn = ctx.lib.sph_tait_f32
        else:
            cm = None
            rho_s, stride = p(self.d('rho')), 1
            perm = _identity_perm(n, self.dest.device)
        with profiling.region("tait"):
            ctx.check(fn(ctx.handle, n, self.dest.num_real_particles, p(perm),
                                       rho_s, stride, self.ro, self.co, self.gamma,
                                       int(bool(accumulate)), p(output1), p(output2), p_s, cs_s,
                                       prho2_s))
        if cm is not None:
            if n0:
                cm.written(self.dest, n0, p_s is not None)
            if n1:
                cm.written(self.dest, n1, cs_s is not None)
            if prho2_s is not None:
                cm.prho2_written(self.dest)
        else:
            self.dest.touch()


class MomentumEquation(SPHFunctionParticle):
    """Pressure gradient plus Monaghan artificial viscosity, with the
    viscous time-step factor dt_fac (pressure_funcs.pyx:246-391)."""

    writes = (True, True, True)
    writes_dt_fac = True
    DST = ('x', 'y', 'z', 'h', 'rho', 'p', 'u', 'v', 'w', 'cs')
    SRC = ('x', 'y', 'z', 'h', 'm', 'rho', 'p', 'u', 'v', 'w', 'cs')

    def __init__(self, source, dest, setup_arrays=True, alpha=1, beta=1, gamma=1.4, eta=0.1,
                 deltap=None, n=1, eps=0.2, **kwargs):
        SPHFunctionParticle.__init__(self, source, dest, setup_arrays, **kwargs)
        self.alpha = float(alpha)
        self.beta = float(beta)
        self.gamma = float(gamma)
        self.eta = float(eta)
        self.deltap = 0.0
        self.with_correction = False
        if deltap is not None:
            self.deltap = float(deltap)
            if deltap < 0:
                raise RuntimeError("deltap cannot be negative!")
            self.with_correction = True
        self.n = float(n)
        self.eps = float(eps)
        self.id = 'momentumequation'
        self.tag = "velocity"
        self.to_reset = ['dt_fac']

    def set_src_dst_reads(self):
        self.src_reads = list(self.SRC)
        self.dst_reads = list(self.DST) + ['tag']

    def setup_arrays(self):
        if self.dest.get_carray('dt_fac') is None:
            self.dest.add_property({'name': 'dt_fac'})
        if 'dt_fac' not in self.dst_reads:
            self.dst_reads.append('dt_fac')

    def pair_params(self):
        return _lib.PairParams(op=_lib.PAIR_MOMENTUM, with_correction=1, alpha=self.alpha,
                               beta=self.beta, eta=self.eta, deltap=self.deltap, n=self.n,
                               eps=self.eps)

    def eval(self, kernel, output1, output2, output3, accumulate=True, names=None):
        if self.with_correction:
            # the artificial-pressure branch (pressure_funcs.pyx:327-365)
            _pair_launch([self], kernel, output1, output2, output3, accumulate, names)
            return
        ctx = self._ctx()
        loc = self.nbr_locator
        cm = loc.cell_manager
        cnt, soff, idx = loc.get_sorted_lists()
        drec = cm.shadows(self.dest, self.DST)
        srec = cm.shadows(self.source, self.SRC)
        # the sources' P/rho^2 record field: computed by the library's
        # pre-pass unless Tait already wrote it for the current p and rho
        fresh = cm.prho2_fresh(self.source)
        p = _lib.ptr
        profiling.count_pairs(loc.pair_count())
        fn = ctx.lib.sph_momentum_av_f32 if cm.precision == "single" else ctx.lib.sph_momentum_av
        with profiling.region("momentum_av"):
            ctx.check(fn(
                ctx.handle, kernel.dim, self.dest.get_number_of_particles(),
                self.dest.num_real_particles, p(cm.bins[self.dest.name].perm), p(cnt), p(soff),
                p(idx),
                p(drec), p(srec), 0 if fresh else cm.bins[self.source.name].n,
                int(cm.uniform_h()), self.alpha,
                self.beta, self.eta,
                int(bool(accumulate)), p(output1), p(output2), p(output3), p(self.d('dt_fac'))))
        if not fresh:
            cm.prho2_written(self.source)
        for name in (names or ()):
            if name:
                cm.written(self.dest, name, False)
        cm.written(self.dest, 'dt_fac', False)

    @staticmethod
    def fusable(funcs):
        f0 = funcs[0]
        key = (f0.alpha, f0.beta, f0.eta, f0.with_correction, f0.deltap, f0.n, f0.eps)
        return all(type(f) is MomentumEquation and f.dest is f0.dest and
                   (f.alpha, f.beta, f.eta, f.with_correction, f.deltap, f.n, f.eps) == key
                   for f in funcs)

    @staticmethod
    def eval_multi(funcs, kernel, output1, output2, output3, accumulate=True, names=None):
        """All functions of one calc (one per source array) in one launch:
        the same sums, in the same order, as their eval calls in turn."""
        f0 = funcs[0]
        if f0.with_correction:
            _pair_launch(funcs, kernel, output1, output2, output3, accumulate, names)
            return
        ctx = f0._ctx()
        cm = f0.nbr_locator.cell_manager
        p = _lib.ptr
        src = _lib.Sources()
        src.nsrc = len(funcs)
        fresh = []
        update_together([f.nbr_locator for f in funcs])
        for k, f in enumerate(funcs):
            loc = f.nbr_locator
            cnt, soff, idx = loc.get_sorted_lists()
            srec = cm.shadows(f.source, f.SRC)
            fr = cm.prho2_fresh(f.source) or f.source in [g.source for g in funcs[:k]]
            fresh.append(fr)
            src.counts[k], src.slice_off[k], src.indices[k] = p(cnt), p(soff), p(idx)
            src.s_rec[k] = p(srec)
            src.ns[k] = 0 if fr else cm.bins[f.source.name].n
            profiling.count_pairs(loc.pair_count())
        drec = cm.shadows(f0.dest, f0.DST)
        if _staged(cm, funcs):
            for k, f in enumerate(funcs):
                _stage_sources(src, k, f, cm)
            with profiling.region("momentum_av"):
                ctx.check(ctx.lib.sph_momentum_av_staged(
                    ctx.handle, ctypes.byref(cm.grid), f0.nbr_locator.radius_scale, kernel.dim,
                    f0.dest.get_number_of_particles(), f0.dest.num_real_particles,
                    p(cm.bins[f0.dest.name].perm), ctypes.byref(src), p(drec),
                    int(cm.uniform_h()), f0.alpha, f0.beta, f0.eta, int(bool(accumulate)),
                    p(output1), p(output2), p(output3), p(f0.d('dt_fac'))))
        else:
            fn = (ctx.lib.sph_momentum_av_multi_f32 if cm.precision == "single"
                  else ctx.lib.sph_momentum_av_multi)
            with profiling.region("momentum_av"):
                ctx.check(fn(ctx.handle, kernel.dim, f0.dest.get_number_of_particles(),
                             f0.dest.num_real_particles, p(cm.bins[f0.dest.name].perm),
                             ctypes.byref(src), p(drec), int(cm.uniform_h()), f0.alpha,
                             f0.beta, f0.eta, int(bool(accumulate)), p(output1), p(output2),
                             p(output3), p(f0.d('dt_fac'))))
        for f, fr in zip(funcs, fresh):
            if not fr:
                cm.prho2_written(f.source)
        for name in (names or ()):
            if name:
                cm.written(f0.dest, name, False)
        cm.written(f0.dest, 'dt_fac', False)


def _pair_launch(funcs, kernel, output1, output2, output3, accumulate, names):
    """One sph_pair_accel launch for the functions of one calc (one per
    source array, all with the same parameters): per destination, each
    source's row summed in list order and added in source order."""
    f0 = funcs[0]
    ctx = f0._ctx()
    cm = f0.nbr_locator.cell_manager
    p = _lib.ptr
    prm = f0.pair_params()
    uses_prho2 = prm.op in (_lib.PAIR_PRESSURE_GRADIENT, _lib.PAIR_MOMENTUM)
    src = _lib.Sources()
    src.nsrc = len(funcs)
    frames = None
    if prm.op == _lib.PAIR_BOUNDARY_FORCE:
        frames = _lib.WallFrames()
    fresh = []
    update_together([f.nbr_locator for f in funcs])
    for k, f in enumerate(funcs):
        loc = f.nbr_locator
        cnt, soff, idx = loc.get_sorted_lists()
        srec = cm.shadows(f.source, f.SRC)
        fr = (not uses_prho2 or cm.prho2_fresh(f.source) or
              f.source in [g.source for g in funcs[:k]])
        fresh.append(fr)
        src.counts[k], src.slice_off[k], src.indices[k] = p(cnt), p(soff), p(idx)
        src.s_rec[k] = p(srec)
        src.ns[k] = 0 if fr else cm.bins[f.source.name].n
        if frames is not None:
            frames.s_perm[k] = p(cm.bins[f.source.name].perm)
            for c in ("tx", "ty", "tz", "nx", "ny", "nz"):
                t = f.source.get_carray(c)
                if t is None or t.dtype != torch.float64:
                    raise AttributeError(f"MonaghanBoundaryForce needs the double property "
                                         f"{c} on source array {f.source.name}")
                getattr(frames, c)[k] = p(t)
        profiling.count_pairs(loc.pair_count())
    drec = cm.shadows(f0.dest, f0.DST)
    fn = ctx.lib.sph_pair_accel_f32 if cm.precision == "single" else ctx.lib.sph_pair_accel
    dt_fac = f0.d('dt_fac') if getattr(f0, 'writes_dt_fac', False) else None
    with profiling.region(f0.id):
        ctx.check(fn(ctx.handle, kernel.dim, f0.dest.get_number_of_particles(),
                     f0.dest.num_real_particles, p(cm.bins[f0.dest.name].perm),
                     ctypes.byref(src), p(drec), int(cm.uniform_h()), ctypes.byref(prm),
                     ctypes.byref(frames) if frames is not None else None,
                     int(bool(accumulate)), p(output1), p(output2), p(output3), p(dt_fac)))
    if uses_prho2:
        for f, fr in zip(funcs, fresh):
            if not fr:
                cm.prho2_written(f.source)
    for name in (names or ()):
        if name:
            cm.written(f0.dest, name, False)
    if dt_fac is not None:
        cm.written(f0.dest, 'dt_fac', False)


class _PairFunction(SPHFunctionParticle):
    """Functions evaluated by sph_pair_accel; every function of a calc with
    the same parameters goes into one launch."""

    writes = (True, True, True)
    writes_dt_fac = False

    def set_src_dst_reads(self):
        self.src_reads = [k for k in self.SRC_READS]
        self.dst_reads = [k for k in self.DST_READS] + ['tag']

    def eval(self, kernel, output1, output2, output3, accumulate=True, names=None):
        _pair_launch([self], kernel, output1, output2, output3, accumulate, names)

    @classmethod
    def fusable(cls, funcs):
        f0 = funcs[0]
        key = f0._param_key()
        return all(type(f) is type(f0) and f.dest is f0.dest and f._param_key() == key
                   for f in funcs)

    @staticmethod
    def eval_multi(funcs, kernel, output1, output2, output3, accumulate=True, names=None):
        _pair_launch(funcs, kernel, output1, output2, output3, accumulate, names)

    def _param_key(self):
        prm = self.pair_params()
        return tuple(getattr(prm, f[0]) for f in prm._fields_)


class SPHPressureGradient(_PairFunction):
    """a_a += -m_b (P_a/rho_a^2 + P_b/rho_b^2) gradW_ab, with Monaghan's
    artificial pressure when deltap is given (pressure_funcs.pyx:19-151)."""

    DST = ('x', 'y', 'z', 'h', 'rho', 'p')
    SRC = ('x', 'y', 'z', 'h', 'm', 'rho', 'p')
    SRC_READS = ('x', 'y', 'z', 'h', 'm', 'rho', 'p')
    DST_READS = ('x', 'y', 'z', 'h', 'rho', 'p')

    def __init__(self, source, dest, setup_arrays=True, deltap=None, n=1, eps=0.2, **kwargs):
        SPHFunctionParticle.__init__(self, source, dest, setup_arrays, **kwargs)
        self.deltap = 0.0
        self.with_correction = False
        if deltap is not None:
            self.deltap = float(deltap)
            if deltap < 0:
                raise RuntimeError("deltap cannot be negative!")
            self.with_correction = True
        self.n = float(n)
        self.eps = float(eps)
        self.id = 'pgrad'
        self.tag = "velocity"

    def pair_params(self):
        return _lib.PairParams(op=_lib.PAIR_PRESSURE_GRADIENT,
                               with_correction=int(self.with_correction), deltap=self.deltap,
                               n=self.n, eps=self.eps)


class MonaghanArtificialViscosity(_PairFunction):
    """Monaghan's artificial viscosity alone, with the viscous time-step factor
    dt_fac (viscosity_funcs.pyx:19-151).  As in the reference, the constructor's
    eta is not stored: the function's eta is 0 unless set afterwards, and
    dt_fac uses the constant 0.01 h_ab^2."""

    writes_dt_fac = True
    DST = ('x', 'y', 'z', 'h', 'rho', 'u', 'v', 'w', 'cs')
    SRC = ('x', 'y', 'z', 'h', 'm', 'rho', 'u', 'v', 'w', 'cs')
    SRC_READS = ('x', 'y', 'z', 'h', 'm', 'rho', 'u', 'v', 'w', 'cs')
    DST_READS = ('x', 'y', 'z', 'h', 'p', 'u', 'v', 'w', 'cs', 'rho')

    def __init__(self, source, dest, setup_arrays=True, alpha=1, beta=1, gamma=1.4, eta=0.1,
                 **kwargs):
        SPHFunctionParticle.__init__(self, source, dest, setup_arrays, **kwargs)
        self.alpha = float(alpha)
        self.beta = float(beta)
        self.gamma = float(gamma)
        self.eta = 0.0
        self.id = 'momavisc'
        self.tag = "velocity"
        self.to_reset = ['dt_fac']

    def setup_arrays(self):
        if self.dest.get_carray('dt_fac') is None:
            self.dest.add_property({'name': 'dt_fac'})
        if 'dt_fac' not in self.dst_reads:
            self.dst_reads.append('dt_fac')

    def pair_params(self):
        return _lib.PairParams(op=_lib.PAIR_ARTIFICIAL_VISCOSITY, alpha=self.alpha,
                               beta=self.beta, eta=self.eta)


class MonaghanBoundaryForce(_PairFunction):
    """Monaghan's repulsive wall force from the wall particles' tangent and
    normal (tx, ty, tz, nx, ny, nz) (boundary_funcs.pyx:9-107)."""

    DST = ('x', 'y', 'z', 'h', 'm', 'cs')
    SRC = ('x', 'y', 'z', 'm')
    SRC_READS = ('x', 'y', 'z', 'h', 'm', 'rho', 'tx', 'ty', 'tz', 'nx', 'ny', 'nz', 'cs')
    DST_READS = ('x', 'y', 'z', 'h', 'm', 'cs')

    def __init__(self, source, dest, setup_arrays=False, delp=1.0, **kwargs):
        SPHFunctionParticle.__init__(self, source, dest, setup_arrays=True, **kwargs)
        self.id = 'monaghanbforce'
        self.tag = "velocity"
        self.delp = float(delp)

    def setup_arrays(self):
        missing = [c for c in ("tx", "ty", "tz", "nx", "ny", "nz")
                   if self.source.get_carray(c) is None]
        if missing:
            raise AttributeError(f"source array {self.source.name} has no property "
                                 f"{missing[0]}")

    def pair_params(self):
        return _lib.PairParams(op=_lib.PAIR_BOUNDARY_FORCE, delp=self.delp)


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

    def eval(self, kernel, output1, output2, output3, accumulate=True, names=None):
        ctx = self._ctx()
        p = _lib.ptr
        ctx.check(ctx.lib.sph_gravity(ctx.handle, self.dest.num_real_particles,
                                      self.dest.get_number_of_particles(), self.gx, self.gy,
                                      self.gz, int(bool(accumulate)), p(output1), p(output2),
                                      p(output3)))
        self.dest.touch(*[n for n in (names or ()) if n])


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

    def eval(self, kernel, output1, output2, output3, accumulate=True, names=None):
        ctx = self._ctx()
        p = _lib.ptr
        ctx.check(ctx.lib.sph_position_stepping(
            ctx.handle, self.dest.num_real_particles, self.dest.get_number_of_particles(),
            p(self.d('u')), p(self.d('v')), p(self.d('w')), p(output1), p(output2), p(output3)))
        self.dest.touch(*[n for n in (names or ()) if n])


class SPHDensityRate(SPHFunctionParticle):
    """Continuity equation: d(rho_a)/dt = sum_b m_b (v_a - v_b) . grad_a W_ab
    (density_funcs.pyx:116-164)."""

    writes = (True, False, False)
    DST = ('x', 'y', 'z', 'h', 'u', 'v', 'w')
    SRC = ('x', 'y', 'z', 'h', 'm', 'u', 'v', 'w')

    def __init__(self, source, dest, setup_arrays=True, **kwargs):
        if kwargs.get('hks'):
            raise NotImplementedError("hks is not part of this path")
        SPHFunctionParticle.__init__(self, source, dest, setup_arrays=True, **kwargs)
        self.id = 'densityrate'
        self.tag = "density"
        self.num_outputs = 1

    def set_src_dst_reads(self):
        self.src_reads = list(self.SRC)
        self.dst_reads = list(self.DST) + ['tag']

    def eval(self, kernel, output1, output2, output3, accumulate=True, names=None):
        ctx = self._ctx()
        loc = self.nbr_locator
        cm = loc.cell_manager
        cnt, soff, idx = loc.get_sorted_lists()
        drec = cm.shadows(self.dest, self.DST)
        srec = cm.shadows(self.source, self.SRC)
        p = _lib.ptr
        profiling.count_pairs(loc.pair_count())
        fn = (ctx.lib.sph_density_rate_f32 if cm.precision == "single"
              else ctx.lib.sph_density_rate)
        with profiling.region("density_rate"):
            ctx.check(fn(ctx.handle, kernel.dim, self.dest.get_number_of_particles(),
                         self.dest.num_real_particles, p(cm.bins[self.dest.name].perm), p(cnt),
                         p(soff), p(idx), p(drec), p(srec), int(cm.uniform_h()),
                         int(bool(accumulate)), p(output1)))
        for name in (names or ()):
            if name:
                cm.written(self.dest, name, False)


class XSPHCorrection(CSPHFunctionParticle):
    """XSPH velocity correction added to the position stepping:
    dx_a/dt += eps sum_b m_b W_ab / rho_ab (v_b - v_a) (xsph_funcs.pyx:36-91)."""

    writes = (True, True, True)
    DST = ('x', 'y', 'z', 'h', 'rho', 'u', 'v', 'w')
    SRC = ('x', 'y', 'z', 'h', 'm', 'rho', 'u', 'v', 'w')

    def __init__(self, source, dest, setup_arrays=True, eps=0.5, **kwargs):
        if kwargs.get('hks'):
            raise NotImplementedError("hks is not part of this path")
        CSPHFunctionParticle.__init__(self, source, dest, setup_arrays, **kwargs)
        self.eps = float(eps)
        self.id = 'xsph'
        self.tag = "position"

    def set_src_dst_reads(self):
        self.src_reads = list(self.SRC)
        self.dst_reads = list(self.DST)

    def eval(self, kernel, output1, output2, output3, accumulate=True, names=None):
        ctx = self._ctx()
        loc = self.nbr_locator
        cm = loc.cell_manager
        cnt, soff, idx = loc.get_sorted_lists()
        drec = cm.shadows(self.dest, self.DST)
        srec = cm.shadows(self.source, self.SRC)
        p = _lib.ptr
        profiling.count_pairs(loc.pair_count())
        fn = ctx.lib.sph_xsph_f32 if cm.precision == "single" else ctx.lib.sph_xsph
        with profiling.region("xsph"):
            ctx.check(fn(ctx.handle, kernel.dim, self.dest.get_number_of_particles(),
                         self.dest.num_real_particles, p(cm.bins[self.dest.name].perm), p(cnt),
                         p(soff), p(idx), p(drec), p(srec), int(cm.uniform_h()), self.eps,
                         int(bool(accumulate)), p(output1), p(output2), p(output3)))
        for name in (names or ()):
            if name:
                cm.written(self.dest, name, False)
