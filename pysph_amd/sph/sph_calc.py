"""SPHCalc: one destination array, its source arrays and one function per
source (sph/sph_calc.pyx:39-440).

``sph(out1, out2, out3)`` follows ``SPHCalc.sph`` (sph_calc.pyx:307-331):
non-integrating calcs start their outputs from zero, the ``to_reset``
properties are zeroed, then every function adds its contribution in
source order.  This class is the HIP implementation; it occupies the slot
of the reference's ``CLCalc`` (sph_calc.pyx:444-714) with no second
backend behind it.
"""
from .. import _lib


class SPHCalc:

    def __init__(self, particles, sources, dest, kernel, funcs, updates=(), integrates=False,
                 dnum=0, nbr_info=True, id="", kernel_gradient_correction=False,
                 kernel_correction=-1, dim=1, snum="", reset_arrays=True):
        if kernel_correction != -1:
            raise NotImplementedError("kernel corrections are not part of this path")
        self.particles = particles
        self.sources = list(sources)
        self.nsrcs = len(self.sources)
        self.dest = dest
        self.kernel = kernel
        self.funcs = list(funcs)
        self.updates = list(updates)
        self.integrates = integrates
        self.dnum = dnum
        self.nbr_info = nbr_info
        self.id = id
        self.dim = dim
        self.snum = snum
        self.reset_arrays = reset_arrays
        self.kernel_correction = kernel_correction
        self.tensor_eval = False
        self.dst_writes = {}
        self.initial_props = {}
        self.setup_internals()

    def setup_internals(self):
        """Bind a locator to every function (sph_calc.pyx:264-305)."""
        func = self.funcs[0]
        self.tag = func.tag
        self.to_reset = list(func.to_reset)
        self.src_reads = func.src_reads
        self.dst_reads = func.dst_reads
        self.nnps_manager = self.particles.nnps_manager
        for f in self.funcs:
            f.cell_manager = self.particles.cell_manager
        self.nbr_locators = []
        if self.nbr_info:
            for src, f in zip(self.sources, self.funcs):
                f.kernel = self.kernel
                loc = self.nnps_manager.get_neighbor_particle_locator(src, self.dest,
                                                                      self.kernel.radius())
                f.nbr_locator = loc
                self.nbr_locators.append(loc)
        else:
            for f in self.funcs:
                f.kernel = self.kernel

    def _outputs(self, names):
        defaults = ('_tmpx', '_tmpy', '_tmpz')
        outs, resolved = [], []
        for k in range(3):
            name = names[k] if k < len(names) and names[k] is not None else defaults[k]
            arr = self.dest.get_carray(name)
            if arr is None:
                raise AttributeError(f"property {name} not present in {self.dest.name}")
            outs.append(arr)
            resolved.append(name)
        return outs, tuple(resolved)

    def _zero(self, ctx, name, arr):
        """Zero one output of the destination, unless it still holds the zeros
        of an earlier call (same tensor, not written since): the temporary
        outputs a function leaves unwritten are zeroed at every evaluation."""
        dest = self.dest
        zeroed = dest.__dict__.setdefault("_zeroed", {})
        mark = (dest.version(name), arr.data_ptr(), arr.numel())
        if zeroed.get(name) == mark:
            return
        ctx.check(ctx.lib.sph_fill(ctx.handle, arr.numel(), 0.0, _lib.ptr(arr)))
        dest.touch(name)
        zeroed[name] = (dest.version(name), arr.data_ptr(), arr.numel())

    def fuses_eos(self, nxt):
        """Whether this calc (SPHRho over one or more sources) and `nxt`
        (TaitEquation on the same array, from the density just computed) can
        run as one launch (SPHRho.eval_multi with eos) with the results of
        their sph() calls in turn."""
        from .funcs import SPHRho, TaitEquation, _staged
        if self.integrates or nxt.integrates or not (self.reset_arrays and nxt.reset_arrays):
            return False
        if self.to_reset or nxt.to_reset or len(nxt.funcs) != 1:
            return False
        funcs = self.funcs
        if not funcs or len(funcs) > _lib.MAX_SRC or not SPHRho.fusable(funcs):
            return False
        tf = nxt.funcs[0]
        if type(tf) is not TaitEquation or nxt.dest is not self.dest or tf.dest is not self.dest:
            return False
        cm = getattr(funcs[0].nbr_locator, "cell_manager", None)
        if cm is None or self.dest.name not in cm.array_indices or _staged(cm, funcs):
            return False
        _, names = self._outputs(self.updates)
        _, nnames = nxt._outputs(nxt.updates)
        if names[0] != "rho" or tuple(nnames[:2]) != ("p", "cs"):
            return False
        # the zeroed outputs of the two calls must not be what either writes
        zeroed = {names[k] for k in range(3) if not funcs[0].writes[k]}
        zeroed |= {nnames[k] for k in range(3) if not tf.writes[k]}
        return not zeroed & {"rho", "p", "cs"}

    def sph_with_eos(self, nxt):
        """self.sph(*self.updates) then nxt.sph(*nxt.updates), fused (see
        fuses_eos)."""
        from .funcs import SPHRho
        outs, names = self._outputs(self.updates)
        nouts, nnames = nxt._outputs(nxt.updates)
        ctx = _lib.context(self.dest.device)
        for k in range(3):
            if not self.funcs[0].writes[k]:
                self._zero(ctx, names[k], outs[k])
        for k in range(3):
            if not nxt.funcs[0].writes[k]:
                nxt._zero(ctx, nnames[k], nouts[k])
        SPHRho.eval_multi(self.funcs, self.kernel, outs[0], outs[1], outs[2], accumulate=False,
                          names=names, eos=(nxt.funcs[0], nouts, nnames))
        if self.dest.is_dirty:
            self.particles.update()

    def sph(self, output_array1=None, output_array2=None, output_array3=None,
            exclude_self=False):
        if exclude_self:
            raise NotImplementedError("exclude_self is not part of this path")
        outs, names = self._outputs((output_array1, output_array2, output_array3))
        ctx = _lib.context(self.dest.device)
        accumulate = True
        if not self.integrates and self.reset_arrays:
            # reset_output_arrays (sph_calc.pyx:431-440): the first function
            # writes 0.0 + result; outputs it does not write are zeroed
            first = self.funcs[0]
            for k in range(3):
                if not first.writes[k]:
                    self._zero(ctx, names[k], outs[k])
            accumulate = False
        for prop in self.to_reset:
            self._zero(ctx, prop, self.dest.get_carray(prop))
        funcs = self.funcs
        multi = getattr(type(funcs[0]), "eval_multi", None) if funcs else None
        if multi is not None and 1 <= len(funcs) <= _lib.MAX_SRC and type(funcs[0]).fusable(funcs):
            # one function per source array: one launch with the same sums (a
            # single source too: the launch also reads the compact rows)
            multi(funcs, self.kernel, outs[0], outs[1], outs[2], accumulate=accumulate,
                  names=names)
        else:
            for f in funcs:
                f.eval(self.kernel, outs[0], outs[1], outs[2], accumulate=accumulate,
                       names=names)
                accumulate = True
        if self.dest.is_dirty:
            self.particles.update()

    def sph_array(self, output1, output2, output3, exclude_self=False):
        """Accumulate every function into the given output tensors."""
        for f in self.funcs:
            f.eval(self.kernel, output1, output2, output3, accumulate=True)
        self.dest.touch()


HIPCalc = SPHCalc
