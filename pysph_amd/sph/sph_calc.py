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

    def sph(self, output_array1=None, output_array2=None, output_array3=None,
            exclude_self=False):
        if exclude_self:
            raise NotImplementedError("exclude_self is not part of this path")
        outs, names = self._outputs((output_array1, output_array2, output_array3))
        ctx = _lib.context(self.dest.device)
        n = self.dest.get_number_of_particles()
        accumulate = True
        if not self.integrates and self.reset_arrays:
            # reset_output_arrays (sph_calc.pyx:431-440): the first function
            # writes 0.0 + result; outputs it does not write are zeroed
            first = self.funcs[0]
            for k in range(3):
                if not first.writes[k]:
                    ctx.check(ctx.lib.sph_fill(ctx.handle, n, 0.0, _lib.ptr(outs[k])))
                    self.dest.touch(names[k])
            accumulate = False
        for prop in self.to_reset:
            ctx.check(ctx.lib.sph_fill(ctx.handle, n, 0.0, _lib.ptr(self.dest.get_carray(prop))))
            self.dest.touch(prop)
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
