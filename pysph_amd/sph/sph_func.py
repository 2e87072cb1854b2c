"""SPH function base classes (sph/sph_func.pyx:26-540).

An SPH function binds a source and a destination ParticleArray and is
evaluated per destination particle.  Here ``eval`` launches one HIP kernel
over all destinations (through libsphhip's C ABI); the per-particle loop of
the reference (``SPHFunction.eval`` at sph_func.pyx:208-227, ``eval_single``
at :427-441 / :513-531) runs as one GPU thread per destination.

``eval(kernel, out1, out2, out3, accumulate=True)`` takes device tensors.
With ``accumulate`` the result is added to the outputs (the reference's
``output += result``); without it the outputs become ``0.0 + result`` --
what the reference produces after ``SPHCalc.reset_output_arrays``.
"""
from .. import _lib


class Function:
    """A function class together with its parameters (sph_func.pyx:26-80)."""

    def __init__(self, sph_func, *args, **kwargs):
        self.sph_func = sph_func
        self.args = args
        self.kwargs = kwargs

    def get_func_class(self):
        return self.sph_func

    def get_func(self, source, dest):
        return self.sph_func(source, dest, *self.args, **self.kwargs)


class SPHFunction:
    """Function of destination properties only (sph_func.pyx:85-378)."""

    # which of the three outputs this function writes
    writes = (True, True, True)

    def __init__(self, source, dest=None, setup_arrays=True, setup_reads=True, *args, **kwargs):
        self.name = ""
        self.id = ""
        self.tag = ""
        self.source = source
        self.dest = source
        self.num_outputs = 3
        self.src_reads = []
        self.dst_reads = []
        self.kernel = None
        self.nbr_locator = None
        if setup_reads:
            self.set_src_dst_reads()
        if setup_arrays:
            self.setup_arrays()
        self.to_reset = []

    @classmethod
    def withargs(cls, *args, **kwargs):
        return Function(cls, *args, **kwargs)

    @classmethod
    def get_func(cls, source, dest):
        return Function(cls).get_func(source, dest)

    @classmethod
    def get_func_class(cls):
        return cls

    def set_src_dst_reads(self):
        pass

    def setup_arrays(self):
        pass

    def setup_iter_data(self):
        pass

    def setup_calc_updates(self, calc):
        return calc.updates

    # device helpers ---------------------------------------------------
    def _ctx(self):
        return _lib.context(self.dest.device)

    def d(self, prop):
        return self.dest.get_carray(prop)

    def s(self, prop):
        return self.source.get_carray(prop)

    def eval(self, kernel, output1, output2, output3, accumulate=True):
        raise NotImplementedError("SPHFunction.eval")


class SPHFunctionParticle(SPHFunction):
    """Function of a (source, destination) pair over neighbour lists
    (sph_func.pyx:383-500)."""

    def __init__(self, source, dest, setup_arrays=True, hks=False, exclude_self=False,
                 nbr_locator=None, *args, **kwargs):
        SPHFunction.__init__(self, source, setup_arrays=False, *args, **kwargs)
        self.dest = dest
        if hks:
            raise NotImplementedError("hks kernel symmetrisation is not part of this path")
        if exclude_self:
            # the reference reads an undefined attribute here (sph_func.pyx:435)
            raise NotImplementedError("exclude_self is not part of this path")
        self.exclude_self = exclude_self
        self.hks = hks
        self.bonnet_and_lok_correction = False
        self.rkpm_first_order_correction = False
        self.nbr_locator = nbr_locator
        if setup_arrays:
            self.setup_arrays()

    def csr(self):
        off, idx = self.nbr_locator.get_csr()
        return off, idx


class CSPHFunctionParticle(SPHFunctionParticle):
    """Numerator/denominator form (sph_func.pyx:504-540); the denominator
    stays zero for SPHRho, so no normalisation happens."""
