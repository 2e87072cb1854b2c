"""Smoothing kernels (pysph.base.kernels): the cubic spline of this path.

``CubicSplineKernel`` keeps the reference's interface (base/kernels.pyx:
``dim``, ``radius()`` = 2 at :449-450, ``get_type()`` = CUBIC_SPLINE at :452,
``_fac``/``fac`` at :441-447 and :75).  The arithmetic itself runs inside
the HIP operator kernels (``sph_cubic_w``/``sph_cubic_gfac`` in
``csrc/sph_internal.h``), which follow kernels.pyx:355-417 in order.
"""
import math

CUBIC_SPLINE = 1
kernel_names = ['CubicSplineKernel']


class KernelBase:
    def __init__(self, dim=1, constant_h=-1):
        self.dim = int(dim)
        self.constant_h = constant_h
        self.has_constant_h = False
        self.fac = self._fac(1.0)

    def radius(self):
        raise NotImplementedError('KernelBase::radius')

    def dimension(self):
        return self.dim

    def _fac(self, h):
        raise NotImplementedError('KernelBase::_fac')


class CubicSplineKernel(KernelBase):
    """Cubic spline of Monaghan (1992); W and gradW evaluated on the device."""

    def _fac(self, h):
        """Normalising factor (base/kernels.pyx:441-447)."""
        if self.dim == 1:
            return (2. / 3.) / h
        elif self.dim == 2:
            return 10 / (7 * math.pi) / (h * h)
        elif self.dim == 3:
            return 1. / math.pi / (h * h * h)
        raise ValueError(self.dim)

    def radius(self):
        return 2.0

    def get_type(self):
        return CUBIC_SPLINE
