"""Import surface of the path's part of pysph.sph (sph/api.py)."""
from .sph_calc import SPHCalc, HIPCalc
from .sph_func import Function, SPHFunction, SPHFunctionParticle, CSPHFunctionParticle
from .funcs import (SPHRho, TaitEquation, MomentumEquation, GravityForce, PositionStepping,
                    SPHDensityRate, XSPHCorrection, SPHPressureGradient,
                    MonaghanArtificialViscosity, MonaghanBoundaryForce)
