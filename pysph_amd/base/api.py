"""Import surface of the path's part of pysph.base (base/api.py:1-48)."""
from .kernels import KernelBase, CubicSplineKernel
from .nnps import (CellManager, FixedDestNbrParticleLocator, NNPSManager, NeighborLocatorType,
                   VarHNbrParticleLocator)
from .particle_array import ParticleArray, get_particle_array
from .particles import Particles
from .particle_types import ParticleType, LocalReal

Fluid = ParticleType.Fluid
Solid = ParticleType.Solid
Boundary = ParticleType.Boundary
Probe = ParticleType.Probe
DummyFluid = ParticleType.DummyFluid
