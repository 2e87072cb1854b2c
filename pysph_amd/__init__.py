"""pysph_amd -- MI355X-native (gfx950) particle hot path of PySPH.

The package mirrors the reference's ``pysph.base`` / ``pysph.sph`` /
``pysph.solver`` API for the path it covers (ParticleArray, Particles,
CubicSplineKernel, SPHRho, TaitEquation, MomentumEquation, GravityForce,
PositionStepping, SPHOperation/SPHIntegration, SPHCalc, the Euler, RK2 and
PredictorCorrector integrators and Solver).  Particle properties are
device tensors; every computation runs in hand-written HIP kernels of
``libsphhip.so`` called through its C ABI (``include/sphhip.h``).
"""
import torch

__version__ = "0.1.0"


def default_device():
    """cuda:0 when a HIP device is visible, else cpu (storage only)."""
    if torch.cuda.is_available():
        return torch.device("cuda", 0)
    return torch.device("cpu")
