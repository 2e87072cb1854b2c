"""Import surface of the path's part of pysph.solver (solver/api.py)."""
from .integrator import (EulerIntegrator, Integrator, PredictorCorrectorIntegrator,
                         RK2Integrator)
from .solver import Solver
from .sph_equation import SPHIntegration, SPHOperation
from .time_step_functions import (TimeStep, ViscousTimeStep, ViscousAndForceBasedTimeStep,
                                  VelocityBasedTimeStep)
from .generators import get_circular_patch, elliptical_drop, dam_break_3d, jittered_lattice_3d
from .utils import savez, savez_compressed, mkdir, load
