"""Particles: the arrays of a simulation with their cell manager and
neighbour-list manager (base/particles.py:36-318).

The reference constructor passes ``max_radius_scale`` to its CellManager,
which that snapshot's CellManager does not accept (particles.py:165-169 vs
cell.pyx:802-805); the intent, a cell size of ``max_radius_scale * max(h)``
with ``max_radius_scale = 2`` (particles.py:110), is what is built here.
"""
from .nnps import CellManager, NeighborLocatorType, NNPSManager


class Particles:

    def __init__(self, arrays=(), in_parallel=False, variable_h=False, load_balancing=True,
                 locator_type=NeighborLocatorType.SPHNeighborLocator, periodic_domain=None,
                 min_cell_size=-1, max_cell_size=0, max_radius_scale=2, update_particles=True,
                 cell_order="morton", precision=None):
        assert len(arrays) > 0, "Particles must be given some arrays!"
        if in_parallel:
            raise NotImplementedError("the MPI path is not part of this package (one GPU)")
        if periodic_domain is not None:
            raise NotImplementedError("periodic domains are not part of this path")
        prec = arrays[0].cl_precision
        for arr in arrays[1:]:
            assert arr.cl_precision == prec, "All arrays must have the same cl_precision"
        # the arrays' cl_precision selects the device arithmetic, as it does
        # for the reference's OpenCL path: "single" -> float records, lists
        # and operators
        self.precision = precision or ("single" if prec == "single" else "double")
        self.variable_h = variable_h
        self.in_parallel = False
        self.load_balancing = load_balancing
        self.locator_type = locator_type
        self.min_cell_size = min_cell_size
        self.max_cell_size = max_cell_size
        self.periodic_domain = periodic_domain
        self.parallel_manager = None
        self.max_radius_scale = max_radius_scale
        self.cell_order = cell_order
        self.arrays = list(arrays)
        self.kernel = None
        self.correction_manager = None
        self.misc_prop_update_functions = []
        self.initialize()

    def initialize(self):
        self.cell_manager = CellManager(arrays_to_bin=self.arrays,
                                        min_cell_size=self.min_cell_size,
                                        max_cell_size=self.max_cell_size,
                                        radius_scale=self.max_radius_scale,
                                        order=self.cell_order, precision=self.precision)
        self.nnps_manager = NNPSManager(cell_manager=self.cell_manager,
                                        variable_h=self.variable_h,
                                        locator_type=self.locator_type)
        self.update()

    def update(self, cache_neighbors=False):
        """Mark locators dirty after particles moved (particles.py:184-227).

        The rebuild itself is lazy, at the next neighbour query."""
        err = self.nnps_manager.update()
        assert err != -1, 'NNPSManager update failed! '
        self.evaluate_misc_properties()

    def evaluate_misc_properties(self):
        for func in self.misc_prop_update_functions:
            func.eval()

    def barrier(self):
        pass

    def update_remote_particle_properties(self, update_prop_list=None):
        pass

    def get_named_particle_array(self, name):
        for a in self.arrays:
            if a.name == name:
                return a
        return None
