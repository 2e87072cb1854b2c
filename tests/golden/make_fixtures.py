"""Writes tests/golden/reference_fixtures.json: the input data and expected
answers that the reference's own tests hold for this path, re-encoded as
data (no reference code is copied or run).

Sources (paths relative to the reference root):
  base/tests/common_data.py:8-43           sample datasets 1 and 2
  base/tests/test_cell.py:29-61            real_to_int / find_cell_id tables
  base/tests/test_cell.py:454-471          cell-size cases
  base/tests/test_cell.py:532-556          number of cells per dataset
  base/tests/test_nnps.py:23-73            brute-force query counts
  base/tests/test_nnps.py:94-181           point queries (radius -> neighbours)
  base/tests/test_nnps.py:221-270          FixedDest locator rows
  base/tests/test_variable_h.py:14-81      variable-h 1D set-up and symmetric rows
  base/tests/test_kernels.py:33-50,111-213 1D cubic-spline quadrature setup
  sph/tests/function_test_template.py:45-58  4-particle unit square
  sph/tests/test_eos_funcs.py:88-116       Tait closed form parameters
  sph/tests/test_pressure_funcs.py:103-112 MomentumEquation parameters
  examples/elliptical_drop_exact.py:1-30   semi-analytic drop ODE constants

Run: python tests/golden/make_fixtures.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))

FIXTURES = {
    "dataset1": {
        "parr1": {"x": [0.25, 0.8, 0.5, 0.8, 0.2, 0.5, 1.5, 1.5],
                  "y": [0.25, 0.1, 0.5, 0.8, 0.9, 1.5, 0.5, 1.5],
                  "z": [0.0] * 8, "h": [1.0] * 8},
        "parr2": {"x": [0.2, 1.2, 1.5, 0.4], "y": [0.0] * 4, "z": [1.6, 1.5, -0.5, 0.4],
                  "h": [1.0] * 4},
    },
    "dataset2": {
        "parr1": {"x": [-0.5, -0.5, 0.5, 0.5, 1.5, 2.5, 2.5],
                  "y": [2.5, -0.5, 1.5, 0.5, 0.5, 0.5, -0.5],
                  "z": [0.0] * 7, "h": [1.0] * 7},
    },
    # (value, step) -> real_to_int
    "real_to_int": [[0.5, 1.0, 0], [0.1, 1.0, 0], [1.0, 1.0, 1], [1.5, 1.0, 1], [1.9, 1.0, 1],
                    [2.1, 1.0, 2], [2.6, 1.0, 2]],
    # point, cell_size -> cell id
    "find_cell_id": [[[0.0, 0.0, 1.0], 1.0, [0, 0, 1]], [[-2.01, 0.0, 1.0], 1.0, [-3, 0, 1]],
                     [[-2.01, -1.01, 1.0], 1.0, [-3, -2, 1]]],
    # max_h, radius_scale, min_cell_size, max_cell_size -> cell_size
    "cell_size": [[1.0, 2.0, -1.0, 0.0, 2.0], [1.0, 2.0, 3.0, 0.0, 3.0],
                  [2.0, 2.0, -0.1, 0.0, 4.0], [1.0, 3.0, 1.0, 2.0, 2.0],
                  [0.0, 2.0, -1.0, 0.0, 1.0]],
    # dataset, cell size -> number of non-empty cells (all arrays binned together)
    "num_cells": [["dataset1", 2.0, 2], ["dataset1", 1.0, 7], ["dataset2", 1.0, 7]],
    # brute-force queries over dataset1 arrays: array, point, radius, exclude, expected set
    "brute_force": [["parr1", [0.4, 0.0, 0.4], 1.0, -1, [0, 1, 2, 3]],
                    ["parr2", [1.0, 1.0, 1.0], 1.4142135623730951, -1, [0, 1, 3]],
                    ["parr2", [1.0, 1.0, 1.0], 1.4142135623730951, 1, [0, 3]]],
    # point queries through the cell structure (cell size 2): array, point, radius, expected
    "point_queries": [["parr1", [0.4, 0.0, 0.4], 0.5, [0]],
                      ["parr1", [0.4, 0.0, 0.4], 1.0, [0, 1, 2, 3]],
                      ["parr1", [1.5, 0.0, -0.5], 4.0, [0, 1, 2, 3, 4, 5, 6, 7]],
                      ["parr2", [1.0, 1.0, 1.0], 0.1, []],
                      ["parr2", [1.0, 1.0, 1.0], 1.0, []],
                      ["parr2", [1.0, 1.0, 1.0], 1.136, [1]],
                      ["parr2", [1.0, 1.0, 1.0], 1.358, [1, 3]],
                      ["parr2", [1.0, 1.0, 1.0], 1.4142135623730951, [0, 1, 3]],
                      ["parr2", [1.0, 1.0, 1.0], 1.88, [0, 1, 2, 3]]],
    # FixedDest rows, source parr1, dest parr2: radius_scale, dest index, expected set
    "locator_rows": [[0.5, 3, [0]], [1.0, 3, [0, 1, 2, 3]], [4.0, 2, [0, 1, 2, 3, 4, 5, 6, 7]]],
    # VarHNbrParticleLocator rows (sorted sets): 11 particles on linspace(0, 1, 11),
    # h = dx + 1e-10 except particle 5 with twice that, radius_scale 2
    "varh_1d": {"n": 11, "h_extra": 1e-10, "wide": 5, "radius_scale": 2.0,
                "rows": [[0, 1, 2], [0, 1, 2, 3, 5], [0, 1, 2, 3, 4, 5], [1, 2, 3, 4, 5],
                         [2, 3, 4, 5, 6], [1, 2, 3, 4, 5, 6, 7, 8, 9], [4, 5, 6, 7, 8],
                         [5, 6, 7, 8, 9], [5, 6, 7, 8, 9, 10], [5, 7, 8, 9, 10], [8, 9, 10]]},
    "kernel_1d": {"x0": -1.0, "x1": 1.0, "n": 21, "h": 0.2,
                  "sum_w": 1.0, "sum_gw": 0.0, "moment_w": 0.0, "moment_gw": -1.0,
                  "places": 10},
    "unit_square": {"x": [0.0, 0.0, 1.0, 1.0], "y": [0.0, 1.0, 1.0, 0.0], "z": [0.0] * 4,
                    "m": [1.0] * 4, "u": [1.0, 0.0, 0.0, -1.0], "v": [0.0] * 4, "w": [0.0] * 4,
                    "p": [0.0, 0.0, 1.0, 1.0], "h": [1.0] * 4, "rho": [1.0] * 4,
                    "cs": [0.0] * 4},
    "tait": {"gamma": 7.0, "co": 1.0, "ro": 1000.0, "places_double": 10},
    "momentum": {"alpha": 1.0, "beta": 1.0, "gamma": 1.4, "eta": 0.1, "places": 6},
    "drop_exact": {"A0": 100.0, "a0": 1.0},
}


def main():
    path = os.path.join(HERE, "reference_fixtures.json")
    with open(path, "w") as f:
        json.dump(FIXTURES, f, indent=1, sort_keys=True)
    print("wrote", path)


if __name__ == "__main__":
    main()
