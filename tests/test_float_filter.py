"""The float distance filter of the list kernel (sph_nnps.hip): its band
B = NN_BAND * D^2 must bound |d^2_float - d^2_FP64| for coordinates within D
of the cell origin, or a list could differ from the FP64 reference lists.
Emulated here in numpy with the kernel's exact operation order."""
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _band():
    src = open(os.path.join(ROOT, "pysph_amd", "csrc", "sph_nnps.hip")).read()
    m = re.search(r"constexpr double NN_BAND = ([0-9.]+) \* 0x1p-24;", src)
    assert m, "NN_BAND not found"
    return float(m.group(1))


def test_band_bounds_the_float_distance_error():
    u = 2.0 ** -24
    band = _band()
    rng = np.random.default_rng(5)
    worst = 0.0
    for trial in range(24):
        O = rng.uniform(-1e4, 1e4, 3) if trial % 2 else rng.uniform(-5, 5, 3)
        D = 10.0 ** rng.uniform(-6, 1)
        n = 50000
        o = O + rng.uniform(0, D / 4, 3)
        lo, hi = (o - 0.999 * D)[:, None], (o + 0.999 * D)[:, None]
        s = np.clip(O[:, None] + rng.uniform(-D, D, (3, n)), lo, hi)
        t = np.clip(O[:, None] + rng.uniform(-D, D, (3, n)), lo, hi)
        fs = (s - o[:, None]).astype(np.float32)      # FP64 shift, then float
        ft = (t - o[:, None]).astype(np.float32)
        g = fs - ft                                    # float arithmetic
        g2 = (g[0] * g[0] + g[1] * g[1]) + g[2] * g[2]
        e = s - t                                      # reference FP64 order
        d2 = (e[0] * e[0] + e[1] * e[1]) + e[2] * e[2]
        worst = max(worst, float(np.max(np.abs(g2.astype(np.float64) - d2)) / (u * D * D)))
    assert worst < band / 2, (worst, band)
