"""Pinning of the oracle's split-form operators and wall force against closed
forms, conservation laws and special cases (CPU only).

SPHPressureGradient  sph/funcs/pressure_funcs.pyx:19-168
MonaghanArtificialViscosity  sph/funcs/viscosity_funcs.pyx:19-157
MomentumEquation deltap  sph/funcs/pressure_funcs.pyx:327-365
MonaghanBoundaryForce  sph/funcs/boundary_funcs.pyx:9-107
"""
import math

import numpy as np
import pytest

from oracle import cpu_ref


def _sigma3(h):
    return 1.0 / (math.pi * h ** 3)


def _dwdr(r, h):
    """dW/dr of the 3D cubic spline, from W(q) = s(1 - 1.5 q^2 + 0.75 q^3),
    s/4 (2 - q)^3 (Monaghan 1992)."""
    q = r / h
    s = _sigma3(h)
    if q < 1:
        return s * (-3.0 * q + 2.25 * q * q) / h
    if q < 2:
        return -0.75 * s * (2 - q) ** 2 / h
    return 0.0


def _w(r, h):
    q = r / h
    s = _sigma3(h)
    if q < 1:
        return s * (1 - 1.5 * q * q + 0.75 * q ** 3)
    if q < 2:
        return 0.25 * s * (2 - q) ** 3
    return 0.0


def _pair(r, h=1.0, **props):
    """Destination 0 at the origin, source 1 at (r, 0, 0); both in one array."""
    d = {"x": np.array([0.0, r]), "y": np.zeros(2), "z": np.zeros(2), "h": np.full(2, h),
         "m": np.array([1.0, 2.0]), "rho": np.array([1000.0, 1010.0]),
         "p": np.array([50.0, 80.0]), "u": np.zeros(2), "v": np.zeros(2), "w": np.zeros(2),
         "cs": np.array([10.0, 12.0])}
    for k, v in props.items():
        d[k] = np.asarray(v, dtype=float)
    offsets = np.array([0, 1], dtype=np.int64)
    indices = np.array([1], dtype=np.int32)
    return d, offsets, indices


def _cloud(n=300, seed=3, uniform_h=True):
    rng = np.random.default_rng(seed)
    d = {k: rng.uniform(0, 1, n) for k in ("x", "y", "z")}
    d["h"] = np.full(n, 0.12) if uniform_h else rng.uniform(0.1, 0.14, n)
    d["m"] = rng.uniform(0.5, 1.5, n)
    d["rho"] = rng.uniform(990, 1010, n)
    d["p"] = rng.uniform(-200, 300, n)
    d["u"], d["v"], d["w"] = rng.standard_normal((3, n))
    d["cs"] = rng.uniform(9, 11, n)
    off, idx = cpu_ref.nnps_brute(d, d, 2.0)
    return d, off, idx


def _out(n):
    return np.zeros(n), np.zeros(n), np.zeros(n)


@pytest.mark.parametrize("r", [0.3, 0.9, 1.0, 1.5, 1.99])
def test_pressure_gradient_two_particles(r):
    d, off, idx = _pair(r)
    ax, ay, az = _out(1)
    cpu_ref.pressure_gradient(3, d, d, off, idx, ax, ay, az)
    s = d["p"][0] / d["rho"][0] ** 2 + d["p"][1] / d["rho"][1] ** 2
    # grad_a W = dW/dr * (x_a - x_b)/r = -dW/dr along x
    want = -d["m"][1] * s * (-_dwdr(r, 1.0))
    assert ax[0] == pytest.approx(want, rel=1e-13)
    assert ay[0] == 0 and az[0] == 0


@pytest.mark.parametrize("deltap", [None, 0.8])
@pytest.mark.parametrize("uniform_h", [True, False])
def test_pressure_gradient_conserves_momentum(deltap, uniform_h):
    d, off, idx = _cloud(uniform_h=uniform_h)
    n = len(d["x"])
    ax, ay, az = _out(n)
    cpu_ref.pressure_gradient(3, d, d, off, idx, ax, ay, az, deltap=deltap, n=4, eps=0.2)
    if not uniform_h:
        # lists are gathered with the destination's h: not symmetric
        return
    scale = np.abs(d["m"] * ax).sum()
    for a in (ax, ay, az):
        assert abs(np.sum(d["m"] * a)) <= 1e-12 * scale


def test_pressure_gradient_artificial_pressure():
    """Positive pressures: the correction adds m_b*(0+0)*fab = 0, so the result
    is the uncorrected one bit for bit.  Negative pressures: the extra term is
    m_b (eps|Pa|/rho_a^2 + eps|Pb|/rho_b^2) (W(r)/W(deltap))^n."""
    d, off, idx = _cloud()
    d["p"] = np.abs(d["p"])
    n = len(d["x"])
    a0, a1 = _out(n), _out(n)
    cpu_ref.pressure_gradient(3, d, d, off, idx, *a0)
    cpu_ref.pressure_gradient(3, d, d, off, idx, *a1, deltap=0.1, n=4, eps=0.2)
    for u, v in zip(a0, a1):
        assert np.array_equal(u, v)
    r, dp, nexp, eps = 0.7, 0.4, 4.0, 0.3
    d, off, idx = _pair(r, p=[-40.0, -70.0])
    a = _out(1)
    cpu_ref.pressure_gradient(3, d, d, off, idx, *a, deltap=dp, n=nexp, eps=eps)
    pa, pb = d["p"]
    ra, rb = d["rho"]
    temp = (-d["m"][1] * (pa / ra ** 2 + pb / rb ** 2) +
            d["m"][1] * (eps * -pa / ra ** 2 + eps * -pb / rb ** 2) *
            (_w(r, 1.0) / _w(dp, 1.0)) ** nexp)
    assert a[0][0] == pytest.approx(temp * -_dwdr(r, 1.0), rel=1e-12)
    # W(deltap) = 0 (deltap beyond the support): every pair raises in the
    # reference and adds nothing
    a = _out(1)
    cpu_ref.pressure_gradient(3, d, d, off, idx, *a, deltap=2.5, n=nexp, eps=eps)
    assert a[0][0] == 0.0


def test_artificial_viscosity_closed_form():
    """Approaching pair: Pi = (-alpha cbar mu + beta mu^2)/rhobar with
    mu = h v.r/r^2 (eta is 0 in the reference's function); receding: 0;
    dt_fac = |h v.r / (r^2 + 0.01 h^2)| either way."""
    r, h, alpha, beta = 0.8, 1.0, 0.1, 0.3
    seen = set()
    for ua in (-3.0, 3.0):  # a at x=0, b at x=r: -x separates them, +x closes them
        d, off, idx = _pair(r, u=[ua, 0.0])
        a, dtf = _out(1), np.zeros(1)
        cpu_ref.artificial_viscosity(3, d, d, off, idx, alpha, beta, 0.0, *a, dtf)
        dot = ua * (0.0 - r)
        assert dtf[0] == pytest.approx(abs(h * dot / (r * r + 0.01 * h * h)), rel=1e-15)
        if dot >= 0:
            assert a[0][0] == 0.0
        else:
            mu = h * dot / (r * r)
            pi = (-alpha * 11.0 * mu + beta * mu * mu) / 1005.0
            assert a[0][0] == pytest.approx(-d["m"][1] * pi * -_dwdr(r, h), rel=1e-12)
        seen.add(dot < 0)
    assert seen == {True, False}
    # eta set on the function enters mu only
    d, off, idx = _pair(r, u=[3.0, 0.0])
    a, dtf = _out(1), np.zeros(1)
    cpu_ref.artificial_viscosity(3, d, d, off, idx, alpha, beta, 0.5, *a, dtf)
    dot = 3.0 * -r
    mu = h * dot / (r * r + 0.25 * h * h)
    pi = (-alpha * 11.0 * mu + beta * mu * mu) / 1005.0
    assert a[0][0] == pytest.approx(-d["m"][1] * pi * -_dwdr(r, h), rel=1e-12)
    assert dtf[0] == pytest.approx(abs(h * dot / (r * r + 0.01 * h * h)), rel=1e-15)


def test_artificial_viscosity_conserves_momentum():
    d, off, idx = _cloud()
    n = len(d["x"])
    a, dtf = _out(n), np.zeros(n)
    cpu_ref.artificial_viscosity(3, d, d, off, idx, 1.0, 1.0, 0.0, *a, dtf)
    scale = np.abs(d["m"] * a[0]).sum()
    assert scale > 0
    for c in a:
        assert abs(np.sum(d["m"] * c)) <= 1e-12 * scale
    assert np.all(dtf >= 0)


def test_momentum_deltap_reduces_and_adds():
    d, off, idx = _cloud()
    n = len(d["x"])
    d["p"] = np.abs(d["p"])
    a0, f0 = _out(n), np.zeros(n)
    a1, f1 = _out(n), np.zeros(n)
    cpu_ref.momentum(3, d, d, off, idx, 0.1, 0.2, 0.1, *a0, f0)
    cpu_ref.momentum_deltap(3, d, d, off, idx, 0.1, 0.2, 0.1, 0.1, 4.0, 0.2, *a1, f1)
    for u, v in zip(a0 + (f0,), a1 + (f1,)):
        assert np.array_equal(u, v)
    r, dp, nexp, eps = 0.6, 0.5, 3.0, 0.25
    d, off, idx = _pair(r, p=[-40.0, 30.0], u=[1.0, 0.0])
    a0, f0 = _out(1), np.zeros(1)
    a1, f1 = _out(1), np.zeros(1)
    cpu_ref.momentum(3, d, d, off, idx, 0.1, 0.0, 0.1, *a0, f0)
    cpu_ref.momentum_deltap(3, d, d, off, idx, 0.1, 0.0, 0.1, dp, nexp, eps, *a1, f1)
    extra = d["m"][1] * (eps * 40.0 / d["rho"][0] ** 2) * (_w(r, 1.0) / _w(dp, 1.0)) ** nexp
    assert a1[0][0] - a0[0][0] == pytest.approx(extra * -_dwdr(r, 1.0), rel=1e-9)
    assert f1[0] == f0[0]
    # W(deltap) = 0: dt_fac still updated, no acceleration from the pair
    a2, f2 = _out(1), np.zeros(1)
    cpu_ref.momentum_deltap(3, d, d, off, idx, 0.1, 0.0, 0.1, 2.5, nexp, eps, *a2, f2)
    assert a2[0][0] == 0.0 and f2[0] == f0[0]


def _gamma(q):
    """Monaghan's wall function of q = y/h, written as 2q - 1.5q^2 on (2/3, 1]."""
    if 0 < q <= 2 / 3:
        return 2 / 3
    if 2 / 3 < q <= 1:
        return 2 * q - 1.5 * q * q
    if 1 < q <= 2:
        return 0.5 * (2 - q) ** 2
    return 0.0


def _wall_case(xoff, y, delp=0.5, h=0.4):
    dst = {"x": np.array([xoff]), "y": np.array([y]), "z": np.zeros(1), "h": np.array([h]),
           "m": np.array([1.5]), "cs": np.array([20.0])}
    src = {"x": np.zeros(1), "y": np.zeros(1), "z": np.zeros(1), "m": np.array([3.0]),
           "tx": np.ones(1), "ty": np.zeros(1), "tz": np.zeros(1), "nx": np.zeros(1),
           "ny": np.ones(1), "nz": np.zeros(1)}
    a = _out(1)
    cpu_ref.boundary_force(dst, src, np.array([0, 1], np.int64), np.array([0], np.int32), delp,
                           *a)
    return a


@pytest.mark.parametrize("q", [0.1, 0.5, 2 / 3, 0.8, 1.0, 1.3, 1.9])
@pytest.mark.parametrize("xoff", [0.0, 0.2, -0.45])
def test_boundary_force_closed_form(q, xoff):
    h, delp = 0.4, 0.5
    y = q * h
    a = _wall_case(xoff, y, delp, h)
    want = (3.0 / 4.5) * _gamma(y / h) * (1 - abs(xoff) / delp) * 0.02 * 400.0 / y
    assert a[1][0] == pytest.approx(want, rel=1e-13)
    assert a[0][0] == 0.0 and a[2][0] == 0.0


def test_boundary_force_special_cases():
    # beyond delp along the wall: no force
    assert _wall_case(0.6, 0.2)[1][0] == 0.0
    # below the wall (q < 0): Gamma = 0
    assert _wall_case(0.1, -0.2)[1][0] == 0.0
    # on the wall plane (y = 0): the reference's division raises, the pair adds nothing
    a = _wall_case(0.1, 0.0)
    assert a[1][0] == 0.0 and not np.isnan(a[1][0])
    # Gamma is continuous at q = 2/3 and q = 1
    for qc in (2 / 3, 1.0):
        lo, hi = _wall_case(0.0, qc * 0.4 * (1 - 1e-12)), _wall_case(0.0, qc * 0.4 * (1 + 1e-12))
        assert lo[1][0] == pytest.approx(hi[1][0], rel=1e-9)
