"""ctypes front end of the FP64 CPU oracle (``oracle/sphref.c``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py``.  The product path never calls
it.

Besides the per-operator restatements, :class:`RefSimulation` restates the
reference time loop for the operation lists this path covers:

* ``Solver.solve`` (solver/solver.py:523-590) with a fixed ``TimeStep``
  (solver/time_step_functions.py:3-6);
* ``Integrator.eval/step`` (solver/integrator.py:287-367), Euler (:420-445),
  RK2 midpoint (:450-498) and PredictorCorrector (:614-683);
* ``SPHCalc.sph`` (sph/sph_calc.pyx:307-358): non-integrating outputs reset
  to zero first, ``to_reset`` props (``dt_fac``) zeroed per calc, one
  function per source array in array order, outputs accumulated ``+=``;
* neighbour lists rebuilt lazily after every position change
  (base/particles.py:184-227, base/nnps.pyx:561-621).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(HERE, "libsphref.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I64 = ctypes.POINTER(ctypes.c_int64)
_I32 = ctypes.POINTER(ctypes.c_int32)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        from . import build as _b
        _b.build()
    lib = ctypes.CDLL(_LIB_PATH)
    i64, i32, dbl, vp = ctypes.c_int64, ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    lib.sphref_real_to_int.argtypes = [dbl, dbl]
    lib.sphref_real_to_int.restype = ctypes.c_int
    lib.sphref_compute_cell_size.argtypes = [dbl, dbl, dbl, dbl]
    lib.sphref_compute_cell_size.restype = dbl
    lib.sphref_nnps.argtypes = [vp, vp, vp, vp, i64, vp, vp, vp, i64, dbl, dbl, i32, vp, vp, i32]
    lib.sphref_nnps.restype = i64
    lib.sphref_nnps_brute.argtypes = [vp, vp, vp, vp, i64, vp, vp, vp, i64, dbl, vp, vp, i32]
    lib.sphref_nnps_varh.argtypes = [vp, vp, vp, vp, i64, vp, vp, vp, vp, i64, i64, dbl, dbl,
                                     vp, vp]
    lib.sphref_nnps_varh.restype = i64
    lib.sphref_nnps_brute.restype = i64
    lib.sphref_cubic_spline.argtypes = [i32, i64] + [vp] * 11
    lib.sphref_cubic_spline.restype = None
    lib.sphref_summation_density.argtypes = [i32, i64] + [vp] * 12 + [i32]
    lib.sphref_summation_density.restype = None
    lib.sphref_pow.argtypes = [dbl, dbl]
    lib.sphref_pow.restype = dbl
    lib.sphref_tait.argtypes = [i64, vp, dbl, dbl, dbl, vp, vp]
    lib.sphref_tait.restype = None
    lib.sphref_momentum.argtypes = [i32, i64] + [vp] * 23 + [dbl, dbl, dbl] + [vp] * 4 + [i32]
    lib.sphref_momentum.restype = None
    lib.sphref_density_rate.argtypes = [i32, i64] + [vp] * 18 + [i32]
    lib.sphref_density_rate.restype = None
    lib.sphref_xsph.argtypes = [i32, i64] + [vp] * 19 + [dbl, vp, vp, vp, i32]
    lib.sphref_xsph.restype = None
    lib.sphref_pressure_gradient.argtypes = ([i32, i64] + [vp] * 15 + [i32, dbl, dbl, dbl] +
                                             [vp] * 3 + [i32])
    lib.sphref_pressure_gradient.restype = None
    lib.sphref_artificial_viscosity.argtypes = ([i32, i64] + [vp] * 21 + [dbl, dbl, dbl] +
                                                [vp] * 4 + [i32])
    lib.sphref_artificial_viscosity.restype = None
    lib.sphref_momentum_deltap.argtypes = ([i32, i64] + [vp] * 23 + [dbl] * 6 + [vp] * 4 +
                                           [i32])
    lib.sphref_momentum_deltap.restype = None
    lib.sphref_boundary_force.argtypes = [i64] + [vp] * 18 + [dbl, vp, vp, vp, i32]
    lib.sphref_boundary_force.restype = None
    lib.sphref_gravity.argtypes = [i64, dbl, dbl, dbl, vp, vp, vp]
    lib.sphref_gravity.restype = None
    lib.sphref_step.argtypes = [i64, vp, vp, dbl, vp]
    lib.sphref_step.restype = None
    lib.sphref_pc_final.argtypes = [i64, vp, vp]
    lib.sphref_pc_final.restype = None
    _lib = lib
    return lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be contiguous"
    return a.ctypes.data


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def real_to_int(v, step):
    """base/cell.pyx:44-61"""
    return _load().sphref_real_to_int(float(v), float(step))


def compute_cell_size(max_h, radius_scale, min_size=-1.0, max_size=0.0):
    """base/cell.pyx:911-941"""
    return _load().sphref_compute_cell_size(float(max_h), float(radius_scale),
                                            float(min_size), float(max_size))


def nnps(dst, src, radius_scale, cell_size, same_array, nthreads=1, nd_real=None):
    """FixedDestNbrParticleLocator lists (base/nnps.pyx:458-621).

    ``dst``/``src`` are dicts with x, y, z (and h for dst).  Returns
    ``(offsets int64[nd+1], indices int32[nnz])`` in reference order."""
    lib = _load()
    dx, dy, dz, dh = (_f64(dst[k]) for k in ("x", "y", "z", "h"))
    sx, sy, sz = (_f64(src[k]) for k in ("x", "y", "z"))
    nd = len(dx) if nd_real is None else int(nd_real)
    offsets = np.zeros(nd + 1, dtype=np.int64)
    nnz = lib.sphref_nnps(_p(dx), _p(dy), _p(dz), _p(dh), nd, _p(sx), _p(sy), _p(sz), len(sx),
                          float(radius_scale), float(cell_size), int(bool(same_array)),
                          _p(offsets), None, nthreads)
    indices = np.zeros(max(nnz, 1), dtype=np.int32)
    lib.sphref_nnps(_p(dx), _p(dy), _p(dz), _p(dh), nd, _p(sx), _p(sy), _p(sz), len(sx),
                    float(radius_scale), float(cell_size), int(bool(same_array)),
                    _p(offsets), _p(indices), nthreads)
    return offsets, indices[:nnz]


def nnps_varh(dst, src, radius_scale, cell_size, ns_real=None):
    """VarHNbrParticleLocator lists (base/nnps.pyx:691-786): each row is the
    destination's own query in stencil order (own index not moved) followed
    by the real sources whose own radius reaches the destination, appended
    in ascending source index.  ``dst``/``src`` need x, y, z, h.  Returns
    ``(offsets int64[nd+1], indices int32[nnz])``."""
    lib = _load()
    dx, dy, dz, dh = (_f64(dst[k]) for k in ("x", "y", "z", "h"))
    sx, sy, sz, sh = (_f64(src[k]) for k in ("x", "y", "z", "h"))
    ns_real = len(sx) if ns_real is None else int(ns_real)
    offsets = np.zeros(len(dx) + 1, dtype=np.int64)
    args = (_p(dx), _p(dy), _p(dz), _p(dh), len(dx), _p(sx), _p(sy), _p(sz), _p(sh), len(sx),
            ns_real, float(radius_scale), float(cell_size))
    nnz = lib.sphref_nnps_varh(*args, _p(offsets), None)
    indices = np.zeros(max(nnz, 1), dtype=np.int32)
    lib.sphref_nnps_varh(*args, _p(offsets), _p(indices))
    return offsets, indices[:nnz]


def nnps_brute(dst, src, radius_scale, nthreads=1, nd_real=None):
    """Brute-force O(N*M) sets, ascending (base/nnps.pyx:78-94 pattern)."""
    lib = _load()
    dx, dy, dz, dh = (_f64(dst[k]) for k in ("x", "y", "z", "h"))
    sx, sy, sz = (_f64(src[k]) for k in ("x", "y", "z"))
    nd = len(dx) if nd_real is None else int(nd_real)
    offsets = np.zeros(nd + 1, dtype=np.int64)
    nnz = lib.sphref_nnps_brute(_p(dx), _p(dy), _p(dz), _p(dh), nd, _p(sx), _p(sy), _p(sz),
                                len(sx), float(radius_scale), _p(offsets), None, nthreads)
    indices = np.zeros(max(nnz, 1), dtype=np.int32)
    lib.sphref_nnps_brute(_p(dx), _p(dy), _p(dz), _p(dh), nd, _p(sx), _p(sy), _p(sz),
                          len(sx), float(radius_scale), _p(offsets), _p(indices), nthreads)
    return offsets, indices[:nnz]


def cubic_spline(dim, a, b, h):
    """W(a, b, h) and gradW for arrays of point pairs (base/kernels.pyx:355-417)."""
    lib = _load()
    a = [_f64(c) for c in a]
    b = [_f64(c) for c in b]
    h = _f64(h)
    n = len(h)
    w = np.empty(n)
    g = [np.empty(n) for _ in range(3)]
    lib.sphref_cubic_spline(int(dim), n, *[_p(c) for c in a], *[_p(c) for c in b], _p(h),
                            _p(w), *[_p(c) for c in g])
    return w, np.stack(g)


def summation_density(dim, dst, src, offsets, indices, rho_out, nthreads=1):
    """rho_out[i] += sum_j m_j W (sph/funcs/density_funcs.pyx:37-73)."""
    lib = _load()
    nd = len(offsets) - 1
    d = [_f64(dst[k]) for k in ("x", "y", "z", "h")]
    s = [_f64(src[k]) for k in ("x", "y", "z", "h", "m")]
    lib.sphref_summation_density(int(dim), nd, *[_p(c) for c in d], *[_p(c) for c in s],
                                 _p(offsets), _p(indices), _p(rho_out), nthreads)


def power(x, e):
    """pow(x, e) of the Tait equation and the artificial pressure: correctly
    rounded for an integral exponent 1..32 (2^-20 <= |x| <= 2^20), else
    the C library's pow."""
    return _load().sphref_pow(float(x), float(e))


def tait(rho, ro, co, gamma, p_out, cs_out, n_real=None):
    """sph/funcs/eos_funcs.pyx:118-129 (accumulating into p_out, cs_out)."""
    n = len(rho) if n_real is None else int(n_real)
    _load().sphref_tait(n, _p(_f64(rho)), float(ro), float(co), float(gamma),
                        _p(p_out), _p(cs_out))


MOM_PROPS = ("x", "y", "z", "h", "rho", "p", "u", "v", "w", "cs")
MOM_SRC_PROPS = ("x", "y", "z", "h", "m", "rho", "p", "u", "v", "w", "cs")


def momentum(dim, dst, src, offsets, indices, alpha, beta, eta, ax, ay, az, dt_fac, nthreads=1):
    """sph/funcs/pressure_funcs.pyx:246-391: a += ..., dt_fac = max(...)."""
    lib = _load()
    nd = len(offsets) - 1
    d = [_f64(dst[k]) for k in MOM_PROPS]
    s = [_f64(src[k]) for k in MOM_SRC_PROPS]
    lib.sphref_momentum(int(dim), nd, *[_p(c) for c in d], *[_p(c) for c in s],
                        _p(offsets), _p(indices), float(alpha), float(beta), float(eta),
                        _p(ax), _p(ay), _p(az), _p(dt_fac), nthreads)


def density_rate(dim, dst, src, offsets, indices, out, nthreads=1):
    """out[i] += sum_j (v_i - v_j) . gradW_ij m_j (sph/funcs/density_funcs.pyx:116-164)."""
    lib = _load()
    nd = len(offsets) - 1
    d = [_f64(dst[k]) for k in ("x", "y", "z", "h", "u", "v", "w")]
    s = [_f64(src[k]) for k in ("x", "y", "z", "h", "m", "u", "v", "w")]
    lib.sphref_density_rate(int(dim), nd, *[_p(c) for c in d], *[_p(c) for c in s],
                            _p(offsets), _p(indices), _p(out), nthreads)


def xsph(dim, dst, src, offsets, indices, eps, ox, oy, oz, nthreads=1):
    """o[i] += sum_j m_j W_ij / rho_ij (v_j - v_i) eps (sph/funcs/xsph_funcs.pyx:36-91)."""
    lib = _load()
    nd = len(offsets) - 1
    d = [_f64(dst[k]) for k in ("x", "y", "z", "h", "rho", "u", "v", "w")]
    s = [_f64(src[k]) for k in ("x", "y", "z", "h", "m", "rho", "u", "v", "w")]
    lib.sphref_xsph(int(dim), nd, *[_p(c) for c in d], *[_p(c) for c in s], _p(offsets),
                    _p(indices), float(eps), _p(ox), _p(oy), _p(oz), nthreads)


def pressure_gradient(dim, dst, src, offsets, indices, ax, ay, az, deltap=None, n=1.0,
                      eps=0.2, nthreads=1):
    """a += -m_b (P_a/rho_a^2 + P_b/rho_b^2 [- artificial pressure]) gradW
    (sph/funcs/pressure_funcs.pyx:61-151)."""
    lib = _load()
    nd = len(offsets) - 1
    d = [_f64(dst[k]) for k in ("x", "y", "z", "h", "rho", "p")]
    s = [_f64(src[k]) for k in ("x", "y", "z", "h", "m", "rho", "p")]
    lib.sphref_pressure_gradient(int(dim), nd, *[_p(c) for c in d], *[_p(c) for c in s],
                                 _p(offsets), _p(indices), int(deltap is not None),
                                 float(deltap or 0.0), float(n), float(eps), _p(ax), _p(ay),
                                 _p(az), nthreads)


def artificial_viscosity(dim, dst, src, offsets, indices, alpha, beta, eta, ax, ay, az,
                         dt_fac, nthreads=1):
    """Monaghan Pi_ab gradW with dt_fac (sph/funcs/viscosity_funcs.pyx:62-151)."""
    lib = _load()
    nd = len(offsets) - 1
    d = [_f64(dst[k]) for k in ("x", "y", "z", "h", "rho", "u", "v", "w", "cs")]
    s = [_f64(src[k]) for k in ("x", "y", "z", "h", "m", "rho", "u", "v", "w", "cs")]
    lib.sphref_artificial_viscosity(int(dim), nd, *[_p(c) for c in d], *[_p(c) for c in s],
                                    _p(offsets), _p(indices), float(alpha), float(beta),
                                    float(eta), _p(ax), _p(ay), _p(az), _p(dt_fac), nthreads)


def momentum_deltap(dim, dst, src, offsets, indices, alpha, beta, eta, deltap, n, eps, ax, ay,
                    az, dt_fac, nthreads=1):
    """MomentumEquation with artificial pressure (sph/funcs/pressure_funcs.pyx:246-391)."""
    lib = _load()
    nd = len(offsets) - 1
    d = [_f64(dst[k]) for k in MOM_PROPS]
    s = [_f64(src[k]) for k in MOM_SRC_PROPS]
    lib.sphref_momentum_deltap(int(dim), nd, *[_p(c) for c in d], *[_p(c) for c in s],
                               _p(offsets), _p(indices), float(alpha), float(beta), float(eta),
                               float(deltap), float(n), float(eps), _p(ax), _p(ay), _p(az),
                               _p(dt_fac), nthreads)


def boundary_force(dst, src, offsets, indices, delp, ax, ay, az, nthreads=1):
    """Monaghan wall force from tangent/normal frames (sph/funcs/boundary_funcs.pyx:55-107)."""
    lib = _load()
    nd = len(offsets) - 1
    d = [_f64(dst[k]) for k in ("x", "y", "z", "h", "m", "cs")]
    s = [_f64(src[k]) for k in ("x", "y", "z", "m", "tx", "ty", "tz", "nx", "ny", "nz")]
    lib.sphref_boundary_force(nd, *[_p(c) for c in d], *[_p(c) for c in s], _p(offsets),
                              _p(indices), float(delp), _p(ax), _p(ay), _p(az), nthreads)


def gravity(n, g, ax, ay, az):
    _load().sphref_gravity(int(n), float(g[0]), float(g[1]), float(g[2]), _p(ax), _p(ay), _p(az))


def step(x0, a, dt, x):
    """solver/integrator.py:356-360"""
    _load().sphref_step(len(x), _p(x0), _p(a), float(dt), _p(x))


def pc_final(x0, x):
    """solver/integrator.py:632-643"""
    _load().sphref_pc_final(len(x), _p(x0), _p(x))


# ---------------------------------------------------------------------------
# whole-run restatement
# ---------------------------------------------------------------------------

class RefArray:
    """Minimal SoA particle store for the oracle (dict of float64 arrays)."""

    def __init__(self, name, ptype, props):
        self.name = name
        self.ptype = ptype
        self.props = {k: np.array(v, dtype=np.float64) for k, v in props.items()}
        n = len(self.props["x"])
        for k, dv in (("x", 0.0), ("y", 0.0), ("z", 0.0), ("u", 0.0), ("v", 0.0), ("w", 0.0),
                      ("m", 1.0), ("h", 1.0), ("p", 0.0), ("e", 0.0), ("rho", 1.0),
                      ("cs", 0.0), ("dt_fac", 0.0)):
            if k not in self.props:
                self.props[k] = np.full(n, dv)
        self.n = n

    def __getitem__(self, k):
        return self.props[k]


class RefSimulation:
    """The WCSPH operation list of this path, driven like Solver.solve.

    Operations, in solver order (sph_equation.py:101-228, solver.py:222-282):
      1. SPHRho            on fluids, from ``density_from`` types, updates rho
         (``density="continuity"``: SPHDensityRate instead, integrating rho)
      2. TaitEquation      on fluids, updates p, cs
      3. MomentumEquation  on fluids, from ``momentum_from`` types, integrates u,v(,w)
      4. GravityForce      on fluids (if ``gravity`` given), integrates u,v(,w)
      5. PositionStepping  on fluids, integrates x,y(,z)
      6. XSPHCorrection    on fluids from fluids (if ``xsph_eps`` given), integrates
         x,y(,z) (Solver.add_operation_xsph, solver.py:186-220)

    ``momentum="split"`` replaces 3. by SPHPressureGradient from ``pgrad_from``
    types followed by MonaghanArtificialViscosity (its eta: ``avisc_eta``) from
    ``momentum_from`` types; ``deltap`` (with ``ap_n``, ``ap_eps``) switches the
    artificial pressure of the momentum equation or the pressure gradient on;
    ``bforce_from`` types add MonaghanBoundaryForce(delp=``bforce_delp``) after
    gravity (the order of sph/examples' dam_break_simple.py).
    """

    def __init__(self, arrays, dim, integrator="rk2", co=1.0, ro=1000.0, gamma=7.0,
                 alpha=1.0, beta=1.0, eta=0.1, gravity=None, density_from=(0,),
                 momentum_from=(0,), cell_radius_scale=2.0, nthreads=1, density="summation",
                 xsph_eps=None, momentum="full", pgrad_from=(0,), avisc_eta=0.0, deltap=None,
                 ap_n=1.0, ap_eps=0.2, bforce_from=(), bforce_delp=1.0, variable_h=False):
        self.arrays = arrays
        self.dim = dim
        self.integrator = integrator
        self.co, self.ro, self.gamma = co, ro, gamma
        self.alpha, self.beta, self.eta = alpha, beta, eta
        self.gravity = gravity
        self.density_from = density_from
        self.momentum_from = momentum_from
        self.cell_radius_scale = cell_radius_scale
        self.nthreads = nthreads
        self.fluids = [a for a in arrays if a.ptype == 0]
        self.vel = ["u", "v", "w"][:dim]
        self.pos = ["x", "y", "z"][:dim]
        self.stepped = self.vel + self.pos  # momentum/gravity updates, then position step
        if density not in ("summation", "continuity"):
            raise ValueError(density)
        self.density = density
        if density == "continuity":
            self.stepped = ["rho"] + self.stepped
        self.xsph_eps = xsph_eps
        if momentum not in ("full", "split"):
            raise ValueError(momentum)
        self.momentum = momentum
        self.pgrad_from = pgrad_from if momentum == "split" else ()
        self.avisc_eta = avisc_eta
        self.deltap, self.ap_n, self.ap_eps = deltap, ap_n, ap_eps
        self.bforce_from = bforce_from
        self.bforce_delp = bforce_delp
        self.variable_h = variable_h
        self.lists = None
        self.dirty = True
        self.pairs = 0

    # -- NNPS -----------------------------------------------------------
    def cell_size(self):
        max_h = max(float(np.max(a["h"])) for a in self.arrays if a.n > 0)
        return compute_cell_size(max_h, self.cell_radius_scale)

    def rebuild(self):
        cs = self.cell_size()
        self.lists = {}
        for d in self.fluids:
            for s in self.arrays:
                used = (set(self.density_from) | set(self.momentum_from) |
                        set(self.pgrad_from) | set(self.bforce_from))
                if s.ptype in used or (self.xsph_eps is not None and s.ptype == 0):
                    if self.variable_h:
                        self.lists[(d.name, s.name)] = nnps_varh(d, s, 2.0, cs)
                    else:
                        self.lists[(d.name, s.name)] = nnps(d, s, 2.0, cs, d is s,
                                                            self.nthreads)
        self.dirty = False

    def _lists(self):
        if self.dirty or self.lists is None:
            self.rebuild()
        return self.lists

    # -- one RHS evaluation (Integrator.eval) ------------------------------
    def eval(self, acc):
        dim = self.dim
        lists = self._lists()
        for d in self.fluids:
            if self.density == "continuity":
                a_rho = acc[d.name]["rho"]
                a_rho[:] = 0.0
                for s in self.arrays:
                    if s.ptype in self.density_from:
                        off, idx = lists[(d.name, s.name)]
                        density_rate(dim, d, s, off, idx, a_rho, self.nthreads)
                        self.pairs += len(idx)
                continue
            d["rho"][:] = 0.0
            for s in self.arrays:
                if s.ptype in self.density_from:
                    off, idx = lists[(d.name, s.name)]
                    summation_density(dim, d, s, off, idx, d["rho"], self.nthreads)
                    self.pairs += len(idx)
        for d in self.fluids:
            d["p"][:] = 0.0
            d["cs"][:] = 0.0
            tait(d["rho"], self.ro, self.co, self.gamma, d["p"], d["cs"])
        for d in self.fluids:
            a = acc[d.name]
            for k in self.vel + self.pos:
                a[k][:] = 0.0
            d["dt_fac"][:] = 0.0
            az = a["w"] if dim == 3 else np.zeros(d.n)
            if self.momentum == "split":
                for s in self.arrays:
                    if s.ptype in self.pgrad_from:
                        off, idx = lists[(d.name, s.name)]
                        pressure_gradient(dim, d, s, off, idx, a["u"], a["v"], az,
                                          deltap=self.deltap, n=self.ap_n, eps=self.ap_eps,
                                          nthreads=self.nthreads)
                        self.pairs += len(idx)
            for s in self.arrays:
                if s.ptype in self.momentum_from:
                    off, idx = lists[(d.name, s.name)]
                    if self.momentum == "split":
                        artificial_viscosity(dim, d, s, off, idx, self.alpha, self.beta,
                                             self.avisc_eta, a["u"], a["v"], az, d["dt_fac"],
                                             self.nthreads)
                    elif self.deltap is not None:
                        momentum_deltap(dim, d, s, off, idx, self.alpha, self.beta, self.eta,
                                        self.deltap, self.ap_n, self.ap_eps, a["u"], a["v"],
                                        az, d["dt_fac"], self.nthreads)
                    else:
                        momentum(dim, d, s, off, idx, self.alpha, self.beta, self.eta,
                                 a["u"], a["v"], az, d["dt_fac"], self.nthreads)
                    self.pairs += len(idx)
            if self.gravity is not None:
                gravity(d.n, self.gravity, a["u"], a["v"], az)
            for s in self.arrays:
                if s.ptype in self.bforce_from:
                    off, idx = lists[(d.name, s.name)]
                    boundary_force(d, s, off, idx, self.bforce_delp, a["u"], a["v"], az,
                                   self.nthreads)
                    self.pairs += len(idx)
            for vk, xk in zip(self.vel, self.pos):
                a[xk][:] = d[vk]
            if self.xsph_eps is not None:
                zero = np.zeros(d.n)
                ax_ = [a[k] for k in self.pos] + [zero] * (3 - dim)
                for s in self.fluids:
                    off, idx = lists[(d.name, s.name)]
                    xsph(dim, d, s, off, idx, self.xsph_eps, ax_[0], ax_[1], ax_[2],
                         self.nthreads)
                    self.pairs += len(idx)

    def _step(self, acc, init, dt):
        for d in self.fluids:
            for k in self.stepped:
                step(init[d.name][k], acc[d.name][k], dt, d[k])
        self.dirty = True

    def integrate(self, dt):
        acc = {d.name: {k: np.zeros(d.n) for k in self.stepped} for d in self.fluids}
        init = {d.name: {k: d[k].copy() for k in self.stepped} for d in self.fluids}
        if self.integrator == "euler":
            self.eval(acc)
            self._step(acc, init, dt)
        elif self.integrator == "rk2":
            self.eval(acc)
            self._step(acc, init, dt / 2)
            self.eval(acc)
            self._step(acc, init, dt)
        elif self.integrator == "pc":
            self.eval(acc)
            self._step(acc, init, dt / 2)
            self.eval(acc)
            self._step(acc, init, dt / 2)
            for d in self.fluids:
                for k in self.stepped:
                    pc_final(init[d.name][k], d[k])
            self.dirty = True
        else:
            raise ValueError(self.integrator)
        return acc

    def run(self, dt, nsteps):
        acc = None
        for _ in range(nsteps):
            acc = self.integrate(dt)
        return acc
