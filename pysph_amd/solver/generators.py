"""Synthetic particle sets of the benchmark configurations (numpy, host).

Every generator returns plain numpy arrays in a dict so the same input can
feed a ParticleArray and a CPU check.

* ``get_circular_patch`` -- the lattice drop of solver/fluid_solver.py:16-53:
  ``numpy.mgrid[-1.05:1.05+1e-4:dx]`` in x and y, particles with
  ``sqrt(x^2+y^2) - 1 > 1e-10`` removed by the reference's swap-with-last
  removal (carray ``remove``), so the particle order is the reference's.
* ``elliptical_drop`` -- the 2D drop of the elliptical-drop runs: radius R,
  h = 1.3 dx, m = rho0 dx^2, u = -s x, v = s y.
* ``jittered_lattice_3d`` -- n^3 lattice at (i+0.5)dx with per-coordinate
  uniform jitter, seeded PCG64.
* ``dam_break_3d`` -- fluid column 1.0 x 0.5 x 2.0 at (i+0.5)dx inside a
  4.0 x 0.5 x 2.5 open tank (floor, two end walls, two side walls) made of
  three layers of wall particles at -0.5dx, -1.5dx, -2.5dx outside.
"""
import math

import numpy as np


def _swap_remove_order(n, remove):
    """Particle order left by removing ``remove`` with swap-with-last
    (carray remove: descending over the sorted indices)."""
    perm = np.arange(n)
    length = n
    for i in np.sort(np.asarray(remove, dtype=np.int64))[::-1]:
        if i < length:
            perm[i] = perm[length - 1]
            length -= 1
    return perm[:length]


def circular_patch_positions(dx, radius=1.0, margin=0.05):
    lo = radius + margin
    x, y = np.mgrid[-lo:lo + 1e-4:dx, -lo:lo + 1e-4:dx]
    x = x.ravel()
    y = y.ravel()
    remove = np.nonzero(np.sqrt(x * x + y * y) - radius > 1e-10)[0]
    order = _swap_remove_order(len(x), remove)
    return x[order].copy(), y[order].copy()


def elliptical_drop(dx=0.025, radius=1.0, rho0=1.0, c0=1400.0, vel=100.0, hdx=1.3):
    """2D drop: h = hdx*dx, m = rho0*dx^2, u = -vel*x, v = vel*y."""
    x, y = circular_patch_positions(dx, radius)
    n = len(x)
    return dict(x=x, y=y, z=np.zeros(n), u=-vel * x, v=vel * y, w=np.zeros(n),
                h=np.full(n, hdx * dx), m=np.full(n, rho0 * dx * dx), rho=np.full(n, rho0),
                p=np.zeros(n), cs=np.full(n, c0))


def get_circular_patch(name="", type=0, dx=0.025 / 1.3, cl_precision="double", **kwargs):
    """The reference's canned drop (fluid_solver.py:16-53) as a ParticleArray."""
    from ..base.particle_array import get_particle_array
    x, y = circular_patch_positions(dx)
    m = np.ones_like(x) * dx * dx
    h = np.ones_like(x) * 2 * dx
    rho = np.ones_like(x)
    p = 0.5 * 1.0 * 100 * 100 * (1 - (x ** 2 + y ** 2))
    cs = np.ones_like(x) * 100.0
    u = -100 * x
    v = 100 * y
    return get_particle_array(x=x, y=y, m=m, rho=rho, h=h, p=p, u=u, v=v, cs=cs, name=name,
                              type=type, cl_precision=cl_precision)


def jittered_lattice_3d(n=100, dx=None, jitter=0.2, seed=20261202, hdx=1.2, rho0=1.0):
    """n^3 particles at (i+0.5)dx (+ U(-jitter*dx, jitter*dx) per coordinate)."""
    if dx is None:
        dx = 1.0 / n
    g = (np.arange(n) + 0.5) * dx
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    N = n ** 3
    rng = np.random.default_rng(seed)
    j = rng.uniform(-jitter * dx, jitter * dx, size=(3, N))
    x = X.ravel() + j[0]
    y = Y.ravel() + j[1]
    z = Z.ravel() + j[2]
    return dict(x=x, y=y, z=z, h=np.full(N, hdx * dx), m=np.full(N, rho0 * dx ** 3),
                rho=np.full(N, rho0))


def dam_break_3d(dx=0.01, fluid=(1.0, 0.5, 2.0), tank=(4.0, 0.5, 2.5), layers=3, rho0=1000.0,
                 hdx=1.3, g=9.81):
    """Fluid column and wall particles of the 3D dam break (z is up)."""
    c0 = 10.0 * math.sqrt(2 * g * fluid[2])
    nf = [int(round(fluid[a] / dx)) for a in range(3)]
    gx, gy, gz = [(np.arange(nf[a]) + 0.5) * dx for a in range(3)]
    X, Y, Z = np.meshgrid(gx, gy, gz, indexing="ij")
    xf, yf, zf = X.ravel(), Y.ravel(), Z.ravel()
    nt = [int(round(tank[a] / dx)) for a in range(3)]
    ix = np.arange(-layers, nt[0] + layers)
    iy = np.arange(-layers, nt[1] + layers)
    iz = np.arange(-layers, nt[2])  # open top
    I, J, K = np.meshgrid(ix, iy, iz, indexing="ij")
    wall = (I < 0) | (I >= nt[0]) | (J < 0) | (J >= nt[1]) | (K < 0)
    xw = (I[wall] + 0.5) * dx
    yw = (J[wall] + 0.5) * dx
    zw = (K[wall] + 0.5) * dx
    h = hdx * dx
    m = rho0 * dx ** 3
    nfl, nw = len(xf), len(xw)
    fl = dict(x=xf, y=yf, z=zf, u=np.zeros(nfl), v=np.zeros(nfl), w=np.zeros(nfl),
              h=np.full(nfl, h), m=np.full(nfl, m), rho=np.full(nfl, rho0), p=np.zeros(nfl),
              cs=np.full(nfl, c0))
    wl = dict(x=xw, y=yw, z=zw, u=np.zeros(nw), v=np.zeros(nw), w=np.zeros(nw),
              h=np.full(nw, h), m=np.full(nw, m), rho=np.full(nw, rho0), p=np.zeros(nw),
              cs=np.full(nw, c0))
    params = dict(dx=dx, h=h, rho0=rho0, c0=c0, gamma=7.0, alpha=0.1, beta=0.0, eta=0.1,
                  g=-g, dt=0.25 * h / c0)
    return fl, wl, params


def _line_frame(angle):
    """Tangent and normal of a straight wall at ``angle`` to +x
    (base/geometry.pyx:62-88)."""
    c, s = np.cos(angle), np.sin(angle)
    return np.array([c, s, 0.0]), np.array([-s, c, 0.0])


def monaghan_walls_2d(lines, dx):
    """Single-row wall particles with tangent/normal frames for Monaghan's
    boundary force, laid out as base/geometry.pyx meshes an open chain of
    straight lines (``mesh_line`` :115-133, corners of ``mesh_geometry``
    :282-351, ``get_particle_array`` :353-405): each line ``(xa, length,
    angle)`` gets the points xa + k*dx*t strictly inside it with frame
    (-t, -n); the joint of consecutive lines gets the normalised mean frame;
    the free ends get their line's frame.  m = 1 and h = 1, as the reference
    sets them."""
    pts, tang, norm = [], [], []
    frames = []
    for xa, length, angle in lines:
        t, n = _line_frame(angle)
        xa = np.array([xa[0], xa[1], 0.0])
        xb = xa + length * t
        frames.append((xa, xb, t, n))
        pos = xa + t * dx
        while (length - (np.linalg.norm(pos - xa) - 1e-10)) > dx:
            pts.append(pos)
            tang.append(-t)
            norm.append(-n)
            pos = pos + t * dx
    nl = len(frames)

    def unit(v):
        return v / np.sqrt(np.dot(v, v))

    for i in range(1, nl - 1):
        xa, xb, t, n = frames[i]
        _, _, tp, np_ = frames[i + 1]
        _, _, tm, nm = frames[i - 1]
        pts.append(xa)
        norm.append(unit((n + np_) * -0.5))
        tang.append(unit((t + tp) * -0.5))
        pts.append(xb)
        norm.append(unit((n + nm) * -0.5))
        tang.append(unit((t + tm) * -0.5))
    xa0, xb0, t0, n0 = frames[0]
    pts.append(xb0)
    norm.append(-n0)
    tang.append(-t0)
    xal, xbl, tl, nl_ = frames[-1]
    pts.append(xal)
    norm.append(-nl_)
    tang.append(-tl)
    P, T, N = np.array(pts), np.array(tang), np.array(norm)
    k = len(P)
    return dict(x=P[:, 0].copy(), y=P[:, 1].copy(), z=np.zeros(k), tx=T[:, 0].copy(),
                ty=T[:, 1].copy(), tz=np.zeros(k), nx=N[:, 0].copy(), ny=N[:, 1].copy(),
                nz=np.zeros(k), m=np.ones(k), h=np.ones(k))


def dam_break_2d_monaghan(h=0.05, ro=1000.0, co=65.0, width=4.0, height=3.0,
                          column=(1.0, 2.0)):
    """The tiny 2D dam break of examples/dam-break/dam_break_simple.py: a
    column of fluid at (dx..w, dx..h) with dx = h/1.25, m = dx^2 ro, and an open
    box (left wall, floor, right wall) of single-row Monaghan wall particles
    at spacing dx.  Returns (fluid, walls, dx); the walls' h is set to the
    fluid's (the reference leaves it at 1, which only sets its cell size)."""
    dx = h / 1.25
    xarr = np.arange(dx, column[0] + dx, dx)
    yarr = np.arange(dx, column[1] + dx, dx)
    x, y = np.meshgrid(xarr, yarr)
    x, y = x.ravel(), y.ravel()
    n = len(x)
    fluid = dict(x=x, y=y, z=np.zeros(n), h=np.full(n, h), m=np.full(n, dx * dx * ro),
                 rho=np.full(n, ro), cs=np.full(n, co))
    walls = monaghan_walls_2d([((0.0, 0.0), height, np.pi / 2),
                               ((width, 0.0), width, np.pi),
                               ((width, height), height, 1.5 * np.pi)], dx)
    walls["h"] = np.full(len(walls["x"]), h)
    return fluid, walls, dx
