"""Closed forms of the momentum equation for a handful of particles (test
infrastructure).

MomentumEquation with Monaghan artificial viscosity
(sph/funcs/pressure_funcs.pyx:246-391) written from its definition, all pairs
within the support, independently of the oracle's expression order:

    a_a = sum_b -m_b (P_a/rho_a^2 + P_b/rho_b^2 + Pi_ab) grad_a W(r_ab, h_ab)
    Pi_ab = (-alpha cbar mu + beta mu^2) / rhobar   when v_ab . r_ab < 0, else 0
    mu = h_ab v_ab . r_ab / (|r_ab|^2 + eta^2 h_ab^2),  h_ab = (h_a + h_b)/2
    cbar = (c_a + c_b)/2,  rhobar = (rho_a + rho_b)/2
    dt_fac_a = max_b |mu_ab|

The kernel gradient is the derivative of the cubic spline W = sigma_d/h^d f(q),
f(q) = 1 - 3/2 q^2 + 3/4 q^3 (q <= 1), 1/4 (2 - q)^3 (1 < q <= 2), 0 beyond:
grad_a W = sigma_d / h^(d+1) f'(q) r_ab / |r_ab| (base/kernels.pyx:349-417).
"""
import math

import numpy as np

SIGMA = {1: 2.0 / 3.0, 2: 10.0 / (7.0 * math.pi), 3: 1.0 / math.pi}


def dfdq(q):
    """f'(q) of the cubic spline's shape function."""
    if q <= 1.0:
        return -3.0 * q + 2.25 * q * q
    if q <= 2.0:
        return -0.75 * (2.0 - q) ** 2
    return 0.0


def grad_w(rab, h, dim):
    """grad_a W(r_ab, h) as a 3-vector (r_ab = x_a - x_b)."""
    r = float(np.sqrt(np.dot(rab, rab)))
    if r == 0.0:
        return np.zeros(3)
    return SIGMA[dim] / h ** (dim + 1) * dfdq(r / h) * np.asarray(rab) / r


def momentum(dim, pos, vel, m, rho, p, cs, h, alpha, beta, eta=0.1,
             pressure_term=None, pi_term=None):
    """(accelerations (n, 3), dt_fac (n,)) of every particle from every other
    one within 2 h_ab.  pressure_term / pi_term replace the pressure and
    viscosity terms (to show that a plausible mistake changes the result)."""
    n = len(m)
    acc = np.zeros((n, 3))
    dtf = np.zeros(n)
    for a in range(n):
        for b in range(n):
            rab = pos[a] - pos[b]
            vab = vel[a] - vel[b]
            hab = 0.5 * (h[a] + h[b])
            r2 = float(np.dot(rab, rab))
            if r2 >= (2.0 * h[a]) ** 2:  # outside the destination's support: not a neighbour
                continue
            dot = float(np.dot(vab, rab))
            mu = hab * dot / (r2 + eta * eta * hab * hab)
            dtf[a] = max(dtf[a], abs(mu))
            pi = 0.0
            if dot < 0.0:
                if pi_term is not None:
                    pi = pi_term(alpha, beta, mu, cs[a], cs[b], rho[a], rho[b])
                else:
                    pi = (-alpha * 0.5 * (cs[a] + cs[b]) * mu + beta * mu * mu) / (
                        0.5 * (rho[a] + rho[b]))
            if pressure_term is not None:
                pt = pressure_term(p[a], p[b], rho[a], rho[b])
            else:
                pt = p[a] / rho[a] ** 2 + p[b] / rho[b] ** 2
            acc[a] += -m[b] * (pt + pi) * grad_w(rab, hab, dim)
    return acc, dtf


def two_particles(approach=True, far=False):
    """Two particles of unequal h, mass, density, pressure and sound speed.
    far=False puts them on the inner branch of the spline (q = 0.66), True on
    the outer one (q = 1.36).  Returns a dict of 2-element arrays."""
    b = np.array([1.2, 0.5, -0.4]) if far else np.array([0.55, -0.3, 0.2])
    pos = np.array([[0.0, 0.0, 0.0], b])
    vel = np.array([[0.5, 0.1, -0.2], [-1.5, 0.4, 0.3]])
    if not approach:
        vel = -vel
    return dict(pos=pos, vel=vel, m=np.array([0.9, 1.4]), rho=np.array([1.2, 0.7]),
                p=np.array([3.0, -0.8]), cs=np.array([2.0, 6.0]), h=np.array([0.9, 1.1]))


def unit_square():
    """The reference's four-particle fixture (sph/tests/
    function_test_template.py:45-58 with get_particle_array's defaults
    h = 1, rho = 1, cs = 0)."""
    return dict(pos=np.array([[0.0, 0.0, 0.0], [0.0, 1.0, 0.0], [1.0, 1.0, 0.0],
                              [1.0, 0.0, 0.0]]),
                vel=np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0],
                              [-1.0, 0.0, 0.0]]),
                m=np.ones(4), rho=np.ones(4), p=np.array([0.0, 0.0, 1.0, 1.0]),
                cs=np.zeros(4), h=np.ones(4))


def as_arrays(c):
    """The case as a dict of per-property arrays (x y z u v w m rho p cs h)."""
    out = {k: c["pos"][:, i].copy() for i, k in enumerate("xyz")}
    out.update({k: c["vel"][:, i].copy() for i, k in enumerate("uvw")})
    for k in ("m", "rho", "p", "cs", "h"):
        out[k] = np.asarray(c[k], dtype=np.float64).copy()
    return out
