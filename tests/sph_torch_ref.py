"""A small pure-PyTorch FP64 SPH (test infrastructure).

The same right-hand side as the HIP path -- SPHRho, TaitEquation,
MomentumEquation with Monaghan viscosity (sph/funcs/density_funcs.pyx:37-73,
eos_funcs.pyx:118-129, pressure_funcs.pyx:246-391), cubic spline
(base/kernels.pyx:349-453) -- written with differentiable torch operations
over an explicit pair list, so torch.autograd supplies reference gradients
for the hand-written reverse kernels.  Uniform h.
"""
import math

import numpy as np
import torch


def pairs_bruteforce(x, y, h, z=None):
    """All ordered pairs (i, j) with |x_i - x_j|^2 < (2h)^2, self included."""
    pts = np.stack([x, y] if z is None else [x, y, z], axis=1)
    d2 = ((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1)
    i, j = np.nonzero(d2 < (2.0 * h) ** 2)
    return torch.from_numpy(i), torch.from_numpy(j)


def _fac(h, dim):
    sigma = {1: 2.0 / 3.0, 2: 10.0 / (7.0 * math.pi), 3: 1.0 / math.pi}[dim]
    return sigma / h ** dim


def cubic_w(r, h, dim):
    q = r / h
    w = torch.where(q > 1.0, 0.25 * (2 - q) ** 3, 1 - 1.5 * q * q * (1 - 0.5 * q))
    return torch.where(q > 2.0, torch.zeros_like(q), w) * _fac(h, dim)


def cubic_gfac(r, h, dim):
    """G with gradW_ab = r_ab G (zero at r = 0)."""
    q = r / h
    rs = torch.where(r > 0, r, torch.ones_like(r))
    outer = -0.75 * (2 - q) ** 2 / (h * rs)
    inner = 3.0 * (0.75 * q - 1) / (h * h)
    g = torch.where(q >= 1.0, outer, inner)
    g = torch.where((q > 2.0) | (q <= 1e-14), torch.zeros_like(g), g)
    return g * _fac(h, dim)


def accelerations(pos, vel, m, h, pairs, alpha, beta=0.0, eta=0.1, co=1400.0, ro=1.0,
                  gamma=7.0):
    """a(x, u) of MomentumEquation + AV, densities by SPHRho over `pairs`.
    pos, vel: (n, dim) tensors; alpha may be a tensor (differentiable)."""
    n, dim = pos.shape
    i, j = pairs
    off = i != j
    i, j = i[off], j[off]
    r = pos[i] - pos[j]
    rab = torch.sqrt((r * r).sum(1))  # r > 0 off the diagonal: sqrt stays differentiable
    w0 = cubic_w(torch.zeros(1, dtype=pos.dtype), h, dim)
    rho = (m * w0).index_add(0, i, m[j] * cubic_w(rab, h, dim))
    B = co * co * ro / gamma
    p = ((rho / ro) ** gamma - 1.0) * B
    cs = (rho / ro) ** (0.5 * (gamma - 1.0)) * co
    P = p / (rho * rho)
    v = vel[i] - vel[j]
    dot = (v * r).sum(1)
    mu = h * dot / ((r * r).sum(1) + eta * eta * h * h)
    cm = 0.5 * (cs[i] + cs[j])
    rm = 0.5 * (rho[i] + rho[j])
    pi = torch.where(dot < 0, (-alpha * cm * mu + beta * mu * mu) / rm, torch.zeros_like(mu))
    s = -m[j] * (P[i] + P[j] + pi)
    term = (s * cubic_gfac(rab, h, dim))[:, None] * r
    return torch.zeros_like(pos).index_add(0, i, term), rho


def rk2_rollout(pos, vel, m, h, alpha, nsteps, dt, **phys):
    """RK2 midpoint rollout (solver/integrator.py:467-498) with the pair
    list rebuilt from the positions before every evaluation (brute force,
    not differentiated); returns the final positions."""
    def pairs_of(p):
        d = p.detach()
        d2 = ((d[:, None, :] - d[None, :, :]) ** 2).sum(-1)
        i, j = torch.nonzero(d2 < (2.0 * h) ** 2, as_tuple=True)
        return i, j

    for _ in range(nsteps):
        a0, _ = accelerations(pos, vel, m, h, pairs_of(pos), alpha, **phys)
        p1 = pos + vel * (dt / 2)
        v1 = vel + a0 * (dt / 2)
        a1, _ = accelerations(p1, v1, m, h, pairs_of(p1), alpha, **phys)
        pos = pos + v1 * dt
        vel = vel + a1 * dt
    return pos
