"""A small pure-PyTorch FP64 SPH (test infrastructure).

The same right-hand side as the HIP path -- SPHRho, TaitEquation,
MomentumEquation with Monaghan viscosity (sph/funcs/density_funcs.pyx:37-73,
eos_funcs.pyx:118-129, pressure_funcs.pyx:246-391), cubic spline
(base/kernels.pyx:349-453) -- written with differentiable torch operations
over an explicit pair list, so torch.autograd supplies reference gradients
for the hand-written reverse kernels.  Uniform h.  Runs on the device of its
input tensors.

The loss of a rollout is only piecewise smooth: the viscosity acts only on
pairs with v_ab . r_ab < 0, so the accelerations have a kink where a pair's
v . r changes sign.  ``rk2_rollout`` can record each evaluation's pairs and
viscosity switches, and ``rk2_rollout_frozen`` replays a batch of perturbed
rollouts with those pairs and switches held fixed: the rollout map of that
branch, smooth in the initial state and alpha, whose derivative at the
recorded point is what autograd (and the adjoint) computes.
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


def pairs_of(p, h):
    """Ordered pairs (i, j), i != j, with |x_i - x_j|^2 < (2h)^2 of the
    positions p (n, dim), on p's device, not differentiated."""
    d = p.detach()
    n = d.shape[0]
    out_i, out_j = [], []
    step = max(1, (1 << 26) // max(n, 1))  # rows per block: ~64M distances at a time
    for r0 in range(0, n, step):
        blk = d[r0:r0 + step]
        d2 = ((blk[:, None, :] - d[None, :, :]) ** 2).sum(-1)
        i, j = torch.nonzero(d2 < (2.0 * h) ** 2, as_tuple=True)
        out_i.append(i + r0)
        out_j.append(j)
    i, j = torch.cat(out_i), torch.cat(out_j)
    off = i != j
    return i[off], j[off]


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
                  gamma=7.0, switch=None):
    """a(x, u) of MomentumEquation + AV, densities by SPHRho over `pairs`.
    pos, vel: (..., n, dim) tensors (leading dimensions: independent
    replicas sharing the pairs); alpha may be a tensor (differentiable, or
    one value per replica with shape (..., 1)).  switch: a boolean mask per
    pair replacing the viscosity condition v_ab . r_ab < 0.  Returns (a, rho,
    the condition as evaluated)."""
    dim = pos.shape[-1]
    i, j = pairs
    off = i != j
    i, j = i[off], j[off]
    if switch is not None:
        switch = switch[off] if switch.shape[-1] == off.shape[0] else switch
    r = pos[..., i, :] - pos[..., j, :]
    rab = torch.sqrt((r * r).sum(-1))  # r > 0 off the diagonal: sqrt stays differentiable
    w0 = cubic_w(torch.zeros(1, dtype=pos.dtype, device=pos.device), h, dim)
    rho = (m * w0).expand(pos.shape[:-1]).clone().index_add(
        pos.dim() - 2, i, m[j] * cubic_w(rab, h, dim))
    B = co * co * ro / gamma
    p = ((rho / ro) ** gamma - 1.0) * B
    cs = (rho / ro) ** (0.5 * (gamma - 1.0)) * co
    P = p / (rho * rho)
    v = vel[..., i, :] - vel[..., j, :]
    dot = (v * r).sum(-1)
    mu = h * dot / ((r * r).sum(-1) + eta * eta * h * h)
    cm = 0.5 * (cs[..., i] + cs[..., j])
    rm = 0.5 * (rho[..., i] + rho[..., j])
    cond = dot < 0
    use = cond if switch is None else switch
    pi = torch.where(use, (-alpha * cm * mu + beta * mu * mu) / rm, torch.zeros_like(mu))
    s = -m[j] * (P[..., i] + P[..., j] + pi)
    term = (s * cubic_gfac(rab, h, dim))[..., None] * r
    return torch.zeros_like(pos).index_add(pos.dim() - 2, i, term), rho, cond


def rk2_rollout(pos, vel, m, h, alpha, nsteps, dt, record=None, **phys):
    """RK2 midpoint rollout (solver/integrator.py:467-498) with the pair
    list rebuilt from the positions before every evaluation (brute force,
    not differentiated); returns the final positions.  record: a list that
    receives (pairs, viscosity condition) of every evaluation."""
    for _ in range(nsteps):
        pr = pairs_of(pos, h)
        a0, _, c0 = accelerations(pos, vel, m, h, pr, alpha, **phys)
        if record is not None:
            record.append((pr, c0.detach()))
        p1 = pos + vel * (dt / 2)
        v1 = vel + a0 * (dt / 2)
        pr = pairs_of(p1, h)
        a1, _, c1 = accelerations(p1, v1, m, h, pr, alpha, **phys)
        if record is not None:
            record.append((pr, c1.detach()))
        pos = pos + v1 * dt
        vel = vel + a1 * dt
    return pos


def rk2_rollout_frozen(pos, vel, m, h, alpha, dt, record, **phys):
    """The rollout of ``rk2_rollout`` over the pairs and viscosity switches of
    `record` (one entry per evaluation), for a batch of replicas: pos, vel of
    shape (B, n, dim), alpha a (B, 1) tensor.  Returns the final positions."""
    assert len(record) % 2 == 0
    for k in range(0, len(record), 2):
        pr, sw = record[k]
        a0, _, _ = accelerations(pos, vel, m, h, pr, alpha, switch=sw, **phys)
        p1 = pos + vel * (dt / 2)
        v1 = vel + a0 * (dt / 2)
        pr, sw = record[k + 1]
        a1, _, _ = accelerations(p1, v1, m, h, pr, alpha, switch=sw, **phys)
        pos = pos + v1 * dt
        vel = vel + a1 * dt
    return pos
