"""Reverse mode of a simulation rollout, for fitting initial velocities or the
artificial viscosity to a target end state.

``Rollout`` advances one simulation (the same kernels as the batched path:
SPHRho, Tait, MomentumEquation + Monaghan viscosity, PositionStepping, RK2
of solver/integrator.py:467-498) for ``nsteps`` steps, keeping the positions
and velocities every ``checkpoint`` steps.  ``Rollout.backward`` walks the
segments in reverse: it recomputes a segment from its checkpoint, keeping
each evaluation's neighbour lists and sorted records, and then applies the
hand-written vector-Jacobian kernels (``sph_batch_vjp``) over those frozen
lists, stage by stage.  For RK2, X1 = X0 + dt/2 F(X0), X2 = X0 + dt F(X1):

    X1bar = J_F(X1)^T (dt X2bar);  X0bar = X2bar + X1bar + J_F(X0)^T (dt/2 X1bar)

``rollout`` wraps this in a ``torch.autograd.Function`` so that the final
positions are differentiable with respect to the initial velocities and
alpha.  The stage arithmetic x0 + a*dt is differentiated as exact; the
neighbour lists are piecewise constant in the positions (the kernel and its
gradient vanish at the support radius), so they are held fixed.
"""
import ctypes

import numpy as np
import torch

from .batch import BatchedSimulation


class Rollout:
    """One simulation of n particles (2D or 3D) with fixed physics except
    alpha; initial positions fixed, initial velocities and alpha free."""

    def __init__(self, pos, m, h, nsteps, dt, checkpoint=20, dim=2, co=1400.0, ro=1.0,
                 gamma=7.0, beta=0.0, eta=0.1, device=None):
        pos = np.asarray(pos, dtype=np.float64)
        n = pos.shape[0]
        self.n, self.dim = n, dim
        self.nsteps, self.dt = int(nsteps), float(dt)
        self.checkpoint = max(1, int(checkpoint))
        d = dict(x=pos[:, 0], y=pos[:, 1], z=pos[:, 2] if dim == 3 else np.zeros(n),
                 h=np.broadcast_to(np.asarray(h, dtype=np.float64), (n,)).copy(),
                 m=np.broadcast_to(np.asarray(m, dtype=np.float64), (n,)).copy(),
                 rho=np.full(n, ro), p=np.zeros(n), cs=np.full(n, co))
        if np.ptp(d["h"]) != 0.0:
            raise NotImplementedError("the reverse mode needs one h for all particles")
        self.sim = BatchedSimulation([d], dim=dim, co=co, ro=ro, gamma=gamma, alpha=0.0,
                                     beta=beta, eta=eta, device=device)
        s = self.sim
        with torch.cuda.stream(s.stream):
            self.pos0 = {k: s.t[k].clone() for k in ("x", "y", "z")}
            # one workspace per evaluation of a segment (lists + records kept)
            nb = s.work.numel()
            self.seg_work = [torch.empty(nb, dtype=torch.uint8, device=s.device)
                             for _ in range(2 * self.checkpoint)]
            self.alpha_bar = torch.zeros(1, dtype=torch.float64, device=s.device)
        self.main_work = (s.desc.work, s.desc.work_bytes)
        for wk in self.seg_work:
            self._use(wk)
            self._call("sph_batch_init_workspace")
        self._use(None)
        self.checkpoints = []
        self.alpha = 0.0

    # -- pieces ---------------------------------------------------------
    def _call(self, name, *args):
        s = self.sim
        s.ctx.check(getattr(s.ctx.lib, name)(s.ctx.handle, ctypes.byref(s.desc), *args))

    def _use(self, work):
        if work is None:
            self.sim.desc.work, self.sim.desc.work_bytes = self.main_work
        else:
            self.sim.desc.work, self.sim.desc.work_bytes = work.data_ptr(), work.numel()

    def _state_keys(self):
        return ("x", "y", "z", "u", "v", "w") if self.dim == 3 else ("x", "y", "u", "v")

    def _set_initial(self, u0, v0, w0, alpha):
        s = self.sim
        with torch.cuda.stream(s.stream):
            for k in ("x", "y", "z"):
                s.t[k].copy_(self.pos0[k])
            s.t["u"].copy_(u0)
            s.t["v"].copy_(v0)
            if w0 is not None:
                s.t["w"].copy_(w0)
        self.alpha = float(alpha)
        s.desc.alpha = self.alpha

    def _step(self, works=None):
        """One RK2 step; evaluation k of it uses works[k] when given."""
        self._call("sph_batch_save_initial")
        self._use(works[0] if works else None)
        self._call("sph_batch_eval")
        self._call("sph_batch_stage", self.dt / 2)
        self._use(works[1] if works else None)
        self._call("sph_batch_eval")
        self._call("sph_batch_stage", self.dt)
        self._use(None)

    # -- forward / backward ---------------------------------------------
    def forward(self, u0, v0, alpha, w0=None):
        """Final positions (device tensors) from initial velocities u0, v0
        (w0 in 3D) and alpha; keeps the checkpoints for backward."""
        s = self.sim
        s.stream.wait_stream(torch.cuda.current_stream(s.device))
        self._set_initial(u0, v0, w0, alpha)
        self.checkpoints = []
        keys = self._state_keys()
        for k in range(self.nsteps):
            if k % self.checkpoint == 0:
                with torch.cuda.stream(s.stream):
                    self.checkpoints.append({q: s.t[q].clone() for q in keys})
            self._step()
        s.check()
        with torch.cuda.stream(s.stream):
            out = tuple(s.t[q].clone() for q in keys[:self.dim])
        s.synchronize()
        return out

    def backward(self, *pos_bar):
        """Gradients of sum(pos_bar . final positions) with respect to the
        initial velocities (one tensor per component) and alpha."""
        s = self.sim
        s.stream.wait_stream(torch.cuda.current_stream(s.device))
        keys = self._state_keys()
        d = self.dim
        dt = self.dt
        vel = keys[d:]
        with torch.cuda.stream(s.stream):  # every tensor op below is ordered with the kernels
            bar = {q: torch.zeros(self.n, dtype=torch.float64, device=s.device) for q in keys}
            for c, q in enumerate(keys[:d]):
                bar[q].copy_(pos_bar[c])
            out = {q: torch.empty_like(bar[q]) for q in keys}
            alpha_total = torch.zeros(1, dtype=torch.float64, device=s.device)
            for seg in reversed(range(len(self.checkpoints))):
                first = seg * self.checkpoint
                nseg = min(self.checkpoint, self.nsteps - first)
                for q in keys:
                    s.t[q].copy_(self.checkpoints[seg][q])
                for j in range(nseg):
                    self._step(self.seg_work[2 * j:2 * j + 2])
                for j in reversed(range(nseg)):
                    # X1bar = J_F(X1)^T (dt X2bar);  X0bar = X2bar + X1bar + J_F(X0)^T (dt/2 X1bar)
                    x1bar = self._vjp(self.seg_work[2 * j + 1], {q: dt * bar[q] for q in keys},
                                      out, alpha_total)
                    add = self._vjp(self.seg_work[2 * j],
                                    {q: (0.5 * dt) * x1bar[q] for q in keys}, out, alpha_total)
                    bar = {q: bar[q] + x1bar[q] + add[q] for q in keys}
            res = tuple(bar[q].clone() for q in vel) + (alpha_total.clone(),)
        s.check()
        return res

    def _vjp(self, work, fbar, out, alpha_total):
        """J_F(X)^T fbar for the evaluation kept in `work`; F = (u, a)."""
        s = self.sim
        keys = self._state_keys()
        d = self.dim
        vel = keys[d:]
        self._use(work)
        p = lambda t: t.data_ptr() if t is not None else None
        abar = [fbar[q] for q in vel]
        xb = [out[q] for q in keys[:d]]
        ub = [out[q] for q in vel]
        s.ctx.check(s.ctx.lib.sph_batch_vjp(
            s.ctx.handle, ctypes.byref(s.desc), p(abar[0]), p(abar[1]),
            p(abar[2]) if d == 3 else None, p(xb[0]), p(xb[1]), p(xb[2]) if d == 3 else None,
            p(ub[0]), p(ub[1]), p(ub[2]) if d == 3 else None, p(self.alpha_bar)))
        self._use(None)
        with torch.cuda.stream(s.stream):
            res = {}
            for c, q in enumerate(keys[:d]):
                res[q] = out[q].clone()
            for c, q in enumerate(vel):
                # dx/dt = u: the position rows of fbar go to the velocities
                res[q] = out[q] + fbar[keys[c]]
            alpha_total += self.alpha_bar
        return res


class _RolloutFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, roll, alpha, *vel0):
        out = roll.forward(*[v.detach() for v in vel0[:2]], float(alpha),
                           w0=vel0[2].detach() if len(vel0) > 2 else None)
        ctx.roll = roll
        return out

    @staticmethod
    def backward(ctx, *pos_bar):
        pos_bar = [g if g is not None else torch.zeros(ctx.roll.n, dtype=torch.float64,
                                                       device=ctx.roll.sim.device)
                   for g in pos_bar]
        res = ctx.roll.backward(*pos_bar)
        vel_bar, alpha_bar = res[:-1], res[-1]
        return (None, alpha_bar.reshape(())) + tuple(vel_bar)


def rollout(roll, alpha, *vel0):
    """Differentiable final positions: rollout(roll, alpha, u0, v0[, w0])."""
    return _RolloutFn.apply(roll, alpha, *vel0)
