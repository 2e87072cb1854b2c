"""GPU checks of the reverse mode: the hand-written vector-Jacobian kernels
against torch.autograd of a pure-PyTorch FP64 SPH (tests/sph_torch_ref.py),
and the rollout gradient against central finite differences."""
import numpy as np
import pytest
import torch

from tests import sph_torch_ref as ref

pytestmark = pytest.mark.gpu


def _drop(dx, radius, vel=100.0, seed=1, noise=5.0):
    from pysph_amd.solver.generators import elliptical_drop
    d = elliptical_drop(dx=dx, radius=radius, vel=vel)
    rng = np.random.default_rng(seed)
    d["u"] = d["u"] + rng.normal(0.0, noise, len(d["x"]))
    d["v"] = d["v"] + rng.normal(0.0, noise, len(d["x"]))
    return d


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_vjp_matches_torch_autograd(gpu):
    import ctypes
    from pysph_amd.batch import BatchedSimulation
    d = _drop(0.08, 0.6, seed=3)
    n = len(d["x"])
    h = float(d["h"][0])
    sim = BatchedSimulation([d], alpha=0.3, beta=0.2)
    sim.ctx.check(sim.ctx.lib.sph_batch_eval(sim.ctx.handle, ctypes.byref(sim.desc)))
    rng = np.random.default_rng(4)
    abar = torch.tensor(rng.normal(0.0, 1.0, (n, 2)), device=sim.device)
    outs = [torch.empty(n, dtype=torch.float64, device=sim.device) for _ in range(4)]
    al = torch.zeros(1, dtype=torch.float64, device=sim.device)
    a_u, a_v = abar[:, 0].contiguous(), abar[:, 1].contiguous()
    sim.stream.wait_stream(torch.cuda.current_stream())
    sim.ctx.check(sim.ctx.lib.sph_batch_vjp(
        sim.ctx.handle, ctypes.byref(sim.desc), a_u.data_ptr(), a_v.data_ptr(), None,
        outs[0].data_ptr(), outs[1].data_ptr(), None, outs[2].data_ptr(), outs[3].data_ptr(),
        None, al.data_ptr()))
    sim.check()
    pos = torch.tensor(np.stack([d["x"], d["y"]], 1), requires_grad=True)
    vel = torch.tensor(np.stack([d["u"], d["v"]], 1), requires_grad=True)
    alpha = torch.tensor(0.3, dtype=torch.float64, requires_grad=True)
    acc, _, _ = ref.accelerations(pos, vel, torch.tensor(d["m"]), h,
                               ref.pairs_bruteforce(d["x"], d["y"], h), alpha, beta=0.2)
    gx, gu, ga = torch.autograd.grad(acc, (pos, vel, alpha), grad_outputs=abar.cpu())
    errs = [_rel(outs[0].cpu(), gx[:, 0]), _rel(outs[1].cpu(), gx[:, 1]),
            _rel(outs[2].cpu(), gu[:, 0]), _rel(outs[3].cpu(), gu[:, 1]),
            _rel(al.cpu(), ga.reshape(1))]
    print(f"VJP vs torch autograd ({n} particles): rel err x y u v alpha "
          + " ".join(f"{e:.2g}" for e in errs))
    assert max(errs) <= 1e-9


def test_rollout_gradient_vs_torch(gpu):
    from pysph_amd.adjoint import Rollout, rollout
    d = _drop(0.1, 0.55, seed=5)
    n = len(d["x"])
    h, dt, steps = float(d["h"][0]), 1e-5, 6
    pos0 = np.stack([d["x"], d["y"]], 1)
    roll = Rollout(pos0, d["m"], h, nsteps=steps, dt=dt, checkpoint=4)
    dev = roll.sim.device
    u0 = torch.tensor(d["u"], device=dev, requires_grad=True)
    v0 = torch.tensor(d["v"], device=dev, requires_grad=True)
    alpha = torch.tensor(0.1, dtype=torch.float64, device=dev, requires_grad=True)
    x, y = rollout(roll, alpha, u0, v0)
    wx = torch.linspace(-1, 1, n, dtype=torch.float64, device=dev)
    loss = (wx * x).sum() + (wx.flip(0) * y).sum()
    gu, gv, ga = torch.autograd.grad(loss, (u0, v0, alpha))
    # pure-PyTorch reference on the CPU
    pos = torch.tensor(pos0)
    vel = torch.tensor(np.stack([d["u"], d["v"]], 1), requires_grad=True)
    alr = torch.tensor(0.1, dtype=torch.float64, requires_grad=True)
    pf = ref.rk2_rollout(pos, vel, torch.tensor(d["m"]), h, alr, steps, dt)
    wxc = wx.cpu()
    lr = (wxc * pf[:, 0]).sum() + (wxc.flip(0) * pf[:, 1]).sum()
    rg, ra = torch.autograd.grad(lr, (vel, alr))
    errs = [_rel(gu.cpu(), rg[:, 0]), _rel(gv.cpu(), rg[:, 1]), _rel(ga.cpu(), ra)]
    pos_err = max(_rel(x.detach().cpu(), pf[:, 0].detach()), _rel(y.detach().cpu(), pf[:, 1].detach()))
    print(f"rollout gradient vs torch ({n} particles, {steps} steps): rel err u v alpha "
          + " ".join(f"{e:.2g}" for e in errs) + f"; positions {pos_err:.2g}")
    assert pos_err <= 1e-12
    assert max(errs) <= 1e-8


def test_rollout_gradient_vs_finite_differences(gpu):
    from pysph_amd.adjoint import Rollout, rollout
    d = _drop(0.05, 0.5, seed=7)
    n = len(d["x"])
    h, dt, steps, eps = float(d["h"][0]), 1e-5, 20, 1e-6
    pos0 = np.stack([d["x"], d["y"]], 1)
    roll = Rollout(pos0, d["m"], h, nsteps=steps, dt=dt, checkpoint=5)
    dev = roll.sim.device
    target = torch.tensor(pos0 * 0.99, device=dev)

    def loss_of(u, v, a):
        x, y = roll.forward(u, v, a)
        return float((((x - target[:, 0]) ** 2 + (y - target[:, 1]) ** 2).mean()).item())

    u0 = torch.tensor(d["u"], device=dev, requires_grad=True)
    v0 = torch.tensor(d["v"], device=dev, requires_grad=True)
    alpha = torch.tensor(0.1, dtype=torch.float64, device=dev, requires_grad=True)
    x, y = rollout(roll, alpha, u0, v0)
    loss = ((x - target[:, 0]) ** 2 + (y - target[:, 1]) ** 2).mean()
    gu, gv, ga = torch.autograd.grad(loss, (u0, v0, alpha))
    rng = np.random.default_rng(8)
    picks = rng.choice(n, 6, replace=False)
    worst = 0.0
    base_u, base_v = d["u"].copy(), d["v"].copy()
    for i in picks:
        for comp, g in (("u", gu), ("v", gv)):
            up, dn = base_u.copy(), base_u.copy()
            vp, vd = base_v.copy(), base_v.copy()
            if comp == "u":
                up[i] += eps
                dn[i] -= eps
            else:
                vp[i] += eps
                vd[i] -= eps
            lp = loss_of(torch.tensor(up, device=dev), torch.tensor(vp, device=dev), 0.1)
            lm = loss_of(torch.tensor(dn, device=dev), torch.tensor(vd, device=dev), 0.1)
            fd = (lp - lm) / (2 * eps)
            e = abs(float(g[i]) - fd) / max(abs(fd), 1e-300)
            worst = max(worst, e)
    lp = loss_of(torch.tensor(base_u, device=dev), torch.tensor(base_v, device=dev), 0.1 + eps)
    lm = loss_of(torch.tensor(base_u, device=dev), torch.tensor(base_v, device=dev), 0.1 - eps)
    fda = (lp - lm) / (2 * eps)
    ea = abs(float(ga) - fda) / abs(fda)
    print(f"rollout gradient vs central FD (eps {eps}, {len(picks)} particles x 2, {steps} "
          f"steps): max rel err velocities {worst:.2g}, alpha {ea:.2g}")
    assert worst <= 1e-4 and ea <= 1e-4
