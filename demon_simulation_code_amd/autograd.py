"""torch.autograd binding of the device trajectory fit (demon_fit_loss_grad).

    fit = TrajectoryFit(x0, y0, mass, targets_x, targets_y, steps=200, dt=1e-3,
                        obs_every=10, ckpt_every=20)
    q = torch.tensor(q_init, dtype=torch.float64, requires_grad=True)
    kappa = torch.tensor(2000.0, dtype=torch.float64, requires_grad=True)
    ...
    loss = fit(q, kappa, confine, gamma)   # forward sweep on the GPU
    loss.backward()                        # recompute + adjoint sweep on the GPU

The physics parameters are the reference's: per-particle charge q [C],
shielding kappa = 1/lambda [1/m] (ShieldedCoulombForce), confinement C [V/m^2]
with F = q C r (ConfinementForce) and drag gamma [1/s] (DragForce).  All the
arithmetic runs in libdemon_gpu.so; this module only moves the parameters and
gradients between torch tensors and the C ABI.
"""
from __future__ import annotations

import numpy as np
import torch

from ._lib import CONFINEMENT, DRAG, SHIELDED_COULOMB, Context


class _FitLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q, kappa, confine, gamma, fit):
        loss, grads = fit._evaluate(q, kappa, confine, gamma, need_grad=any(ctx.needs_input_grad[:4]))
        ctx.grads = grads
        ctx.like = (q, kappa, confine, gamma)
        return torch.tensor(loss, dtype=torch.float64, device=q.device)

    @staticmethod
    def backward(ctx, g):
        q, kappa, confine, gamma = ctx.like
        gr = ctx.grads
        scale = g.detach().to("cpu", torch.float64).item()
        out = []
        for name, like in (("q", q), ("kappa", kappa), ("confine", confine), ("gamma", gamma)):
            v = gr.get(name)
            if v is None:
                out.append(None)
            else:
                out.append((torch.as_tensor(np.asarray(v) * scale, dtype=like.dtype)).to(like.device).reshape(like.shape))
        return (*out, None)


class TrajectoryFit:
    """Loss L(q, kappa, C, gamma) of a trajectory against targets at every
    obs_every-th step, L = sum |r - r*|^2 / (2 n n_obs), on one GPU.

    dt_check (default True) keeps the reference integrator's adaptive time
    step: every step's dt comes from Integrator::modifyTimeStep
    (Integrator.cpp:68-155), as in Runge_Kutta4::moveParticles
    (Runge_Kutta4.cpp:25), and the adjoint replays the dt of each step.
    dt_check=False steps with the fixed dt instead, which differs from the
    reference whenever a tested pair comes within 100 um."""

    def __init__(self, x0, y0, mass, targets_x, targets_y, *, steps, dt, obs_every, ckpt_every,
                 vx0=None, vy0=None, device: int = 0, dt_check: bool = True):
        self.n = len(x0)
        self.x0 = np.ascontiguousarray(x0, dtype=np.float64)
        self.y0 = np.ascontiguousarray(y0, dtype=np.float64)
        self.vx0 = None if vx0 is None else np.ascontiguousarray(vx0, dtype=np.float64)
        self.vy0 = None if vy0 is None else np.ascontiguousarray(vy0, dtype=np.float64)
        self.mass = np.ascontiguousarray(mass, dtype=np.float64)
        self.tx = np.ascontiguousarray(targets_x, dtype=np.float64)
        self.ty = np.ascontiguousarray(targets_y, dtype=np.float64)
        self.steps, self.dt, self.obs_every, self.ckpt_every = int(steps), float(dt), int(obs_every), int(ckpt_every)
        self.sim = Context(self.n, shielding=1.0, forces=CONFINEMENT | DRAG | SHIELDED_COULOMB,
                           dt_check=bool(dt_check), device=device)

    def close(self):
        self.sim.close()

    def _evaluate(self, q, kappa, confine, gamma, need_grad: bool):
        qn = q.detach().to("cpu", torch.float64).numpy().copy()
        self.sim.set_params(shielding=float(kappa), confine=float(confine), drag_gamma=float(gamma))
        self.sim.upload(self.x0, self.y0, self.vx0, self.vy0, qn, self.mass)
        return self.sim.fit_loss_grad(self.steps, self.dt, self.obs_every, self.ckpt_every, self.tx,
                                      self.ty, grad=need_grad)

    def __call__(self, q, kappa, confine, gamma):
        return _FitLoss.apply(q, kappa, confine, gamma, self)

    def timing(self):
        """(forward_ms, backward_ms) of the last evaluation."""
        return self.sim.fit_timing()
