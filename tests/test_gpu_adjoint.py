"""Device adjoint (demon_fit_loss_grad) against the oracle's CPU reference
gradient and against central finite differences of the device loss.

Bars (north star): within 1e-6 relative of central finite differences
(h = 1e-6 relative) -- or within the difference quotient's own rounding
error, 8 eps |L| / h, for parameters whose gradient is that small -- and
equal to the FP64 CPU reference gradient (here to 1e-8 relative).
"""
import numpy as np
import pytest

import demon_simulation_code_amd as dm
from tests.helpers import CONFINE, GAMMA, KAPPA, M0, Q0, lattice

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def _problem(n, steps, obs, seed=3, jitter_q=0.1):
    x, y = lattice(n, seed=seed)
    rng = np.random.default_rng(seed)
    q_true = Q0 * (1 + jitter_q * rng.uniform(-1, 1, n))
    m = np.full(n, M0)
    # targets: trajectory of the true parameters
    with dm.Context(n, shielding=KAPPA, confine=CONFINE, gamma=GAMMA, dt_check=False) as c:
        c.upload(x, y, None, None, q_true, m)
        tx, ty = [], []
        t = 0.0
        for _ in range(steps // obs):
            t = c.rk4_steps(obs, 1e-3, t)
            gx, gy, _, _ = c.download()
            tx.append(gx)
            ty.append(gy)
    q0 = Q0 * (1 + 0.1 * np.random.default_rng(seed + 1).uniform(-1, 1, n))
    return x, y, m, q0, np.array(tx), np.array(ty)


def _ctx(n, x, y, m, q, kappa=KAPPA * 1.02, confine=CONFINE * 0.97, gamma=GAMMA * 1.05):
    c = dm.Context(n, shielding=kappa, confine=confine, gamma=gamma, dt_check=False)
    c.upload(x, y, None, None, q, m)
    return c


@pytest.mark.parametrize("ckpt", [20, 7, 40])
def test_adjoint_matches_cpu_reference_gradient(oracle, ckpt):
    n, steps, obs = 256, 40, 10
    x, y, m, q, tx, ty = _problem(n, steps, obs)
    kappa, conf, gam = KAPPA * 1.02, CONFINE * 0.97, GAMMA * 1.05
    with _ctx(n, x, y, m, q) as c:
        loss, g = c.fit_loss_grad(steps, 1e-3, obs, ckpt, tx, ty, state_grad=True)
    z = np.zeros(n)
    L, gq, gk, gc, gg, gs = oracle.fit_loss_grad(x, y, z, z, q, m, kappa, conf, gam, 1e-3, steps, obs,
                                                  tx, ty, state_grad=True)
    assert loss == pytest.approx(L, rel=1e-9)
    assert np.max(np.abs(g["q"] - gq)) <= 1e-8 * np.max(np.abs(gq))
    assert g["kappa"] == pytest.approx(gk, rel=1e-8)
    assert g["confine"] == pytest.approx(gc, rel=1e-8)
    assert g["gamma"] == pytest.approx(gg, rel=1e-8)
    for name, ref in zip(("x0", "y0", "vx0", "vy0"), gs):
        assert np.max(np.abs(g[name] - ref)) <= 1e-8 * np.max(np.abs(ref)), name


def test_loss_call_restores_state_and_is_deterministic():
    n, steps, obs = 512, 30, 10
    x, y, m, q, tx, ty = _problem(n, steps, obs, seed=5)
    with _ctx(n, x, y, m, q) as c:
        l1, g1 = c.fit_loss_grad(steps, 1e-3, obs, 20, tx, ty)
        s = c.download()
        l2, g2 = c.fit_loss_grad(steps, 1e-3, obs, 20, tx, ty)
        l3, _ = c.fit_loss_grad(steps, 1e-3, obs, 20, tx, ty, grad=False)
    assert np.array_equal(s[0], x) and np.array_equal(s[1], y)
    assert l1 == l2 == l3
    assert np.array_equal(g1["q"], g2["q"]) and g1["kappa"] == g2["kappa"]


def test_adjoint_vs_central_differences_config5():
    """N = 4096, 200 steps, observations every 10th step, checkpoints every 20:
    32 random parameters (charges and the three globals) by central differences."""
    n, steps, obs, ckpt = 4096, 200, 10, 20
    x, y, m, q, tx, ty = _problem(n, steps, obs, seed=11)
    kappa, conf, gam = KAPPA * 1.02, CONFINE * 0.97, GAMMA * 1.05
    with _ctx(n, x, y, m, q, kappa, conf, gam) as c:
        L, g = c.fit_loss_grad(steps, 1e-3, obs, ckpt, tx, ty)
        fwd_ms, bwd_ms = c.fit_timing()

        def loss_at(**kw):
            qq = kw.pop("q", q)
            c.set_params(shielding=kw.get("kappa", kappa), confine=kw.get("confine", conf),
                         drag_gamma=kw.get("gamma", gam))
            c.upload(x, y, None, None, qq, m)
            return c.fit_loss_grad(steps, 1e-3, obs, ckpt, tx, ty, grad=False)[0]

        rng = np.random.default_rng(0)
        checks = [("kappa", kappa), ("confine", conf), ("gamma", gam)]
        checks += [("q", int(i)) for i in rng.choice(n, 29, replace=False)]
        worst = 0.0
        for name, v in checks:
            if name == "q":
                h = abs(q[v]) * 1e-6
                qp, qm = q.copy(), q.copy()
                qp[v] += h
                qm[v] -= h
                fd = (loss_at(q=qp) - loss_at(q=qm)) / (2 * h)
                ad = g["q"][v]
            else:
                h = abs(v) * 1e-6
                fd = (loss_at(**{name: v + h}) - loss_at(**{name: v - h})) / (2 * h)
                ad = g[name]
            bound = 1e-6 * abs(ad) + 8 * EPS * abs(L) / h
            worst = max(worst, abs(fd - ad) / bound)
            assert abs(fd - ad) <= bound, (name, v, fd, ad)
    print(f"fit N={n}: forward {fwd_ms:.1f} ms, recompute+adjoint {bwd_ms:.1f} ms, "
          f"ratio {(fwd_ms + bwd_ms) / fwd_ms:.2f}; worst FD error / bound {worst:.3f}")


def test_torch_autograd_binding():
    torch = pytest.importorskip("torch")
    from demon_simulation_code_amd.autograd import TrajectoryFit

    n, steps, obs = 256, 20, 10
    x, y, m, q, tx, ty = _problem(n, steps, obs, seed=7)
    fit = TrajectoryFit(x, y, m, tx, ty, steps=steps, dt=1e-3, obs_every=obs, ckpt_every=10)
    qt = torch.tensor(q, dtype=torch.float64, requires_grad=True)
    kt = torch.tensor(KAPPA * 1.02, dtype=torch.float64, requires_grad=True)
    ct = torch.tensor(CONFINE * 0.97, dtype=torch.float64, requires_grad=True)
    gt = torch.tensor(GAMMA * 1.05, dtype=torch.float64, requires_grad=True)
    loss = fit(qt, kt, ct, gt)
    (2.0 * loss).backward()
    with _ctx(n, x, y, m, q) as c:
        L, g = c.fit_loss_grad(steps, 1e-3, obs, 10, tx, ty)
    fit.close()
    assert loss.item() == L
    assert np.allclose(qt.grad.numpy(), 2 * g["q"], rtol=0, atol=0)
    assert kt.grad.item() == 2 * g["kappa"] and gt.grad.item() == 2 * g["gamma"]
