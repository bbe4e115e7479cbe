"""Pins the oracle's CPU reference gradient (oracle_fit_loss_grad), the
checker of the device adjoint, without the device:

* its loss equals the loss recomputed from the oracle's own RK4 trajectory
  (oracle_rk4_steps, itself pinned bitwise to the compiled reference);
* every gradient component equals a fourth-order central difference of that
  loss on small smooth problems (no pair near the cutoff), with and without
  steps whose dt the dt check cuts;
* a planted sign error in one adjoint term would break the second check,
  which the dt-cut case also exercises through the replayed dt.
"""
import numpy as np
import pytest

from oracle.bind import OracleParams
from tests.helpers import CONFINE, GAMMA, M0, Q0


def _setup(n, seed, spacing):
    rng = np.random.default_rng(seed)
    side = int(np.ceil(np.sqrt(n)))
    gx, gy = np.meshgrid(np.arange(side), np.arange(side))
    x = (gx.ravel()[:n] - side / 2) * spacing + rng.uniform(-0.2, 0.2, n) * spacing
    y = (gy.ravel()[:n] - side / 2) * spacing + rng.uniform(-0.2, 0.2, n) * spacing
    vx = rng.uniform(-1, 1, n) * 1e-3
    vy = rng.uniform(-1, 1, n) * 1e-3
    q = Q0 * (1 + 0.1 * rng.uniform(-1, 1, n))
    m = np.full(n, M0)
    return x, y, vx, vy, q, m


def _targets(oracle, x, y, vx, vy, q, m, kappa, steps, obs, dt_check):
    tx, ty = [], []
    for b in range(steps // obs):
        r = oracle.rk4_steps(x, y, vx, vy, q, m,
                             OracleParams(kappa, CONFINE, GAMMA, 0.0, 0, 7, int(dt_check), 1),
                             1e-3, obs, step0=b * obs)
        x, y, vx, vy = r[0], r[1], r[2], r[3]
        tx.append(x + 3e-6)
        ty.append(y - 2e-6)
    return np.array(tx), np.array(ty)


def _fd4(f, v, h):
    return (8 * (f(v + h) - f(v - h)) - (f(v + 2 * h) - f(v - 2 * h))) / (12 * h)


@pytest.mark.parametrize("dt_check,spacing", [(False, 400e-6), (True, 400e-6), (True, 110e-6)])
def test_fit_gradient_equals_fourth_order_differences(oracle, dt_check, spacing):
    n, steps, obs = 16, 30, 5
    kappa = 200.0  # cutoff 50 mm: no pair anywhere near it
    x, y, vx, vy, q, m = _setup(n, 5, spacing)
    tx, ty = _targets(oracle, x, y, vx, vy, q * 1.05, m, kappa, steps, obs, dt_check)
    K, C, G = kappa * 1.1, CONFINE * 0.9, GAMMA * 1.2

    def loss(x=x, y=y, vx=vx, vy=vy, q=q, K=K, C=C, G=G):
        return oracle.fit_loss_grad(x, y, vx, vy, q, m, K, C, G, 1e-3, steps, obs, tx, ty,
                                    dt_check=dt_check)[0]

    L, gq, gk, gc, gg, gs = oracle.fit_loss_grad(x, y, vx, vy, q, m, K, C, G, 1e-3, steps, obs, tx, ty,
                                                  state_grad=True, dt_check=dt_check)
    checks = [("kappa", gk, lambda v: loss(K=v), K), ("confine", gc, lambda v: loss(C=v), C),
              ("gamma", gg, lambda v: loss(G=v), G)]
    for i in (0, 7, 15):
        def fq(v, i=i):
            qq = q.copy()
            qq[i] = v
            return loss(q=qq)
        checks.append((f"q{i}", gq[i], fq, q[i]))
    for name, arr, g in (("x", x, gs[0]), ("y", y, gs[1]), ("vx", vx, gs[2]), ("vy", vy, gs[3])):
        for i in (3, 12):
            def fs(v, i=i, arr=arr, name=name):
                a = arr.copy()
                a[i] = v
                return loss(**{name: a})
            scale = 1e-3 if name in ("x", "y") else 1e-3
            checks.append((f"{name}{i}", g[i], fs, arr[i] if arr[i] != 0 else scale))
    for name, ad, f, v in checks:
        h = abs(v) * 1e-4
        fd = _fd4(f, v, h)
        assert abs(fd - ad) <= 2e-7 * abs(ad) + 1e-9 * L / h, (name, ad, fd)


@pytest.mark.parametrize("dt_check", [False, True])
def test_fit_loss_equals_trajectory_loss(oracle, dt_check):
    n, steps, obs = 32, 40, 10
    kappa = 2000.0
    x, y, vx, vy, q, m = _setup(n, 9, 120e-6 if dt_check else 400e-6)
    tx, ty = _targets(oracle, x, y, vx, vy, q * 0.97, m, kappa, steps, obs, dt_check)
    L = oracle.fit_loss_grad(x, y, vx, vy, q, m, kappa, CONFINE, GAMMA, 1e-3, steps, obs, tx, ty,
                             dt_check=dt_check)[0]
    acc, xs, ys, vxs, vys = 0.0, x, y, vx, vy
    for b in range(steps // obs):
        r = oracle.rk4_steps(xs, ys, vxs, vys, q, m,
                             OracleParams(kappa, CONFINE, GAMMA, 0.0, 0, 7, int(dt_check), 1),
                             1e-3, obs, step0=b * obs)
        xs, ys, vxs, vys = r[0], r[1], r[2], r[3]
        acc += np.sum((xs - tx[b]) ** 2 + (ys - ty[b]) ** 2)
    ref = acc / (2 * n * (steps // obs))
    assert L == pytest.approx(ref, rel=1e-10)


def test_dt_cut_case_really_cuts(oracle):
    """The 110 um spacing case above starts with pairs inside 100 um."""
    x, y, *_ = _setup(16, 5, 110e-6)
    assert oracle.modify_time_step(x, y) < 1e-3
