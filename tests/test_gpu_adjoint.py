"""Device adjoint (demon_fit_loss_grad) against the oracle's CPU reference
gradient and against central finite differences of the device loss.

Bars: equal to the FP64 CPU reference gradient (1e-8 relative on short
trajectories; 1e-6 on the 200-step collapsing benchmark fit, where the two
forwards' last-bit differences grow with the trajectory), and within 1e-6
relative of central finite differences at h = 1e-6 relative -- or within
the difference quotient's measured rounding error where that is larger.
"""
import numpy as np
import pytest

import demon_simulation_code_amd as dm
from tests.helpers import CONFINE, GAMMA, KAPPA, M0, Q0, lattice

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def _problem(n, steps, obs, seed=3, jitter_q=0.1, kappa=KAPPA, start=None):
    x, y = lattice(n, seed=seed) if start is None else start
    rng = np.random.default_rng(seed)
    q_true = Q0 * (1 + jitter_q * rng.uniform(-1, 1, n))
    m = np.full(n, M0)
    # targets: trajectory of the true parameters
    with dm.Context(n, shielding=kappa, confine=CONFINE, gamma=GAMMA) as c:
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
    c = dm.Context(n, shielding=kappa, confine=confine, gamma=gamma)
    c.upload(x, y, None, None, q, m)
    return c


def _grad_dict(loss, g, L, gq, gk, gc, gg, gs):
    dev = {"loss": loss, "q": g["q"], "kappa": g["kappa"], "confine": g["confine"], "gamma": g["gamma"],
           "x0": g["x0"], "y0": g["y0"], "vx0": g["vx0"], "vy0": g["vy0"]}
    cpu = {"loss": L, "q": gq, "kappa": gk, "confine": gc, "gamma": gg,
           "x0": gs[0], "y0": gs[1], "vx0": gs[2], "vy0": gs[3]}
    return dev, cpu


def _rel(a, b):
    """max |a - b| / max |b| (scalars or arrays)."""
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(np.asarray(b))))


def _vs_cpu(oracle, n, steps, obs, ckpt, seed, tol, conditioning=False):
    """Device loss and gradient against the oracle's CPU reference gradient.
    With `conditioning`, the bar per component is max(tol, 10 s), where s is
    how far the device's own result moves when the initial x is changed by
    one ulp: a chaotic trajectory amplifies last-bit differences, and the two
    forwards (device and CPU) differ in their last bits from the first step."""
    x, y, m, q, tx, ty = _problem(n, steps, obs, seed=seed)
    kappa, conf, gam = KAPPA * 1.02, CONFINE * 0.97, GAMMA * 1.05
    with _ctx(n, x, y, m, q) as c:
        loss, g = c.fit_loss_grad(steps, 1e-3, obs, ckpt, tx, ty, state_grad=True)
        fwd_ms, bwd_ms = c.fit_timing()
        if conditioning:
            c.upload(np.nextafter(x, np.inf), y, None, None, q, m)
            loss2, g2 = c.fit_loss_grad(steps, 1e-3, obs, ckpt, tx, ty, state_grad=True)
    z = np.zeros(n)
    ref = oracle.fit_loss_grad(x, y, z, z, q, m, kappa, conf, gam, 1e-3, steps, obs, tx, ty,
                               state_grad=True, dt_check=True)
    dev, cpu = _grad_dict(loss, g, *ref)
    errs = {k: _rel(dev[k], cpu[k]) for k in dev}
    bars = {k: tol for k in dev}
    if conditioning:
        dev2, _ = _grad_dict(loss2, g2, *ref)
        for k in dev:
            bars[k] = max(tol, 10.0 * _rel(dev2[k], dev[k]))
    print(f"N={n} steps={steps} ckpt={ckpt}: forward {fwd_ms:.1f} ms, recompute+adjoint {bwd_ms:.1f} ms; "
          + ", ".join(f"{k} {errs[k]:.1e} (bar {bars[k]:.1e})" for k in dev))
    for k in dev:
        assert errs[k] <= bars[k], (k, errs[k], bars[k])


@pytest.mark.parametrize("ckpt", [20, 7, 40])
def test_adjoint_matches_cpu_reference_gradient(oracle, ckpt):
    _vs_cpu(oracle, 256, 40, 10, ckpt, seed=3, tol=1e-8)


@pytest.mark.timeout(600)
def test_adjoint_matches_cpu_reference_gradient_config5(oracle):
    """The benchmark fit at full size: N = 4096 on the 500 um lattice, 200 steps
    with the dt check, observations every 10th step, checkpoints every 20.  The
    cloud implodes, and a one-ulp change of the start moves the gradient by
    about 1e-5 relative (the CPU gradient moves alike), so the bar is ten
    times the device's own one-ulp sensitivity."""
    _vs_cpu(oracle, 4096, 200, 10, 20, seed=11, tol=1e-8, conditioning=True)


@pytest.mark.parametrize("n", [256, 8192])
def test_kept_segments_do_not_change_the_gradient(monkeypatch, n):
    """The forward sweep keeps the stage records of as many trailing checkpoint
    segments as DEMON_FIT_KEEP_BYTES allows and the backward sweep recomputes
    the others: loss and gradient are the same to the bit whether 1, 2 or all
    6 segments (40 steps, checkpoints every 7; the last one partial) are kept.
    N = 8192 runs the forward with the Newton-3 pair kernel, N = 256 with the
    gather kernel."""
    steps, obs, ckpt = 40, 10, 7
    x, y, m, q, tx, ty = _problem(n, steps, obs, seed=3)
    npad = -(-n // 1024) * 1024                   # n rounded up to 1024
    seg = ckpt * 4 * (5 * npad + npad // 16) * 8  # bytes of one segment's records
    runs = []
    with _ctx(n, x, y, m, q) as c:
        assert c.pair_mode() == (dm.PAIRS_NEWTON3 if n == 8192 else dm.PAIRS_GATHER)
        for keep in (0, 2 * seg + seg // 2, None):
            if keep is None:
                monkeypatch.delenv("DEMON_FIT_KEEP_BYTES", raising=False)
            else:
                monkeypatch.setenv("DEMON_FIT_KEEP_BYTES", str(keep))
            runs.append(c.fit_loss_grad(steps, 1e-3, obs, ckpt, tx, ty, state_grad=True))
            assert np.isfinite(runs[-1][0]) and np.all(np.isfinite(runs[-1][1]["q"]))
            s = c.download()
            assert np.array_equal(s[0], x) and np.array_equal(s[1], y)
    (l0, g0), rest = runs[0], runs[1:]
    assert g0["q"].any() and g0["kappa"] != 0.0
    for l1, g1 in rest:
        assert l1 == l0
        assert g1.keys() == g0.keys()
        for k in g0:
            assert np.array_equal(np.asarray(g1[k]), np.asarray(g0[k])), k


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


@pytest.mark.timeout(600)
def test_adjoint_vs_central_differences_config5():
    """N = 4096, 200 steps with the dt check, observations every 10th step,
    checkpoints every 20: 32 random parameters (29 charges and the three
    globals) by central differences at h = 1e-6 relative.

    The instance is a smooth one: lambda = 5 mm (cutoff 50 mm, beyond the
    cloud's diameter) and the start state relaxed to equilibrium (a 280 um
    lattice damped at gamma = 30/s for 1.5 s), so no pair crosses the hard
    cutoff, no step's dt is cut and the trajectory is not chaotic.  On the
    cold 500 um lattice with lambda = 500 um the cloud implodes within 40 ms:
    the reference's cutoff at 10 lambda makes the force jump, every parameter
    moves pairs across it, close passes cut dt tenfold and chaos amplifies
    rounding, so a difference quotient there measures jumps and noise, not
    the derivative; there the device gradient is checked against the CPU
    reference gradient instead.

    The difference quotient is the fourth-order central one, [8(L(+h) -
    L(-h)) - (L(+2h) - L(-2h))] / 12h, so that truncation (h^4) stays far
    below 1e-6.  Its rounding error is bounded from the loss's own rounding
    noise, measured as the largest second difference |L(v+d) - 2 L(v) +
    L(v-d)| / 2 over six tiny relative steps d = 1e-12 .. 3e-10 (the
    curvature part is below 1e-19 |L|), times the stencil's noise gain 18/12
    / h and a factor 4 for the sampling of that maximum."""
    n, steps, obs, ckpt = 4096, 200, 10, 20
    kappa_true, spacing = 200.0, 280e-6
    x, y = lattice(n, seed=11, spacing=spacing, jitter=0.1 * spacing)
    m = np.full(n, M0)
    q_true = Q0 * (1 + 0.1 * np.random.default_rng(11).uniform(-1, 1, n))
    with dm.Context(n, shielding=kappa_true, confine=CONFINE, gamma=30.0) as c:
        c.upload(x, y, None, None, q_true, m)
        c.rk4_steps(1500, 1e-3, 0.0)
        x, y, vx, vy = c.download()
        assert c.stats().dt_reductions == 0
    x, y, m, q, tx, ty = _problem(n, steps, obs, seed=11, kappa=kappa_true, start=(x, y))
    kappa, conf, gam = kappa_true * 1.02, CONFINE * 0.97, GAMMA * 1.05
    with _ctx(n, x, y, m, q, kappa, conf, gam) as c:
        L, g = c.fit_loss_grad(steps, 1e-3, obs, ckpt, tx, ty)
        fwd_ms, bwd_ms = c.fit_timing()
        stats = c.stats()

        def loss_at(**kw):
            qq = kw.pop("q", q)
            c.set_params(shielding=kw.get("kappa", kappa), confine=kw.get("confine", conf),
                         drag_gamma=kw.get("gamma", gam))
            c.upload(x, y, None, None, qq, m)
            return c.fit_loss_grad(steps, 1e-3, obs, ckpt, tx, ty, grad=False)[0]

        def shifted(name, v, d):
            if name == "q":
                qq = q.copy()
                qq[v] += d
                return loss_at(q=qq)
            return loss_at(**{name: v + d})

        rng = np.random.default_rng(0)
        checks = [("kappa", kappa), ("confine", conf), ("gamma", gam)]
        checks += [("q", int(i)) for i in rng.choice(n, 29, replace=False)]
        rows, worst, plain = [], 0.0, 0
        for name, v in checks:
            base = abs(q[v]) if name == "q" else abs(v)
            ad = g["q"][v] if name == "q" else g[name]
            noise = 0.0
            for rel in (1e-12, 3e-12, 1e-11, 3e-11, 1e-10, 3e-10):
                d = base * rel
                noise = max(noise, abs(shifted(name, v, d) - 2 * L + shifted(name, v, -d)) / 2)
            noise = max(noise, EPS * abs(L))
            h = base * 1e-6
            fd = (8 * (shifted(name, v, h) - shifted(name, v, -h))
                  - (shifted(name, v, 2 * h) - shifted(name, v, -2 * h))) / (12 * h)
            rel_err = abs(fd - ad) / abs(ad)
            bound = 1e-6 * abs(ad) + 4 * 1.5 * noise / h
            plain += rel_err <= 1e-6
            worst = max(worst, abs(fd - ad) / bound)
            rows.append((name, v, ad, fd, rel_err, bound / abs(ad)))
    for r in rows:
        print("  %-8s %6s adjoint %+.9e  FD %+.9e  rel %.1e  bound %.1e" % (r[0], r[1] if r[0] == "q" else "", *r[2:]))
    print(f"fit N={n}: forward {fwd_ms:.1f} ms, recompute+adjoint {bwd_ms:.1f} ms, "
          f"ratio {(fwd_ms + bwd_ms) / fwd_ms:.2f}; {plain}/{len(rows)} within 1e-6 plain, "
          f"worst FD error / bound {worst:.3f}; dt reductions {stats.dt_reductions}")
    assert stats.dt_reductions == 0
    for name, v, ad, fd, rel_err, b in rows:
        assert rel_err <= b, (name, v, ad, fd, rel_err, b)


@pytest.mark.parametrize("close_pair", [False, True])
def test_torch_autograd_binding(close_pair):
    """The torch.autograd wrapper returns the C ABI's loss and gradients to
    the bit, with the reference's dt check on by default.  close_pair puts
    particles 0 and 1 (one group of four, so a pair the reference's check
    tests) 60 um apart: the first steps' dt is then cut tenfold until the
    pair has separated beyond 100 um, and a wrapper stepping with the fixed
    dt would differ."""
    torch = pytest.importorskip("torch")
    from demon_simulation_code_amd.autograd import TrajectoryFit

    n, steps, obs = 256, 20, 10
    x, y, m, q, tx, ty = _problem(n, steps, obs, seed=7)
    if close_pair:
        x, y = x.copy(), y.copy()
        x[1], y[1] = x[0] + 60e-6, y[0]
        with _ctx(n, x, y, m, q) as c:
            c.rk4_steps(steps, 1e-3, 0.0)
            assert c.stats().dt_reductions > 0
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
    assert np.array_equal(qt.grad.numpy(), 2 * g["q"])
    assert kt.grad.item() == 2 * g["kappa"] and gt.grad.item() == 2 * g["gamma"]
    assert ct.grad.item() == 2 * g["confine"]
    if close_pair:
        # the fixed-dt fit is a different function here
        fixed = TrajectoryFit(x, y, m, tx, ty, steps=steps, dt=1e-3, obs_every=obs, ckpt_every=10,
                              dt_check=False)
        lf = fixed(qt.detach(), kt.detach(), ct.detach(), gt.detach()).item()
        fixed.close()
        assert lf != L


def test_adjoint_pair_counts_equal_forward_counts():
    """The adjoint pair kernel visits the forward kernel's cutoff set: over a
    fit, its counts equal the forward counts of the same trajectory (the
    forward count of pairs inside the cutoff is the oracle's exact count);
    with timing on its launches (4 per step) are timed, with timing off not."""
    n, steps, obs, ckpt = 1024, 20, 10, 5
    x, y, m, q, tx, ty = _problem(n, steps, obs)
    with _ctx(n, x, y, m, q) as c:
        c.rk4_steps(steps, 1e-3, 0.0)
        want = c.pair_counts()
        c.upload(x, y, None, None, q, m)
        c.fit_loss_grad(steps, 1e-3, obs, ckpt, tx, ty)
        ak = c.fit_adjoint_kernel()
        c.set_timing(False)
        c.fit_loss_grad(steps, 1e-3, obs, ckpt, tx, ty)
        ak2 = c.fit_adjoint_kernel()
    assert 0 < want[0] < want[1]
    assert (ak["near_pairs"], ak["tested_pairs"]) == want
    assert ak["launches"] == 4 * steps and ak["pair_ms"] > 0
    assert ak2["launches"] == 0 and (ak2["near_pairs"], ak2["tested_pairs"]) == want
