"""Pins of the CPU oracle (no GPU needed).

Each oracle function is checked against something other than its own
formula: the reference's own outputs (golden fixtures made by
tests/golden/make_golden.py from the reference's sources; live when
oracle/_ref is built), closed forms, invariants and long-double brute force.
"""
from pathlib import Path

import numpy as np
import pytest

from oracle.bind import OracleParams, Reference

GOLD = Path(__file__).resolve().parent / "golden"
E = -1.602e-19
Q0 = 6000 * E
M0 = 2.8e-14
KAPPA = 2000.0
COUL = 1.0 / (4.0 * np.pi * 8.8541878e-12)
CONFINE = 5e-11 / abs(Q0)


def bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.int64), np.asarray(b).view(np.int64))


# --- golden vectors from the reference ------------------------------------

@pytest.mark.parametrize("n", [64, 256, 1024])
def test_yukawa_matches_reference_golden_bitwise(oracle, n):
    d = np.load(GOLD / f"forces_n{n}.npz")
    fx, fy = oracle.yukawa(d["x"], d["y"], d["q"], float(d["kappa"]))
    assert bits_equal(fx, d["f1x"]) and bits_equal(fy, d["f1y"])
    # force3 reads the stage caches; same arithmetic
    assert bits_equal(fx, d["f3x"]) and bits_equal(fy, d["f3y"])


@pytest.mark.parametrize("n", [64, 256, 1024])
def test_rk4_matches_reference_golden_bitwise(oracle, n):
    d = np.load(GOLD / f"rk4_n{n}.npz")
    op = OracleParams(float(d["kappa"]), float(d["confine"]), float(d["gamma"]), 0.0, 0, 7, 1)
    state = (d["x"], d["y"], np.zeros(n), np.zeros(n))
    done = 0
    for target in (1, 10, 100):
        if f"x_{target}" not in d:
            continue
        *state, t, _ = oracle.rk4_steps(*state, d["q"], d["m"], op, float(d["dt"]), target - done,
                                        t=done * float(d["dt"]))
        done = target
        for name, arr in zip(("x", "y", "vx", "vy"), state):
            assert bits_equal(arr, d[f"{name}_{target}"]), (name, target)


def test_dtcheck_matches_reference_golden(oracle):
    d = np.load(GOLD / "dtcheck.npz")
    for a, b, sep, want in d["cases"]:
        xx, yy = d["x"].copy(), d["y"].copy()
        xx[int(b)] = xx[int(a)] + sep
        yy[int(b)] = yy[int(a)]
        assert oracle.modify_time_step(xx, yy, 1e-4, 1e-3) == want


# --- live comparison with the reference build (skipped where not built) ---

@pytest.mark.skipif(not Reference.available(), reason="oracle/_ref not built (needs the reference sources)")
def test_random_configs_bitwise_vs_live_reference(oracle):
    ref = Reference()
    rng = np.random.default_rng(0)
    for trial in range(6):
        n = 4 * int(rng.integers(2, 40))
        x = rng.uniform(-2e-3, 2e-3, n)
        y = rng.uniform(-2e-3, 2e-3, n)
        q = Q0 * rng.uniform(0.5, 1.5, n)
        kappa = float(rng.uniform(500, 8000))
        fx, fy = oracle.yukawa(x, y, q, kappa)
        rx, ry = ref.yukawa(x, y, q, kappa, stage=1 + trial % 4)
        assert bits_equal(fx, rx) and bits_equal(fy, ry)
        assert oracle.modify_time_step(x, y, 1e-4, 1e-3) == ref.modify_time_step(x, y, 1e-4, 1e-3)
        m = M0 * rng.uniform(0.8, 1.2, n)
        op = OracleParams(kappa, CONFINE, 2.0, 0.0, 0, 7, 1)
        z = np.zeros(n)
        o = oracle.rk4_steps(x, y, z, z, q, m, op, 1e-4, 5)
        r = ref.rk4_steps(x, y, z, z, q, m, kappa, CONFINE, 2.0, 0.0, 7, 1e-4, 5)
        for a, b in zip(o[:4], r[:4]):
            assert bits_equal(a, b)
        assert o[4] == r[4]


@pytest.mark.skipif(not Reference.available(), reason="oracle/_ref not built (needs the reference sources)")
def test_move_loop_step_count_vs_live_reference(oracle):
    """moveParticles(end) drives by `while (t < end)`; the oracle's count is the reference's."""
    ref = Reference()
    n = 16
    rng = np.random.default_rng(1)
    x = rng.uniform(-3e-3, 3e-3, n)
    y = rng.uniform(-3e-3, 3e-3, n)
    q, m, z = np.full(n, Q0), np.full(n, M0), np.zeros(n)
    op = OracleParams(KAPPA, CONFINE, 2.0, 0.0, 0, 7, 1)
    for end in (0.01, 0.1, 0.37):
        o = oracle.rk4_move(x, y, z, z, q, m, op, 1e-3, end)
        r = ref.rk4_move(x, y, z, z, q, m, KAPPA, CONFINE, 2.0, 0.0, 7, 1e-3, end)
        assert o[4] == r[4]
        for a, b in zip(o[:4], r[:4]):
            assert bits_equal(a, b)


# --- closed forms and invariants -------------------------------------------

def test_two_body_closed_form(oracle):
    for r in (100e-6, 500e-6, 2e-3, 4.9e-3):
        x = np.array([0.0, r])
        y = np.array([0.0, 0.0])
        q = np.array([Q0, 1.3 * Q0])
        fx, fy = oracle.yukawa(x, y, q, KAPPA)
        u = r * KAPPA
        f = COUL * Q0 * 1.3 * Q0 * (1 + u) * np.exp(-u) / r**2
        assert fx[0] == pytest.approx(-f, rel=1e-14)
        assert fx[1] == pytest.approx(f, rel=1e-14)
        assert fy[0] == 0.0 and fy[1] == 0.0


def test_cutoff_is_hard(oracle):
    r_cut = 10.0 / KAPPA
    for r, live in ((np.nextafter(r_cut, 0), True), (r_cut, False), (np.nextafter(r_cut, 1), False),
                    (2 * r_cut, False)):
        x = np.array([0.0, r])
        fx, _ = oracle.yukawa(x, np.zeros(2), np.array([Q0, Q0]), KAPPA)
        assert (fx[0] != 0.0) == live


def test_newton3_and_long_double(oracle):
    rng = np.random.default_rng(5)
    n = 64
    x = rng.uniform(-2e-3, 2e-3, n)
    y = rng.uniform(-2e-3, 2e-3, n)
    q = Q0 * rng.uniform(0.8, 1.2, n)
    fx, fy = oracle.yukawa(x, y, q, KAPPA)
    gx, gy = oracle.yukawa_gather(x, y, q, KAPPA)
    lx, ly = oracle.yukawa_longdouble(x, y, q, KAPPA)
    scale = np.abs(lx).max() + np.abs(ly).max()
    assert abs(fx.sum()) / scale < 1e-13 and abs(fy.sum()) / scale < 1e-13
    for ax, ay in ((fx, fy), (gx, gy)):
        assert np.max(np.hypot(ax - lx, ay - ly)) / scale < 1e-13


def test_rows_equal_gather(oracle):
    rng = np.random.default_rng(6)
    n = 200
    x, y = rng.uniform(-3e-3, 3e-3, (2, n))
    q = Q0 * rng.uniform(0.8, 1.2, n)
    gx, gy = oracle.yukawa_gather(x, y, q, KAPPA)
    rows = np.array([0, 7, 199, 100])
    rx, ry = oracle.yukawa_rows(x, y, q, KAPPA, rows)
    assert bits_equal(rx, gx[rows]) and bits_equal(ry, gy[rows])


def test_symmetric_hexagon_centre_force_vanishes(oracle):
    a = 500e-6
    ang = np.arange(6) * np.pi / 3
    x = np.r_[0.0, a * np.cos(ang)]
    y = np.r_[0.0, a * np.sin(ang)]
    fx, fy = oracle.yukawa_gather(x, y, np.full(7, Q0), KAPPA)
    single = COUL * Q0 * Q0 * (1 + a * KAPPA) * np.exp(-a * KAPPA) / a**2
    assert abs(fx[0]) < 1e-14 * single and abs(fy[0]) < 1e-14 * single


def test_confinement_drag_closed_form(oracle):
    rng = np.random.default_rng(7)
    n = 10
    x, y, vx, vy = rng.normal(size=(4, n)) * 1e-3
    q = Q0 * rng.uniform(0.5, 1.5, n)
    m = M0 * rng.uniform(0.5, 1.5, n)
    op = OracleParams(KAPPA, CONFINE, 2.0, 0.0, 0, 1 | 2, 0)
    fx, fy = oracle.stage_force(x, y, vx, vy, q, m, op)
    np.testing.assert_allclose(fx, q * CONFINE * x - 2.0 * m * vx, rtol=1e-15, atol=0)
    np.testing.assert_allclose(fy, q * CONFINE * y - 2.0 * m * vy, rtol=1e-15, atol=0)


def test_rk4_linear_system_amplification_polynomial(oracle):
    """Confinement + drag only: one RK4 step is T4(hA) applied to (x, v)."""
    n = 4
    rng = np.random.default_rng(8)
    x, y, vx, vy = rng.normal(size=(4, n)) * 1e-3
    q = np.full(n, Q0)
    m = np.full(n, M0)
    gamma, h = 2.0, 1e-3
    op = OracleParams(KAPPA, CONFINE, gamma, 0.0, 0, 1 | 2, 0)
    o = oracle.rk4_steps(x, y, vx, vy, q, m, op, h, 1)
    A = np.array([[0.0, 1.0], [Q0 * CONFINE / M0, -gamma]])
    hA = h * A
    T = np.eye(2) + hA + hA @ hA / 2 + hA @ hA @ hA / 6 + hA @ hA @ hA @ hA / 24
    for i in range(n):
        ex = T @ np.array([x[i], vx[i]])
        ey = T @ np.array([y[i], vy[i]])
        assert o[0][i] == pytest.approx(ex[0], rel=1e-13)
        assert o[2][i] == pytest.approx(ex[1], rel=1e-12)
        assert o[1][i] == pytest.approx(ey[0], rel=1e-13)
        assert o[3][i] == pytest.approx(ey[1], rel=1e-12)


def test_two_body_energy_drift_is_fourth_order(oracle):
    """gamma = 0, confinement + Yukawa: energy error after a fixed time scales ~ h^4."""
    x0 = np.array([-0.4e-3, 0.4e-3])
    y0 = np.zeros(2)
    vx0 = np.zeros(2)
    vy0 = np.array([0.01, -0.01])
    q, m = np.full(2, Q0), np.full(2, M0)

    def energy(x, y, vx, vy):
        r = np.hypot(x[0] - x[1], y[0] - y[1])
        u = r * KAPPA
        pot = COUL * Q0 * Q0 * np.exp(-u) / r - (Q0 * CONFINE / 2) * np.sum(x**2 + y**2)
        return 0.5 * M0 * np.sum(vx**2 + vy**2) + pot

    e0 = energy(x0, y0, vx0, vy0)
    errs = []
    for h in (4e-4, 2e-4):
        op = OracleParams(KAPPA, CONFINE, 0.0, 0.0, 0, 1 | 2 | 4, 0)
        o = oracle.rk4_steps(x0, y0, vx0, vy0, q, m, op, h, int(round(0.04 / h)))
        errs.append(abs(energy(*o[:4]) - e0))
    ratio = errs[0] / errs[1]
    assert 10 < ratio < 40, ratio  # 2^4 = 16


# --- dt check ----------------------------------------------------------------

def test_dtcheck_coverage_predicate(oracle):
    """Literal lane-shuffle replay vs the closed-form predicate, pair by pair."""
    rng = np.random.default_rng(9)
    n = 24
    base_x = np.arange(n) * 1e-3
    base_y = np.zeros(n)
    for i in range(n):
        for j in range(i + 1, n):
            x, y = base_x.copy(), base_y.copy()
            x[j] = x[i] + 30e-6
            y[j] = 0.0 if rng.random() < 0.5 else 1e-6
            reduced = oracle.modify_time_step(x, y, 1e-4, 1e-3) < 1e-3
            assert reduced == oracle.covers(i, j), (i, j)
    # hole examples named in the analysis of Integrator.cpp:112-129
    assert not oracle.covers(3, 4) and not oracle.covers(0, 5) and oracle.covers(0, 4)


def test_dtcheck_min_matches_replay(oracle):
    rng = np.random.default_rng(10)
    for _ in range(20):
        n = 4 * int(rng.integers(2, 12))
        x = rng.uniform(0, 2e-3, n)
        y = rng.uniform(0, 2e-3, n)
        s = oracle.dtcheck_min_s(x, y)
        d = np.float32(np.sqrt(np.float32(s)))
        dist, dt = np.float32(1e-4), 1e-3
        while d <= dist:
            dist = np.float32(dist / np.float32(10))
            dt = dt / 10.0
        assert oracle.modify_time_step(x, y, 1e-4, 1e-3) == dt


# --- RNG ---------------------------------------------------------------------

def test_philox_known_answers(oracle):
    """Random123 known-answer vectors for philox4x32-10."""
    assert oracle.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert oracle.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert oracle.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_thermal_moments(oracle):
    n, heat = 200000, 1.0
    fx, fy = oracle.thermal(n, heat, 123, 7, 2)
    mag = np.hypot(fx, fy)
    assert np.mean(mag) == pytest.approx(0.5, abs=3e-3)          # E[U] = 1/2
    assert np.mean(mag**2) == pytest.approx(1 / 3, abs=3e-3)     # E[U^2] = 1/3
    assert abs(np.mean(fx)) < 3e-3 and abs(np.mean(fy)) < 3e-3   # isotropy
    assert np.mean(fx**2) == pytest.approx(1 / 6, abs=3e-3)
    gx, _ = oracle.thermal(n, heat, 123, 7, 3)
    assert abs(np.corrcoef(fx, gx)[0, 1]) < 0.01                  # stages independent


# --- Magnetic and TimeVaryingDrag (driver order, stage times) ---------------

MAGNETIC, TV_DRAG = 1024, 256


@pytest.mark.skipif(not Reference.available(), reason="oracle/_ref not built (needs the reference sources)")
def test_magnetic_tvdrag_bitwise_vs_live_reference(oracle):
    """Confinement, Drag/TimeVaryingDrag, Magnetic and Yukawa in the driver's
    order, with the dt check, against the reference's own classes."""
    ref = Reference()
    rng = np.random.default_rng(11)
    for trial, forces in enumerate((1 | 2 | 4 | MAGNETIC, 1 | 4 | TV_DRAG | MAGNETIC, 1 | 4 | TV_DRAG)):
        n = 4 * int(rng.integers(3, 30))
        x = rng.uniform(-2e-3, 2e-3, n)
        y = rng.uniform(-2e-3, 2e-3, n)
        vx = rng.normal(size=n) * 1e-4
        vy = rng.normal(size=n) * 1e-4
        q = Q0 * rng.uniform(0.5, 1.5, n)
        m = M0 * rng.uniform(0.8, 1.2, n)
        kappa, bfield, scale = float(rng.uniform(500, 8000)), float(rng.uniform(-2, 2)), float(rng.uniform(-5, 5))
        op = OracleParams(kappa, CONFINE, 2.0, 0.0, 0, forces, 1, 0, bfield, scale)
        o = oracle.rk4_steps(x, y, vx, vy, q, m, op, 1e-4, 6, t=0.25)
        r = ref.rk4_steps_ext(x, y, vx, vy, q, m, kappa, CONFINE, 2.0, 0.0, bfield, scale, forces, 1e-4, 6,
                              t=0.25)
        for a, b in zip(o[:4], r[:4]):
            assert bits_equal(a, b), trial
        assert o[4] == r[4]


def test_magnetic_rk4_step_is_amplification_polynomial(oracle):
    """Magnetic force alone is linear: dv/dt = w (vy, -vx) with w = qB/m, and
    one RK4 step is T4(hA) on (x, y, vx, vy)."""
    n = 3
    rng = np.random.default_rng(12)
    x, y, vx, vy = rng.normal(size=(4, n)) * 1e-3
    q, m = np.full(n, Q0), np.full(n, M0)
    bfield, h = 0.7, 1e-3
    op = OracleParams(KAPPA, CONFINE, 2.0, 0.0, 0, MAGNETIC, 0, 0, bfield, 0.0)
    o = oracle.rk4_steps(x, y, vx, vy, q, m, op, h, 1)
    w = Q0 * bfield / M0
    A = np.zeros((4, 4))
    A[0, 2] = A[1, 3] = 1.0
    A[2, 3], A[3, 2] = w, -w
    hA = h * A
    T = np.eye(4)
    term = np.eye(4)
    for k in range(1, 5):
        term = term @ hA / k
        T = T + term
    for i in range(n):
        e = T @ np.array([x[i], y[i], vx[i], vy[i]])
        got = np.array([o[0][i], o[1][i], o[2][i], o[3][i]])
        np.testing.assert_allclose(got, e, rtol=1e-13, atol=1e-13 * np.abs(e).max())


def test_tvdrag_uses_the_stage_times(oracle):
    """TimeVaryingDrag alone: dv/dt = -(s t + g) v.  One RK4 step from t0
    evaluates the rate at t0, t0 + h/2, t0 + h/2, t0 + h."""
    n = 2
    vx = np.array([1e-3, -2e-3])
    vy = np.array([3e-4, 5e-4])
    z = np.zeros(n)
    q, m = np.full(n, Q0), np.full(n, M0)
    s, g, h, t0 = 40.0, 2.0, 1e-2, 0.3
    op = OracleParams(KAPPA, CONFINE, g, 0.0, 0, TV_DRAG, 0, 0, 0.0, s)
    o = oracle.rk4_steps(z, z, vx, vy, q, m, op, h, 1, t=t0)

    def rate(t):
        return -(s * t + g)

    for v0, got in ((vx, o[2]), (vy, o[3])):
        k1 = h * rate(t0) * v0
        k2 = h * rate(t0 + h / 2) * (v0 + k1 / 2)
        k3 = h * rate(t0 + h / 2) * (v0 + k2 / 2)
        k4 = h * rate(t0 + h) * (v0 + k3)
        np.testing.assert_allclose(got, v0 + (k1 + 2 * k2 + 2 * k3 + k4) / 6, rtol=1e-14)
    assert o[4] == t0 + h


# --- Runge_Kutta2 (the driver's -I) -----------------------------------------

@pytest.mark.skipif(not Reference.available(), reason="oracle/_ref not built (needs the reference sources)")
def test_rk2_bitwise_vs_live_reference(oracle):
    ref = Reference()
    rng = np.random.default_rng(21)
    for forces in (7, 1 | 4 | TV_DRAG | MAGNETIC):
        n = 4 * int(rng.integers(3, 30))
        x = rng.uniform(-2e-3, 2e-3, n)
        y = rng.uniform(-2e-3, 2e-3, n)
        vx, vy = rng.normal(size=(2, n)) * 1e-4
        q = Q0 * rng.uniform(0.5, 1.5, n)
        m = M0 * rng.uniform(0.8, 1.2, n)
        kappa = float(rng.uniform(500, 8000))
        op = OracleParams(kappa, CONFINE, 2.0, 0.0, 0, forces, 1, 0, 0.6, -2.0)
        o = oracle.rk2_steps(x, y, vx, vy, q, m, op, 1e-4, 7, t=0.1)
        r = ref.rk4_steps_ext(x, y, vx, vy, q, m, kappa, CONFINE, 2.0, 0.0, 0.6, -2.0, forces, 1e-4, 7, t=0.1,
                              order=2)
        for a, b in zip(o[:4], r[:4]):
            assert bits_equal(a, b), forces
        assert o[4] == r[4]


def test_rk2_step_is_midpoint_polynomial(oracle):
    """Confinement + drag only: one Runge_Kutta2 step is 1 + hA + (hA)^2/2."""
    n = 4
    rng = np.random.default_rng(22)
    x, y, vx, vy = rng.normal(size=(4, n)) * 1e-3
    q, m = np.full(n, Q0), np.full(n, M0)
    gamma, h = 2.0, 1e-3
    op = OracleParams(KAPPA, CONFINE, gamma, 0.0, 0, 1 | 2, 0)
    o = oracle.rk2_steps(x, y, vx, vy, q, m, op, h, 1)
    hA = h * np.array([[0.0, 1.0], [Q0 * CONFINE / M0, -gamma]])
    T = np.eye(2) + hA + hA @ hA / 2
    for i in range(n):
        ex, ey = T @ np.array([x[i], vx[i]]), T @ np.array([y[i], vy[i]])
        assert o[0][i] == pytest.approx(ex[0], rel=1e-13) and o[2][i] == pytest.approx(ex[1], rel=1e-12)
        assert o[1][i] == pytest.approx(ey[0], rel=1e-13) and o[3][i] == pytest.approx(ey[1], rel=1e-12)


# --- the further force plugins ---------------------------------------------

RECT, VOID, ROT, DRIVE, ELEC, VERT, GRAV = 8192, 2048, 128, 64, 4096, 16384, 8
TH_LOCAL, TV_THERMAL = 32, 512


def plugin_params(forces, kappa=3000.0, seed=0):
    """Constants that make each further plugin comparable to the confinement
    force on a mm-sized cloud of the test particles."""
    op = OracleParams(kappa, CONFINE, 2.0, 0.0, seed, forces, 1, 0, 0.5, -3.0)
    op.confine_x, op.confine_y = CONFINE, 3 * CONFINE
    op.void_decay = 2000.0
    op.rot_inner, op.rot_outer, op.rot_const = 5e-4, 1.5e-3, 1e-13
    op.drive_const, op.drive_amplitude, op.drive_shift = 1e-5, 1e-13, 7e-3
    op.electric_field, op.plasma_radius = -CONFINE, 2e-3
    op.vert_electric_field, op.vert_decay = 1e2, 3e-3
    op.gravity = -9.8
    return op


def plugin_cloud(rng, n):
    x = rng.uniform(-2e-3, 2e-3, n)
    y = rng.uniform(-2e-3, 2e-3, n)
    vx, vy = rng.normal(size=(2, n)) * 1e-4
    q = Q0 * rng.uniform(0.5, 1.5, n)
    m = M0 * rng.uniform(0.8, 1.2, n)
    return x, y, vx, vy, q, m


def test_plugins_match_reference_golden_bitwise(oracle):
    """The reference's own plugin classes' forceS(t) and 6-step RK4 / RK2
    runs, recorded in tests/golden/plugins.npz, against the oracle."""
    d = np.load(GOLD / "plugins.npz")
    op = OracleParams()
    for name, v in zip(d["par_names"], d["par"]):
        setattr(op, str(name), float(v))
    op.dt_check = 1
    st = [d[k] for k in ("x", "y", "vx", "vy", "q", "m")]
    for f in d["force_sets"]:
        op.forces = int(f)
        fx, fy = oracle.stage_force(*st, op, t=0.37)
        assert bits_equal(fx, d[f"force_{f}_x"]) and bits_equal(fy, d[f"force_{f}_y"]), f
    for order in (4, 2):
        run = oracle.rk4_steps if order == 4 else oracle.rk2_steps
        for f in d["runs"]:
            op.forces = int(f)
            o = run(*st, op, 1e-4, 6, t=0.25)
            for name, arr in zip(("x", "y", "vx", "vy"), o[:4]):
                assert bits_equal(arr, d[f"rk{order}_{f}_{name}"]), (order, f, name)
            assert o[4] == float(d[f"rk{order}_{f}_t"])


@pytest.mark.skipif(not Reference.available(), reason="oracle/_ref not built (needs the reference sources)")
@pytest.mark.parametrize("forces", [RECT, VOID, ROT, DRIVE, ELEC, VERT, GRAV, 1 | VOID])
def test_plugin_force_bitwise_vs_live_reference(oracle, forces):
    """Each plugin's forceS(t) on one state, the reference's own class against
    the oracle, bit for bit (the same libm exp/sin/sqrt on both sides)."""
    ref = Reference()
    rng = np.random.default_rng(forces)
    x, y, vx, vy, q, m = plugin_cloud(rng, 48)
    op = plugin_params(forces)
    fo = oracle.stage_force(x, y, vx, vy, q, m, op, t=0.37)
    fx, fy = ref.force_all(x, y, vx, vy, q, m, op, t=0.37, reps=4)
    assert np.abs(fo[0]).max() + np.abs(fo[1]).max() > 1e-14
    for k in range(4):
        assert bits_equal(fo[0], fx[k]) and bits_equal(fo[1], fy[k]), k


@pytest.mark.skipif(not Reference.available(), reason="oracle/_ref not built (needs the reference sources)")
@pytest.mark.parametrize("order", [4, 2])
def test_plugin_trajectories_bitwise_vs_live_reference(oracle, order):
    """All deterministic plugins inside Runge_Kutta4 / Runge_Kutta2 with the dt
    check, in the driver's construction order, against the reference."""
    ref = Reference()
    rng = np.random.default_rng(40 + order)
    for forces in (VOID | 2 | 4 | DRIVE | ROT | GRAV, RECT | 2 | 4 | ELEC | VERT | MAGNETIC,
                   1 | VOID | 4 | TV_DRAG | GRAV):
        x, y, vx, vy, q, m = plugin_cloud(rng, 4 * int(rng.integers(3, 16)))
        op = plugin_params(forces, kappa=float(rng.uniform(500, 8000)))
        run = oracle.rk4_steps if order == 4 else oracle.rk2_steps
        o = run(x, y, vx, vy, q, m, op, 1e-4, 6, t=0.25)
        r = ref.rk_all(x, y, vx, vy, q, m, op, 1e-4, 6, t=0.25, order=order)
        for a, b in zip(o[:4], r[:4]):
            assert bits_equal(a, b), forces
        assert o[4] == r[4]


def test_plugin_closed_forms(oracle):
    """Rect is the linear field q (cX x, cY y); Void adds the radial term
    -q d e^{-d r} r_hat to the confinement; Rotational is tangential with
    magnitude c inside the shear layer and zero outside; Gravitational is
    m g along y; Driving is A sin(k x - w t) along x; Electric and
    VertElectric are the exponentially decaying fields."""
    rng = np.random.default_rng(3)
    x, y, vx, vy, q, m = plugin_cloud(rng, 64)
    z = np.zeros(64)

    def force(f, t=0.0):
        return oracle.stage_force(x, y, vx, vy, q, m, plugin_params(f), t=t)

    r = np.hypot(x, y)
    fx, fy = force(RECT)
    np.testing.assert_allclose(fx, q * CONFINE * x, rtol=1e-15)
    np.testing.assert_allclose(fy, q * 3 * CONFINE * y, rtol=1e-15)
    fx, fy = force(VOID)
    d = 2000.0
    np.testing.assert_allclose(fx, q * CONFINE * x - q * d * np.exp(-d * r) * x / r, rtol=1e-13)
    np.testing.assert_allclose(fy, q * CONFINE * y - q * d * np.exp(-d * r) * y / r, rtol=1e-13)
    fx, fy = force(ROT)
    inside = (r > 5e-4) & (r < 1.5e-3)
    assert inside.any() and (~inside).any()
    np.testing.assert_allclose(np.hypot(fx, fy)[inside], 1e-13, rtol=1e-15)
    np.testing.assert_allclose((fx * x + fy * y)[inside], 0.0, atol=1e-30)
    assert np.all(fx[~inside] == 0) and np.all(fy[~inside] == 0)
    assert np.all((x * fy - y * fx)[inside] > 0)   # counter-clockwise for c > 0
    fx, fy = force(GRAV)
    assert np.all(fx == 0) and np.array_equal(fy, m * -9.8)
    fx, fy = force(DRIVE, t=0.013)
    np.testing.assert_allclose(fx, 1e-13 * np.sin(2 * np.pi / 0.002 * x - 2 * np.pi * 10 * 0.013),
                               rtol=1e-12, atol=1e-27)
    assert np.all(fy == 0)
    fx, fy = force(ELEC)
    np.testing.assert_allclose(fx, q * -CONFINE * np.exp(-r / 2e-3) * x, rtol=1e-14)
    fx, fy = force(VERT)
    assert np.all(fx == 0)
    np.testing.assert_allclose(fy, q * 1e2 * np.exp(y / 3e-3), rtol=1e-14)
    # a zero damping constant makes the reference's 0/0 a NaN (DrivingForce.cpp:57)
    op = plugin_params(DRIVE)
    op.drive_const = 0.0
    assert np.isnan(oracle.stage_force(x, y, z, z, q, m, op)[0]).all()


@pytest.mark.skipif(not Reference.available(), reason="oracle/_ref not built (needs the reference sources)")
def test_thermal_localized_selection_vs_live_reference(oracle):
    """ThermalForceLocalized uses heat_in strictly inside the radius (on the
    stage position) and heat_out elsewhere: with heat_in << heat_out the
    largest of 16 draws classifies each particle, identically for the
    reference's classes and the oracle; a particle exactly on the circle is
    outside.  The magnitude law E|F|^2 = h^2/3 holds for both."""
    ref = Reference()
    rng = np.random.default_rng(6)
    n, R = 400, 1e-3
    x, y = rng.uniform(-2e-3, 2e-3, (2, n))
    x[0], y[0] = R, 0.0
    x[1], y[1] = np.nextafter(R, 0.0), 0.0
    z = np.zeros(n)
    q, m = np.full(n, Q0), np.full(n, M0)
    op = OracleParams(3000.0, 0.0, 0.0, 0.0, 99, TH_LOCAL, 0, 0)
    op.heat_in, op.heat_out, op.heat_radius = 1e-14, 1e-11, R
    fx, fy = ref.force_all(x, y, z, z, q, m, op, reps=16)
    mag_r = np.hypot(fx, fy).max(axis=0)
    draws = [oracle.stage_force(x, y, z, z, q, m, op, step=s, stage=1 + s % 4) for s in range(16)]
    mag_o = np.max([np.hypot(a, b) for a, b in draws], axis=0)
    inside = np.hypot(x, y) < R
    assert not inside[0] and inside[1]
    assert np.array_equal(mag_r < 1e-14, inside)
    assert np.array_equal(mag_o < 1e-14, inside)
    for f2 in ((fx ** 2 + fy ** 2)[:, ~inside], np.array([a ** 2 + b ** 2 for a, b in draws])[:, ~inside]):
        assert f2.mean() / 1e-22 == pytest.approx(1 / 3, rel=0.05)


@pytest.mark.skipif(not Reference.available(), reason="oracle/_ref not built (needs the reference sources)")
def test_tv_thermal_magnitude_at_stage_time_vs_live_reference(oracle):
    """TimeVaryingThermalForce at time t draws with magnitude h = s t + o:
    zero where h(t) = 0, bounded by |h| and E|F|^2 = h^2/3 otherwise, for the
    reference's classes and the oracle alike; RK4 stage times as TimeVaryingDrag."""
    ref = Reference()
    rng = np.random.default_rng(7)
    n = 400
    x, y = rng.uniform(-2e-3, 2e-3, (2, n))
    z = np.zeros(n)
    q, m = np.full(n, Q0), np.full(n, M0)
    op = OracleParams(3000.0, 0.0, 0.0, 0.0, 5, TV_THERMAL, 0, 0)
    op.heat_scale, op.heat_offset = 4e-12, -1e-12
    fx, fy = ref.force_all(x, y, z, z, q, m, op, t=0.25, reps=4)
    assert np.all(fx == 0) and np.all(fy == 0)
    fo = oracle.stage_force(x, y, z, z, q, m, op, t=0.25)
    assert np.all(fo[0] == 0) and np.all(fo[1] == 0)
    for t in (1.0, 2.5):
        h = 4e-12 * t - 1e-12
        fx, fy = ref.force_all(x, y, z, z, q, m, op, t=t, reps=20)
        draws = np.array([np.hypot(*oracle.stage_force(x, y, z, z, q, m, op, step=s, t=t)) for s in range(20)])
        for mag in (np.hypot(fx, fy), draws):
            assert mag.max() <= abs(h)
            assert (mag ** 2).mean() / h ** 2 == pytest.approx(1 / 3, rel=0.03)


# --- FFTAnalysis -a -----------------------------------------------------------

def test_average_fft_is_the_summed_dft(oracle):
    """The restated averageFFT equals numpy's FFT (a textbook transform) of
    each particle's series, summed over particles, |.| / nuse for k < rows/2;
    odd and even row counts; a stride wider than nuse."""
    rng = np.random.default_rng(8)
    for rows, stride, nuse in ((64, 5, 3), (77, 4, 4), (2, 1, 1), (1, 3, 2)):
        series = rng.normal(size=(rows, stride)) * 1e-3
        amp = oracle.average_fft(series, nuse)
        want = np.abs(np.fft.fft(series[:, :nuse], axis=0).sum(axis=1))[: rows // 2] / nuse
        assert amp.shape == (rows // 2,)
        np.testing.assert_allclose(amp, want, rtol=0, atol=1e-15 * np.abs(series).sum())


def test_average_fft_of_a_cosine(oracle):
    """A pure cosine of frequency f0 = k0 / (rows dt) puts rows/2 * A into bin
    k0 and nothing elsewhere (every particle in phase); the reference labels
    bin i with i / (rows * 0.01) Hz."""
    rows, k0, amp0 = 200, 13, 2e-4
    t = np.arange(rows)
    x = amp0 * np.cos(2 * np.pi * k0 * t / rows)
    series = np.stack([x, x, x], axis=1)
    amp = oracle.average_fft(series, 3)
    assert amp[k0] == pytest.approx(rows / 2 * amp0, rel=1e-12)
    others = np.delete(amp, k0)
    assert others.max() < 1e-12 * amp[k0]
