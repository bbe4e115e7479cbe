"""Pins of the oracle's cell-list form (no GPU needed).

oracle_yukawa_cells_add restates the reference's serial Yukawa loop
(ShieldedCoulombForce.cpp:43-56) over the pairs inside the cutoff only; it
must equal the all-pairs restatement oracle_yukawa_add -- which equals the
compiled reference bit for bit (tests/test_oracle.py) -- bit for bit, so
that it can step the 131072-particle benchmark cloud on the CPU.  The same
for the cell-list dt check (Integrator.cpp:68-155) and for whole RK4
trajectories with the two forms.
"""
import numpy as np
import pytest

from oracle.bind import OracleParams
from tests.test_oracle import CONFINE, KAPPA, M0, Q0, bits_equal


def hexagonal(n, spacing, jitter, seed):
    from demon_simulation_code_amd import hex_lattice

    return hex_lattice(n, spacing, jitter, seed)


def _cases():
    rng = np.random.default_rng(3)
    out = []
    # the benchmark lattice (most pairs beyond the cutoff)
    x, y = hexagonal(2048, 500e-6, 50e-6, 1)
    out.append(("lattice", x, y, np.full(2048, Q0)))
    # a compact cloud: every pair inside the cutoff
    x, y = hexagonal(1024, 60e-6, 6e-6, 2)
    out.append(("dense", x, y, np.full(1024, Q0)))
    # uneven charges of both signs, random positions
    n = 1500
    x, y = rng.uniform(-0.02, 0.02, (2, n))
    q = Q0 * rng.uniform(-1.2, 1.2, n)
    out.append(("mixed", x, y, q))
    # pairs at the cutoff and one ulp either side of it, in every lane role
    r_cut = 10.0 / KAPPA
    base_x, base_y = hexagonal(512, 700e-6, 0.0, 3)
    x, y = base_x.copy(), base_y.copy()
    for k, (a, b) in enumerate(((0, 1), (2, 5), (7, 4), (10, 300), (301, 11), (40, 41))):
        r = (np.nextafter(r_cut, 0), r_cut, np.nextafter(r_cut, 1))[k % 3]
        x[b], y[b] = x[a] + r, y[a]
    out.append(("cutoff", x, y, np.full(512, Q0)))
    return out


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_cells_equal_all_pairs_bitwise(oracle, case):
    _, x, y, q = case
    rng = np.random.default_rng(4)
    # forces already holding other plugins' terms (the stage sum adds onto them)
    f0x, f0y = rng.normal(0, 1e-12, (2, len(x)))
    ax, ay = oracle.yukawa(x, y, q, KAPPA, f0x.copy(), f0y.copy())
    cx, cy = oracle.yukawa_cells(x, y, q, KAPPA, f0x.copy(), f0y.copy())
    assert bits_equal(ax, cx) and bits_equal(ay, cy)


def test_cells_rejects_odd_n(oracle):
    x, y = hexagonal(9, 500e-6, 0.0, 1)
    with pytest.raises(AssertionError):
        oracle.yukawa_cells(x, y, np.full(9, Q0), KAPPA)


def test_cells_dtcheck_equals_all_pairs(oracle):
    rng = np.random.default_rng(8)
    for trial in range(12):
        n = 4 * int(rng.integers(8, 64))
        x, y = hexagonal(n, 400e-6, 40e-6, 10 + trial)
        # close pairs: covered ones, the coverage hole (3, 4), and none
        for a, b, r in ((0, 1, 80e-6), (3, 4, 20e-6), (5, n - 2, 9e-6), (8, 13, 0.7e-6))[: trial % 5]:
            x[b], y[b] = x[a] + r * np.cos(trial), y[a] + r * np.sin(trial)
        s_all = oracle.dtcheck_min_s(x, y)
        s_cell = oracle.dtcheck_min_s_cells(x, y)
        if s_all <= np.float32(1.0001e-8):
            assert s_cell == s_all
        else:  # beyond the reach of the check: no reduction either way
            assert s_cell >= s_all
        assert oracle.modify_time_step_cells(x, y, 1e-4, 1e-3) == oracle.modify_time_step(x, y, 1e-4, 1e-3)


def test_cells_trajectory_equals_all_pairs_bitwise(oracle):
    """12 RK4 steps with the dt check, a pair 50 um apart so that dt is cut,
    through oracle_rk4_steps with the all-pairs and the cell-list forms."""
    n = 1024
    x, y = hexagonal(n, 500e-6, 50e-6, 7)
    x[1], y[1] = x[0] + 50e-6, y[0]
    q, m, z = np.full(n, Q0), np.full(n, M0), np.zeros(n)
    runs = []
    for gather in (0, 2):
        op = OracleParams(KAPPA, CONFINE, 2.0, 0.0, 0, 1 | 2 | 4, 1, gather)
        runs.append(oracle.rk4_steps(x, y, z, z, q, m, op, 1e-3, 12))
    a, b = runs
    assert a[4] < 12e-3  # dt was cut
    for u, v in zip(a[:4], b[:4]):
        assert bits_equal(u, v)
    assert a[4] == b[4] and a[5] == b[5]
