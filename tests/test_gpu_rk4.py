"""GPU parity of Runge_Kutta4::moveParticles (with the dt check) against the oracle.

Bar (north star): positions within 1e-9 relative after 100 steps, thermal off
(max_i |dr_i| / max_i |r_i|).  Step counts and dt reductions must match exactly.
"""
import numpy as np
import pytest

import demon_simulation_code_amd as dm
from oracle.bind import OracleParams
from tests.helpers import CONFINE, GAMMA, KAPPA, M0, charges, lattice, pos_rel_err

pytestmark = pytest.mark.gpu

FORCES = dm.CONFINEMENT | dm.DRAG | dm.SHIELDED_COULOMB
POS_TOL = 1e-9


def _run_both(oracle, x, y, q, m, steps, dt=1e-3, shape=None, graph=False, forces=FORCES, heat=0.0,
              mode=None):
    n = len(x)
    vx = np.zeros(n)
    vy = np.zeros(n)
    with dm.Context(n, shielding=KAPPA, confine=CONFINE, gamma=GAMMA, heat=heat, seed=5,
                    forces=forces) as c:
        c.upload(x, y, vx, vy, q, m)
        if mode is not None:
            c.set_pair_mode(mode)
        if shape:
            c.set_pair_shape(*shape)
        c.use_graph(graph)
        t = c.rk4_steps(steps, dt, 0.0)
        g = c.download()
        st = c.stats()
    op = OracleParams(KAPPA, CONFINE, GAMMA, heat, 5, forces, 1)
    o = oracle.rk4_steps(x, y, vx, vy, q, m, op, dt, steps)
    return g, t, st, o


@pytest.mark.parametrize("n,graph,mode", [(256, False, None), (1024, False, None), (1024, True, None),
                                          (1000, False, None), (1024, False, dm.PAIRS_GATHER),
                                          (1024, True, dm.PAIRS_GATHER)])
def test_rk4_100_steps(oracle, n, graph, mode):
    x, y = lattice(n, seed=n)
    q = charges(n)
    m = np.full(n, M0)
    g, t, st, o = _run_both(oracle, x, y, q, m, 100, graph=graph, mode=mode)
    assert st.steps == 100
    assert t == o[4]
    assert pos_rel_err(o[0], o[1], g[0], g[1]) <= POS_TOL
    vscale = np.max(np.hypot(o[2], o[3]))
    assert np.max(np.hypot(o[2] - g[2], o[3] - g[3])) / vscale <= 1e-8


def test_dt_check_reduction_and_hole(oracle):
    """A covered close pair divides dt by 10; pair (3, 4) sits in the coverage hole."""
    n = 64
    x, y = lattice(n, seed=9)
    q = charges(n, spread=0.0)
    m = np.full(n, M0)
    cases = [((0, 1), 50e-6, 1e-4), ((3, 4), 50e-6, 1e-3), ((10, 40), 5e-6, 1e-5), ((5, 6), 0.9e-6, 1e-6)]
    for (a, b), sep, want_dt in cases:
        xx, yy = x.copy(), y.copy()
        xx[b] = xx[a] + sep
        yy[b] = yy[a]
        assert oracle.modify_time_step(xx, yy, 1e-4, 1e-3) == pytest.approx(want_dt, rel=1e-12)
        g, t, st, o = _run_both(oracle, xx, yy, q, m, 1)
        assert st.last_dt == oracle.modify_time_step(xx, yy, 1e-4, 1e-3)
        assert t == o[4]
        assert pos_rel_err(o[0], o[1], g[0], g[1]) <= POS_TOL


def test_rk4_move_loop_semantics(oracle):
    """moveParticles(end) runs while (t < end) with double accumulation of t;
    the step count and final time must equal the oracle's (100 steps to 0.1 s)."""
    n = 256
    x, y = lattice(n, seed=3)
    q = charges(n)
    m = np.full(n, M0)
    vx = np.zeros(n)
    op = OracleParams(KAPPA, CONFINE, GAMMA, 0.0, 5, FORCES, 1)
    o = oracle.rk4_move(x, y, vx, vx, q, m, op, 1e-3, 0.1)
    with dm.Context(n, shielding=KAPPA, confine=CONFINE, gamma=GAMMA, forces=FORCES) as c:
        c.upload(x, y, vx, vx, q, m)
        t, steps = c.rk4_move(0.0, 0.1, 1e-3)
        g = c.download()
    assert steps == o[5] == 100
    assert t == o[4]
    assert pos_rel_err(o[0], o[1], g[0], g[1]) <= POS_TOL


def test_thermal_rk4_matches_oracle_draws(oracle):
    """Thermal on: the counter-based draws are the oracle's, so a short run still matches."""
    n = 256
    x, y = lattice(n, seed=4)
    q = charges(n)
    m = np.full(n, M0)
    forces = FORCES | dm.THERMAL
    g, t, st, o = _run_both(oracle, x, y, q, m, 20, forces=forces, heat=3e-15)
    assert pos_rel_err(o[0], o[1], g[0], g[1]) <= POS_TOL


def test_rk4_state_is_finite_at_full_size():
    n = 131072
    x, y = lattice(n)
    q = charges(n)
    m = np.full(n, M0)
    with dm.Context(n, shielding=KAPPA, confine=CONFINE, gamma=GAMMA, forces=FORCES) as c:
        c.upload(x, y, None, None, q, m)
        t = c.rk4_steps(3, 1e-3, 0.0)
        g = c.download()
    assert t == pytest.approx(3e-3)
    assert all(np.isfinite(a).all() for a in g)
