"""GPU parity of the further force plugins (RectConfinement, ConfinementVoid,
Rotational, Driving, Electric, VertElectric, Gravitational,
ThermalLocalized, TimeVaryingThermal) against the oracle, per stage force and
inside Runge_Kutta4 / Runge_Kutta2 with the dt check.

Bars: per evaluation, the elementwise plugins within 1e-13 of the sum of
|terms| (the device's exp, sin and sqrt may differ from the host libm's in
the last bit), the linear ones (Rect, Gravitational) bit for bit; positions
within 1e-9 relative after the runs, with identical step counts and times.
"""
import numpy as np
import pytest

import demon_simulation_code_amd as dm
from oracle.bind import OracleParams
from tests.helpers import CONFINE, GAMMA, KAPPA, M0, charges, force_rel_err, lattice, pair_abs_sum, pos_rel_err

pytestmark = pytest.mark.gpu

ELEM_TOL = 1e-13
POS_TOL = 1e-9

# constants for a 1024-particle lattice of 500 um spacing (radius ~8 mm): each
# plugin is comparable to the confinement force somewhere in the cloud
PLUGIN = dict(confine_x=CONFINE, confine_y=2 * CONFINE, void_decay=300.0, rot_inner=2e-3, rot_outer=6e-3,
              rot_const=2e-14, drive_const=1e-5, drive_amplitude=2e-14, drive_shift=7e-3,
              electric_field=-0.5 * CONFINE, plasma_radius=5e-3, vert_electric_field=20.0, vert_decay=8e-3,
              gravity=-0.5, heat_in=3e-15, heat_out=1e-15, heat_radius=4e-3, heat_scale=2e-15,
              heat_offset=1e-15)
BFIELD, DRAG_SCALE, HEAT, SEED = 0.8, -1.5, 2e-15, 9


def oracle_params(forces):
    op = OracleParams(KAPPA, CONFINE, GAMMA, HEAT, SEED, forces, 1, 0, BFIELD, DRAG_SCALE)
    for k, v in PLUGIN.items():
        setattr(op, k, v)
    return op


def context(n, forces):
    return dm.Context(n, shielding=KAPPA, confine=CONFINE, gamma=GAMMA, heat=HEAT, seed=SEED, forces=forces,
                      magnetic_field=BFIELD, drag_scale=DRAG_SCALE, **PLUGIN)


def state(n, seed):
    x, y = lattice(n, seed=seed)
    rng = np.random.default_rng(seed)
    vx, vy = rng.normal(0, 5e-3, (2, n))
    q = charges(n, seed=seed)
    m = np.full(n, M0) * (1 + 0.05 * rng.uniform(-1, 1, n))
    return x, y, vx, vy, q, m


def term_scale(x, y, vx, vy, q, m):
    """A bound of the sum of |terms| of every elementwise plugin here."""
    r = np.hypot(x, y)
    p = PLUGIN
    return (np.abs(q) * (CONFINE * 2 + p["confine_y"] + abs(p["electric_field"])) * r
            + np.abs(q) * p["void_decay"] * np.exp(-p["void_decay"] * r)
            + np.abs(q) * p["vert_electric_field"] * np.exp(y / p["vert_decay"])
            + m * (GAMMA + 5.0) * np.hypot(vx, vy) + np.abs(q * BFIELD) * np.hypot(vx, vy)
            + p["rot_const"] + p["drive_amplitude"] + m * abs(p["gravity"]) + 1e-14)


@pytest.mark.parametrize("forces", [dm.RECT_CONFINEMENT, dm.GRAVITATIONAL, dm.RECT_CONFINEMENT | dm.GRAVITATIONAL])
def test_linear_plugins_bit_exact(oracle, forces):
    n = 1024
    x, y, vx, vy, q, m = state(n, 21)
    with context(n, forces) as c:
        c.upload(x, y, vx, vy, q, m)
        for stage in (1, 3):
            gx, gy = c.force_stage(forces, stage, 0, x, y, vx, vy, t=0.4)
            fx, fy = oracle.stage_force(x, y, vx, vy, q, m, oracle_params(forces), stage=stage, t=0.4)
            assert np.array_equal(fx, gx) and np.array_equal(fy, gy)


@pytest.mark.parametrize("forces", [dm.CONFINEMENT_VOID, dm.ROTATIONAL, dm.DRIVING, dm.ELECTRIC, dm.VERT_ELECTRIC,
                                    dm.CONFINEMENT | dm.CONFINEMENT_VOID,
                                    dm.CONFINEMENT_VOID | dm.DRAG | dm.DRIVING | dm.MAGNETIC | dm.RECT_CONFINEMENT
                                    | dm.ROTATIONAL | dm.TIME_VARYING_DRAG | dm.ELECTRIC | dm.GRAVITATIONAL
                                    | dm.VERT_ELECTRIC])
def test_elementwise_plugins(oracle, forces):
    """Each plugin (and all of them together, in the driver's order) at
    three stage times against the oracle."""
    n = 1024
    x, y, vx, vy, q, m = state(n, 22)
    scale = term_scale(x, y, vx, vy, q, m)
    with context(n, forces) as c:
        c.upload(x, y, vx, vy, q, m)
        for stage, t in ((1, 0.3), (2, 0.3005), (4, 0.301)):
            gx, gy = c.force_stage(forces, stage, 2, x, y, vx, vy, t=t)
            fx, fy = oracle.stage_force(x, y, vx, vy, q, m, oracle_params(forces), step=2, stage=stage, t=t)
            assert np.abs(fx).max() + np.abs(fy).max() > 0
            assert force_rel_err(fx, fy, gx, gy, scale) <= ELEM_TOL, stage
    if forces == dm.ROTATIONAL:
        r = np.hypot(x, y)
        outside = (r <= PLUGIN["rot_inner"]) | (r >= PLUGIN["rot_outer"])
        assert outside.any() and np.all(gx[outside] == 0) and np.all(gy[outside] == 0)


@pytest.mark.parametrize("forces", [dm.THERMAL_LOCALIZED, dm.TIME_VARYING_THERMAL,
                                    dm.THERMAL | dm.THERMAL_LOCALIZED | dm.TIME_VARYING_THERMAL])
def test_thermal_plugins_share_the_oracle_draws(oracle, forces):
    """ThermalLocalized and TimeVaryingThermal draw from their own Philox
    domains: the device's kicks are the oracle's to the last bits of cos/sin,
    and three thermal plugins together are three independent kicks."""
    n = 1024
    x, y, vx, vy, q, m = state(n, 23)
    with context(n, forces) as c:
        c.upload(x, y, vx, vy, q, m)
        for stage, step, t in ((1, 0, 0.2), (3, 7, 1.5)):
            gx, gy = c.force_stage(forces, stage, step, x, y, vx, vy, t=t)
            fx, fy = oracle.stage_force(x, y, vx, vy, q, m, oracle_params(forces), step=step, stage=stage, t=t)
            hmax = HEAT + PLUGIN["heat_in"] + PLUGIN["heat_scale"] * t + PLUGIN["heat_offset"]
            assert np.max(np.hypot(fx - gx, fy - gy)) <= 1e-14 * hmax
            assert np.max(np.hypot(gx, gy)) <= hmax
    if forces == dm.THERMAL_LOCALIZED:
        inside = np.hypot(x, y) < PLUGIN["heat_radius"]
        assert np.max(np.hypot(gx, gy)[inside]) <= PLUGIN["heat_in"]
        assert np.max(np.hypot(gx, gy)[~inside]) <= PLUGIN["heat_out"]
        assert np.max(np.hypot(gx, gy)[inside]) > PLUGIN["heat_out"]


def test_time_dependent_plugins_need_the_stage_time():
    n = 256
    x, y, vx, vy, q, m = state(n, 24)
    with context(n, dm.DRIVING | dm.TIME_VARYING_THERMAL | dm.CONFINEMENT) as c:
        c.upload(x, y, vx, vy, q, m)
        for f in (dm.DRIVING, dm.TIME_VARYING_THERMAL):
            with pytest.raises(dm.DemonError):
                c.force_stage(f, 1, 0, x, y, vx, vy)
        c.force_stage(dm.CONFINEMENT, 1, 0, x, y, vx, vy)


RUN_A = (dm.CONFINEMENT_VOID | dm.DRAG | dm.DRIVING | dm.ROTATIONAL | dm.SHIELDED_COULOMB | dm.THERMAL_LOCALIZED
         | dm.GRAVITATIONAL)
RUN_B = (dm.MAGNETIC | dm.RECT_CONFINEMENT | dm.SHIELDED_COULOMB | dm.TIME_VARYING_DRAG | dm.TIME_VARYING_THERMAL
         | dm.ELECTRIC | dm.VERT_ELECTRIC)
RUNS = [
    # -V -w -S -G -L: the void replaces the confinement; the localized heating
    (RUN_A, 4, False),
    # -R -B -E -F -v -D: rectangular trap with the fields and the time-varying heat and drag
    (RUN_B, 4, False),
    (RUN_A, 2, False),
    # the captured step (hipGraph) reads the stage times on the device
    (RUN_B, 4, True),
    (RUN_A, 2, True),
]


@pytest.mark.parametrize("forces,order,graph", RUNS)
def test_plugin_trajectories(oracle, forces, order, graph):
    n = 1024
    x, y, vx, vy, q, m = state(n, 25)
    steps, t0 = 60, 0.5
    with context(n, forces) as c:
        c.upload(x, y, vx, vy, q, m)
        c.use_graph(graph)
        t = (c.rk4_steps if order == 4 else c.rk2_steps)(steps, 1e-3, t0)
        g = c.download()
        assert c.stats().steps == steps
    run = oracle.rk4_steps if order == 4 else oracle.rk2_steps
    o = run(x, y, vx, vy, q, m, oracle_params(forces), 1e-3, steps, t=t0)
    assert t == o[4]
    assert pos_rel_err(o[0], o[1], g[0], g[1]) <= POS_TOL
    vscale = np.max(np.hypot(o[2], o[3]))
    assert np.max(np.hypot(o[2] - g[2], o[3] - g[3])) / vscale <= 1e-8
    # the run moved the cloud: the plugins were not no-ops
    assert np.max(np.hypot(o[0] - x, o[1] - y)) > 1e-6


def test_elementwise_plugins_full_size(oracle):
    """Every elementwise plugin at N = 131072 (the benchmark's size), one
    stage force against the oracle."""
    n = 131072
    x, y, vx, vy, q, m = state(n, 26)
    forces = (dm.CONFINEMENT_VOID | dm.DRAG | dm.DRIVING | dm.MAGNETIC | dm.RECT_CONFINEMENT | dm.ROTATIONAL
              | dm.TIME_VARYING_DRAG | dm.ELECTRIC | dm.GRAVITATIONAL | dm.VERT_ELECTRIC | dm.THERMAL_LOCALIZED
              | dm.TIME_VARYING_THERMAL)
    with context(n, forces) as c:
        c.upload(x, y, vx, vy, q, m)
        gx, gy = c.force_stage(forces, 3, 11, x, y, vx, vy, t=0.75)
    fx, fy = oracle.stage_force(x, y, vx, vy, q, m, oracle_params(forces), step=11, stage=3, t=0.75)
    scale = term_scale(x, y, vx, vy, q, m) + PLUGIN["heat_in"] + PLUGIN["heat_scale"] * 0.75 + PLUGIN["heat_offset"]
    assert force_rel_err(fx, fy, gx, gy, scale) <= ELEM_TOL


def test_plugins_ragged_sizes(oracle):
    """Sizes that fill no block evenly: the elementwise plugins at n = 1001 and
    a 20-step RK4 run of all of them at n = 1000 (the dt check needs n % 4 = 0
    in the reference)."""
    forces = (dm.CONFINEMENT_VOID | dm.DRAG | dm.DRIVING | dm.RECT_CONFINEMENT | dm.ROTATIONAL | dm.ELECTRIC
              | dm.GRAVITATIONAL | dm.VERT_ELECTRIC | dm.THERMAL_LOCALIZED)
    x, y, vx, vy, q, m = state(1001, 27)
    with context(1001, forces) as c:
        c.upload(x, y, vx, vy, q, m)
        gx, gy = c.force_stage(forces, 2, 4, x, y, vx, vy, t=0.1)
    fx, fy = oracle.stage_force(x, y, vx, vy, q, m, oracle_params(forces), step=4, stage=2, t=0.1)
    scale = term_scale(x, y, vx, vy, q, m) + PLUGIN["heat_in"]
    assert force_rel_err(fx, fy, gx, gy, scale) <= ELEM_TOL
    n = 1000
    x, y, vx, vy, q, m = state(n, 28)
    full = forces | dm.SHIELDED_COULOMB
    with context(n, full) as c:
        c.upload(x, y, vx, vy, q, m)
        t = c.rk4_steps(20, 1e-3, 0.0)
        g = c.download()
    o = oracle.rk4_steps(x, y, vx, vy, q, m, oracle_params(full), 1e-3, 20)
    assert t == o[4]
    assert pos_rel_err(o[0], o[1], g[0], g[1]) <= POS_TOL
