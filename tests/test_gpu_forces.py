"""GPU parity of the per-stage forces (Force::force1..4) against the oracle.

Bar (north star): FP64 forces within 1e-12 relative per evaluation, measured
condition-aware: |dF_i| / (sum_j |F_ij| + |F_conf,i| + |F_drag,i|).
FP32-pair variant: within 1e-5 relative on the same scale.
"""
import numpy as np
import pytest

import demon_simulation_code_amd as dm
from oracle.bind import OracleParams
from tests.helpers import CONFINE, GAMMA, KAPPA, M0, Q0, charges, force_rel_err, lattice, pair_abs_sum

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-5


MODES = [dm.PAIRS_NEWTON3, dm.PAIRS_GATHER]


def _ctx(n, q, forces=dm.CONFINEMENT | dm.DRAG | dm.SHIELDED_COULOMB, precision=dm.FP64,
         heat=0.0, x=None, y=None, shape=None, mode=None):
    c = dm.Context(n, shielding=KAPPA, confine=CONFINE, gamma=GAMMA, heat=heat, seed=11,
                   forces=forces, precision=precision)
    if x is None:
        x, y = lattice(n)
    c.upload(x, y, None, None, q, np.full(n, M0))
    if mode is not None:
        c.set_pair_mode(mode)
    if shape:
        c.set_pair_shape(*shape)
    return c


def test_default_mode_by_size():
    """Gather for small clouds (too few superblock rows for the Newton-3
    kernel to fill the chip), Newton-3 from about 6000 particles on."""
    with _ctx(1024, charges(1024)) as c:
        assert c.pair_mode() == dm.PAIRS_GATHER
    with _ctx(8192, charges(8192)) as c:
        assert c.pair_mode() == dm.PAIRS_NEWTON3


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("n", [2, 3, 64, 255, 256, 1000, 1024, 4096])
@pytest.mark.parametrize("stage", [1, 4])
def test_yukawa_stage_fp64(oracle, n, stage, mode):
    x, y = lattice(n, seed=n)
    q = charges(n)
    with _ctx(n, q, x=x, y=y, mode=mode) as c:
        gx, gy = c.force_stage(dm.SHIELDED_COULOMB, stage, 0, x, y)
    fx, fy = oracle.yukawa(x, y, q, KAPPA)
    scale = pair_abs_sum(x, y, q, KAPPA) + 1e-300
    assert force_rel_err(fx, fy, gx, gy, scale) <= FP64_TOL


@pytest.mark.parametrize("mode,shape", [(dm.PAIRS_GATHER, (1, 1)), (dm.PAIRS_GATHER, (1, 7)),
                                        (dm.PAIRS_GATHER, (2, 3)), (dm.PAIRS_GATHER, (4, 1)),
                                        (dm.PAIRS_GATHER, (4, 64)), (dm.PAIRS_NEWTON3, (1, 1)),
                                        (dm.PAIRS_NEWTON3, (1, 2)), (dm.PAIRS_NEWTON3, (1, 5))])
def test_yukawa_pair_shapes(oracle, mode, shape):
    n = 3000
    x, y = lattice(n, seed=5)
    q = charges(n)
    with _ctx(n, q, x=x, y=y, shape=shape, mode=mode) as c:
        gx, gy = c.force_stage(dm.SHIELDED_COULOMB, 2, 0, x, y)
    fx, fy = oracle.yukawa(x, y, q, KAPPA)
    assert force_rel_err(fx, fy, gx, gy, pair_abs_sum(x, y, q, KAPPA)) <= FP64_TOL


@pytest.mark.parametrize("n,shape,slices", [(1000, (1, 0), 64), (1000, (1, 40), 32), (1000, (2, 0), 64),
                                            (1000, (4, 30), 16), (2000, (1, 0), 64), (2000, (2, 48), 32),
                                            (2000, (1, 20), 16), (2000, (4, 0), 64)])
@pytest.mark.parametrize("precision,tol", [(dm.FP64, FP64_TOL), (dm.FP32_PAIR, FP32_TOL)])
def test_yukawa_short_j_slices(oracle, n, shape, slices, precision, tol):
    """Gather kernel with j slices of a quarter, a half or a whole box block
    (16, 32 or 64 j; N = 1000 pads to 1024, N = 2000 to 2048, the last block
    partly padding): the requested slice count is rounded to such lengths."""
    x, y = lattice(n, seed=7)
    q = charges(n)
    with _ctx(n, q, x=x, y=y, shape=shape, mode=dm.PAIRS_GATHER, precision=precision) as c:
        assert c.pair_shape() == (shape[0], slices)
        gx, gy = c.force_stage(dm.SHIELDED_COULOMB, 2, 0, x, y)
    fx, fy = oracle.yukawa(x, y, q, KAPPA)
    assert force_rel_err(fx, fy, gx, gy, pair_abs_sum(x, y, q, KAPPA)) <= tol


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("precision,tol", [(dm.FP64, FP64_TOL), (dm.FP32_PAIR, FP32_TOL)])
def test_yukawa_dense_cluster(oracle, mode, precision, tol):
    """A compact cluster (60 um lattice, 2.1 mm across, inside one cutoff
    radius): every pair contributes, every block pair takes the straight-line
    path, and the Newton-3 kernel evaluates every tile."""
    n = 4096
    x, y = lattice(n, seed=9, spacing=60e-6, jitter=6e-6)
    q = charges(n)
    with _ctx(n, q, x=x, y=y, mode=mode, precision=precision) as c:
        gx, gy = c.force_stage(dm.SHIELDED_COULOMB, 1, 0, x, y)
        near, tested = c.pair_counts()
    fx, fy = oracle.yukawa(x, y, q, KAPPA)
    assert near == n * (n - 1)  # every ordered pair contributes
    assert tested == n * n      # no block pair culled (the diagonal blocks include i == j)
    assert force_rel_err(fx, fy, gx, gy, pair_abs_sum(x, y, q, KAPPA)) <= tol


def test_all_stage_forces_fp64(oracle):
    n = 1024
    x, y = lattice(n, seed=3)
    rng = np.random.default_rng(4)
    vx, vy = rng.normal(0, 1e-2, n), rng.normal(0, 1e-2, n)
    q = charges(n)
    m = np.full(n, M0) * (1 + 0.05 * rng.uniform(-1, 1, n))
    forces = dm.CONFINEMENT | dm.DRAG | dm.SHIELDED_COULOMB | dm.THERMAL
    heat = 3.0e-15
    c = dm.Context(n, shielding=KAPPA, confine=CONFINE, gamma=GAMMA, heat=heat, seed=99, forces=forces)
    with c:
        c.upload(x, y, vx, vy, q, m)
        for stage in (1, 2, 3, 4):
            gx, gy = c.force_stage(forces, stage, 17, x, y, vx, vy)
            op = OracleParams(KAPPA, CONFINE, GAMMA, heat, 99, forces, 1)
            fx, fy = oracle.stage_force(x, y, vx, vy, q, m, op, step=17, stage=stage)
            scale = (pair_abs_sum(x, y, q, KAPPA) + np.abs(q * CONFINE) * np.hypot(x, y)
                     + m * GAMMA * np.hypot(vx, vy) + heat)
            assert force_rel_err(fx, fy, gx, gy, scale) <= FP64_TOL, stage


def test_force_stage_adds_into_arrays(oracle):
    """Force::forceS adds into Cloud::forceX/Y; it never overwrites."""
    n = 256
    x, y = lattice(n)
    q = charges(n)
    base_x = np.linspace(-1e-12, 1e-12, n)
    base_y = np.linspace(1e-12, -1e-12, n)
    with _ctx(n, q, x=x, y=y) as c:
        gx, gy = c.force_stage(dm.SHIELDED_COULOMB, 1, 0, x, y, fx=base_x.copy(), fy=base_y.copy())
    fx, fy = oracle.yukawa(x, y, q, KAPPA, fx=base_x.copy(), fy=base_y.copy())
    scale = pair_abs_sum(x, y, q, KAPPA) + np.abs(base_x) + np.abs(base_y)
    assert force_rel_err(fx, fy, gx, gy, scale) <= FP64_TOL


@pytest.mark.parametrize("mode", MODES)
def test_cutoff_boundary_exact(oracle, mode):
    """Pairs placed 1 ulp either side of u = 10 are classified as the reference does."""
    lam = 1.0 / KAPPA
    r_cut = 10.0 * lam
    xs, ys = [], []
    for k in range(16):
        base = 0.05 * k
        r = r_cut
        for _ in range(k % 4):
            r = np.nextafter(r, np.inf if k % 2 else -np.inf)
        xs += [base, base + r]
        ys += [0.0, 0.0]
    x, y = np.array(xs), np.array(ys)
    n = len(x)
    q = np.full(n, Q0)
    with _ctx(n, q, x=x, y=y, mode=mode) as c:
        gx, gy = c.force_stage(dm.SHIELDED_COULOMB, 1, 0, x, y)
    fx, fy = oracle.yukawa(x, y, q, KAPPA)
    np.testing.assert_array_equal(fx == 0.0, gx == 0.0)
    scale = pair_abs_sum(x, y, q, KAPPA) + 1e-300
    assert force_rel_err(fx, fy, gx, gy, scale) <= FP64_TOL


@pytest.mark.parametrize("mode", MODES)
def test_newton3_full_size(mode):
    """N = 131072: sum of Yukawa forces vanishes (Newton's third law) to rounding."""
    n = 131072
    x, y = lattice(n)
    q = charges(n)
    with _ctx(n, q, x=x, y=y, mode=mode) as c:
        gx, gy = c.force_stage(dm.SHIELDED_COULOMB, 1, 0, x, y)
    rows = np.random.default_rng(0).choice(n, 512, replace=False)
    scale = pair_abs_sum(x, y, q, KAPPA, rows).mean() * n
    assert abs(gx.sum()) / scale < 1e-12 and abs(gy.sum()) / scale < 1e-12


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("spacing", [500e-6, 12e-6])
@pytest.mark.parametrize("precision,tol", [(dm.FP64, FP64_TOL), (dm.FP32_PAIR, FP32_TOL)])
def test_sampled_rows_full_size(oracle, mode, spacing, precision, tol):
    """N = 131072 (the large-cloud size, in both precisions of the pair
    arithmetic): 512 sampled particles against the oracle's FP64 gather, on
    the benchmark lattice (500 um: most block pairs culled) and on a lattice
    as compact as the imploded benchmark cloud (12 um: 4.6 mm across, every
    pair inside the cutoff).  FP32-pair: within 1e-5 of the FP64 oracle."""
    n = 131072
    x, y = lattice(n, spacing=spacing, jitter=0.1 * spacing)
    q = charges(n)
    rows = np.random.default_rng(1).choice(n, 512, replace=False)
    with _ctx(n, q, x=x, y=y, mode=mode, precision=precision) as c:
        gx, gy = c.force_stage(dm.SHIELDED_COULOMB, 1, 0, x, y)
    fx, fy = oracle.yukawa_rows(x, y, q, KAPPA, rows)
    scale = pair_abs_sum(x, y, q, KAPPA, rows)
    err = force_rel_err(fx, fy, gx[rows], gy[rows], scale)
    print(f"N={n} spacing={spacing * 1e6:.0f} um precision={precision} mode={mode}: "
          f"max condition-aware force error {err:.2e} (bar {tol:.0e})")
    assert err <= tol
    if precision == dm.FP32_PAIR:
        assert err > 1e-12  # it really is the FP32 path


def test_newton3_equals_gather_full_size():
    """N = 131072 on a 40 um lattice (mixed culled, straight-line and
    prefiltered tiles): the two kernels agree to rounding on every particle."""
    n = 131072
    x, y = lattice(n, spacing=40e-6, jitter=4e-6)
    q = charges(n)
    out = []
    for mode in MODES:
        with _ctx(n, q, x=x, y=y, mode=mode) as c:
            out.append(c.force_stage(dm.SHIELDED_COULOMB, 1, 0, x, y))
    rows = np.random.default_rng(2).choice(n, 256, replace=False)
    scale = pair_abs_sum(x, y, q, KAPPA, rows).mean()
    d = np.hypot(out[0][0] - out[1][0], out[0][1] - out[1][1])
    assert d.max() / scale <= 1e-12


@pytest.mark.parametrize("n", [1024, 4096])
def test_yukawa_fp32_pair(oracle, n):
    x, y = lattice(n, seed=2)
    q = charges(n)
    with _ctx(n, q, x=x, y=y, precision=dm.FP32_PAIR) as c:
        gx, gy = c.force_stage(dm.SHIELDED_COULOMB, 1, 0, x, y)
    fx, fy = oracle.yukawa(x, y, q, KAPPA)
    err = force_rel_err(fx, fy, gx, gy, pair_abs_sum(x, y, q, KAPPA))
    assert err <= FP32_TOL
    assert err > 0.0  # it really is the FP32 path


def test_magnetic_and_time_varying_drag_stage_forces(oracle):
    """MagneticForce and TimeVaryingDragForce at a stage time, in the driver's
    order: bit-exact with the oracle without the pair sums, and within the FP64
    bar with them and the thermal force (Force::forceS(t) through
    demon_force_stage_at)."""
    n = 1024
    x, y = lattice(n, seed=5)
    rng = np.random.default_rng(6)
    vx, vy = rng.normal(0, 1e-2, n), rng.normal(0, 1e-2, n)
    q = charges(n)
    m = np.full(n, M0) * (1 + 0.05 * rng.uniform(-1, 1, n))
    bfield, scale_d = 0.9, -2.5
    elem = dm.CONFINEMENT | dm.MAGNETIC | dm.TIME_VARYING_DRAG
    full = elem | dm.SHIELDED_COULOMB | dm.THERMAL | dm.DRAG
    heat = 2.0e-15
    c = dm.Context(n, shielding=KAPPA, confine=CONFINE, gamma=GAMMA, heat=heat, seed=7, forces=full,
                   magnetic_field=bfield, drag_scale=scale_d)
    with c:
        c.upload(x, y, vx, vy, q, m)
        for stage, t in ((1, 0.25), (2, 0.2505), (4, 0.251)):
            gx, gy = c.force_stage(elem, stage, 3, x, y, vx, vy, t=t)
            op = OracleParams(KAPPA, CONFINE, GAMMA, heat, 7, elem, 1, 0, bfield, scale_d)
            fx, fy = oracle.stage_force(x, y, vx, vy, q, m, op, step=3, stage=stage, t=t)
            assert np.array_equal(fx, gx) and np.array_equal(fy, gy), stage
            gx, gy = c.force_stage(full, stage, 3, x, y, vx, vy, t=t)
            op = OracleParams(KAPPA, CONFINE, GAMMA, heat, 7, full, 1, 0, bfield, scale_d)
            fx, fy = oracle.stage_force(x, y, vx, vy, q, m, op, step=3, stage=stage, t=t)
            scale = (pair_abs_sum(x, y, q, KAPPA) + np.abs(q * CONFINE) * np.hypot(x, y)
                     + m * (GAMMA + abs(scale_d * t + GAMMA)) * np.hypot(vx, vy)
                     + np.abs(q * bfield) * np.hypot(vx, vy) + heat)
            assert force_rel_err(fx, fy, gx, gy, scale) <= FP64_TOL, stage
        with pytest.raises(dm.DemonError):
            c.force_stage(dm.TIME_VARYING_DRAG, 1, 0, x, y, vx, vy)  # needs the stage time


@pytest.mark.parametrize("mode", MODES)
def test_near_pair_count_is_exact(oracle, mode):
    """The in-kernel count of ordered pairs inside the cutoff (the P_near of
    the roofline's FLOP count) equals the oracle's count under the
    reference's rounding, on a sparse lattice where most tiles are culled and
    on two evaluations in a row (the counts add up)."""
    n = 4096
    x, y = lattice(n, seed=21)
    q = charges(n)
    want = oracle.near_pairs(x, y, KAPPA)
    assert 0 < want < n * (n - 1) // 4
    with _ctx(n, q, x=x, y=y, mode=mode) as c:
        c.force_stage(dm.SHIELDED_COULOMB, 1, 0, x, y)
        near1, tested1 = c.pair_counts()
        c.force_stage(dm.SHIELDED_COULOMB, 2, 0, x, y)
        near2, tested2 = c.pair_counts()
    assert near1 == want and near2 == 2 * want
    assert tested2 == 2 * tested1 and want <= tested1 < n * n
