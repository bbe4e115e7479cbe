"""Writes tests/golden/*.npz from the reference's own code.

The reference (AU-PSL/demon_simulation_code) ships no tests, fixtures or
golden vectors.  This script compiles its hot-path sources in place (serial
build, oracle/Makefile target `ref`) and records, for fixed inputs stored in
the same files:
  forces_n{N}.npz   ShieldedCoulombForce::force1 and ::force3 on the inputs
  rk4_n{N}.npz      Runge_Kutta4 (Confinement + Drag + ShieldedCoulomb, with
                    Integrator::modifyTimeStep) after 1, 10 and 100 steps
  dtcheck.npz       Integrator::modifyTimeStep on constructed close pairs
  plugins.npz       the deterministic further force plugins (RectConfinement,
                    ConfinementVoid, Rotational, Driving, Electric,
                    VertElectric, Gravitational) built as the driver builds
                    them: forceS(t) on one state per force set, and 6-step
                    Runge_Kutta4 / Runge_Kutta2 runs with the dt check
Run only where /root/reference exists:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from oracle.bind import REF_PAR, OracleParams, Reference, build  # noqa: E402

E = -1.602e-19
Q0 = 6000 * E
M0 = 2.8e-14
KAPPA = 2000.0
CONFINE = 5e-11 / abs(Q0)
GAMMA = 2.0
DT = 1e-3


def lattice(n, seed):
    """Jittered hexagonal patch (numpy generator, independent of the product's)."""
    a = 500e-6
    L = int(np.sqrt(n)) + 6
    i, j = np.meshgrid(np.arange(-L, L + 1), np.arange(-L, L + 1), indexing="ij")
    i, j = i.ravel(), j.ravel()
    o = np.lexsort((i, j, i * i + i * j + j * j))[:n]
    x = a * (i[o] + 0.5 * j[o])
    y = a * (np.sqrt(3) / 2) * j[o]
    x, y = x - x.mean(), y - y.mean()
    rng = np.random.default_rng(seed)
    return x + rng.uniform(-50e-6, 50e-6, n), y + rng.uniform(-50e-6, 50e-6, n)


def main():
    build(ref=True)
    ref = Reference(openmp=False)
    for n, seed in ((64, 1), (256, 2), (1024, 3)):
        x, y = lattice(n, seed)
        q = Q0 * (1 + 0.1 * np.random.default_rng(seed + 100).uniform(-1, 1, n))
        f1x, f1y = ref.yukawa(x, y, q, KAPPA, stage=1)
        f3x, f3y = ref.yukawa(x, y, q, KAPPA, stage=3)
        np.savez_compressed(HERE / f"forces_n{n}.npz", x=x, y=y, q=q, kappa=KAPPA,
                            f1x=f1x, f1y=f1y, f3x=f3x, f3y=f3y)
        m = M0 * (1 + 0.05 * np.random.default_rng(seed + 200).uniform(-1, 1, n))
        z = np.zeros(n)
        out = dict(x=x, y=y, q=q, m=m, kappa=KAPPA, confine=CONFINE, gamma=GAMMA, dt=DT)
        state = (x, y, z, z)
        done = 0
        steps = (1, 10, 100) if n <= 256 else (1, 10)
        for target in steps:
            *state, t = ref.rk4_steps(*state, q, m, KAPPA, CONFINE, GAMMA, 0.0, 7, DT, target - done,
                                      t=done * DT if done else 0.0)
            done = target
            for name, arr in zip(("x", "y", "vx", "vy"), state):
                out[f"{name}_{target}"] = arr
        np.savez_compressed(HERE / f"rk4_n{n}.npz", **out)
    # dt check on constructed pairs
    x, y = lattice(64, 9)
    cases = []
    for a, b, sep in ((0, 1, 50e-6), (3, 4, 50e-6), (10, 40, 5e-6), (5, 6, 0.9e-6), (0, 5, 30e-6),
                      (2, 3, 99.99e-6), (2, 3, 100.01e-6), (60, 63, 7e-7)):
        xx, yy = x.copy(), y.copy()
        xx[b] = xx[a] + sep
        yy[b] = yy[a]
        cases.append((a, b, sep, ref.modify_time_step(xx, yy, 1e-4, 1e-3)))
    np.savez_compressed(HERE / "dtcheck.npz", x=x, y=y, cases=np.array(cases))
    plugins(ref)
    print("golden fixtures written to", HERE)


PLUGIN_FORCE_SETS = (8192, 2048, 128, 64, 4096, 16384, 8, 1 | 2048)
PLUGIN_RUNS = (2048 | 2 | 4 | 64 | 128 | 8, 8192 | 2 | 4 | 4096 | 16384 | 1024, 1 | 2048 | 4 | 256 | 8)


def plugins(ref):
    rng = np.random.default_rng(77)
    n = 48
    x, y = rng.uniform(-2e-3, 2e-3, (2, n))
    vx, vy = rng.normal(size=(2, n)) * 1e-4
    q = Q0 * rng.uniform(0.5, 1.5, n)
    m = M0 * rng.uniform(0.8, 1.2, n)
    op = OracleParams(3000.0, CONFINE, GAMMA, 0.0, 0, 0, 1, 0, 0.5, -3.0)
    consts = dict(confine_x=CONFINE, confine_y=3 * CONFINE, void_decay=2000.0, rot_inner=5e-4, rot_outer=1.5e-3,
                  rot_const=1e-13, drive_const=1e-5, drive_amplitude=1e-13, drive_shift=7e-3,
                  electric_field=-CONFINE, plasma_radius=2e-3, vert_electric_field=1e2, vert_decay=3e-3,
                  gravity=-9.8)
    for k, v in consts.items():
        setattr(op, k, v)
    out = dict(x=x, y=y, vx=vx, vy=vy, q=q, m=m, par=np.array([getattr(op, k) for k in REF_PAR]),
               par_names=np.array(REF_PAR), force_sets=np.array(PLUGIN_FORCE_SETS), runs=np.array(PLUGIN_RUNS))
    for f in PLUGIN_FORCE_SETS:
        op.forces = f
        fx, fy = ref.force_all(x, y, vx, vy, q, m, op, t=0.37, reps=1)
        out[f"force_{f}_x"], out[f"force_{f}_y"] = fx[0], fy[0]
    for order in (4, 2):
        for f in PLUGIN_RUNS:
            op.forces = f
            r = ref.rk_all(x, y, vx, vy, q, m, op, 1e-4, 6, t=0.25, order=order)
            for name, arr in zip(("x", "y", "vx", "vy"), r[:4]):
                out[f"rk{order}_{f}_{name}"] = arr
            out[f"rk{order}_{f}_t"] = r[4]
    np.savez_compressed(HERE / "plugins.npz", **out)


if __name__ == "__main__":
    main()
