"""Writes tests/golden/*.npz from the reference's own code.

The reference (AU-PSL/demon_simulation_code) ships no tests, fixtures or
golden vectors.  This script compiles its hot-path sources in place (serial
build, oracle/Makefile target `ref`) and records, for fixed inputs stored in
the same files:
  forces_n{N}.npz   ShieldedCoulombForce::force1 and ::force3 on the inputs
  rk4_n{N}.npz      Runge_Kutta4 (Confinement + Drag + ShieldedCoulomb, with
                    Integrator::modifyTimeStep) after 1, 10 and 100 steps
  dtcheck.npz       Integrator::modifyTimeStep on constructed close pairs
Run only where /root/reference exists:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from oracle.bind import Reference, build  # noqa: E402

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
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
