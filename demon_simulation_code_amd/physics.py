"""Physical constants and the mapping of the benchmark workloads onto the
reference's parameter forms.

Constants follow the reference, not CODATA:
  electron charge -1.602e-19 C   (Cloud.cpp:15)
  eps0 8.8541878e-12 F/m         (Cloud.cpp:16)
  1/(4 pi eps0)                  (ShieldedCoulombForce.cpp:21)
Confinement in the reference is F = q C r with C in V/m^2
(ConfinementForce.cpp:48-54); a spring constant k [N/m] maps to C = k/|q|.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

ELECTRON_CHARGE = -1.602e-19
EPSILON0 = 8.8541878e-12
COULOMB = 1.0 / (4.0 * math.pi * EPSILON0)


def confine_from_k(k: float, charge: float) -> float:
    """C [V/m^2] such that q C r = -k r for a particle of charge q < 0."""
    return k / abs(charge)


@dataclass(frozen=True)
class Workload:
    """A single-cloud benchmark workload (hexagonal lattice, thermal off)."""

    name: str
    n: int
    steps: int
    dt: float = 1.0e-3
    spacing: float = 500e-6
    jitter: float = 50e-6
    lam: float = 500e-6          # Debye length; kappa = 1/lam
    charge_e: float = 6000.0     # charge in units of |e|
    mass: float = 2.8e-14
    k_spring: float = 5e-11      # N/m
    gamma: float = 2.0
    seed: int = 1

    @property
    def kappa(self) -> float:
        return 1.0 / self.lam

    @property
    def charge(self) -> float:
        return self.charge_e * ELECTRON_CHARGE

    @property
    def confine(self) -> float:
        return confine_from_k(self.k_spring, self.charge)


# The reference's CPU-runnable case, the large single cloud (the headline
# metric's workload), and the differentiable fit's size.
PLAIN_1K = Workload("hex1024_fp64", n=1024, steps=1000)
THERMAL_16K = Workload("hex16384_fp64", n=16384, steps=5000)
LARGE_131K = Workload("hex131072_fp64", n=131072, steps=200)
ADJOINT_4K = Workload("hex4096_fp64", n=4096, steps=200)


def ensemble(clouds: int = 256, seed: int = 4, n_min: int = 64, n_max: int = 8192,
             multiple_of: int = 1):
    """The batched workload: per cloud N log-uniform in [n_min, n_max],
    lambda uniform in [250, 1000] um, charge uniform in [3000, 9000] e,
    k uniform in [2e-11, 1e-10] N/m, gamma = 2 /s.  Returns a list of dicts
    (n, kappa, charge, confine, gamma)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    out = []
    for _ in range(clouds):
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        n = max(n_min, (n + multiple_of - 1) // multiple_of * multiple_of)
        lam = rng.uniform(250e-6, 1000e-6)
        q = rng.uniform(3000.0, 9000.0) * ELECTRON_CHARGE
        k = rng.uniform(2e-11, 1e-10)
        out.append(dict(n=n, kappa=1.0 / lam, charge=q, confine=confine_from_k(k, q), gamma=2.0))
    return out
