"""demon_kinetic_energy_noalloc (the entry point that allocates no device
memory) against demon_kinetic_energy, through the C ABI: the same kernels,
so the same doubles bit for bit, on two calls in a row with the second on a
smaller cloud (the fixed scratch keeps the first call's contents), and a
clear error beyond its fixed size."""
import ctypes

import numpy as np
import pytest

import demon_simulation_code_amd as dm
from demon_simulation_code_amd._lib import lib
from tests.helpers import CONFINE, GAMMA, KAPPA, M0, Q0, lattice

pytestmark = pytest.mark.gpu


def _energies(c, fn):
    tot, th = ctypes.c_double(), ctypes.c_double()
    rc = fn(c._h, ctypes.byref(tot), ctypes.byref(th))
    return rc, tot.value, th.value


def _cloud(n, seed):
    x, y = lattice(n, seed=seed)
    rng = np.random.default_rng(seed)
    vx, vy = rng.normal(0.0, 1e-3, (2, n)) + np.array([[2e-4], [-1e-4]])
    m = M0 * rng.uniform(0.5, 1.5, n)
    c = dm.Context(n, shielding=KAPPA, confine=CONFINE, gamma=GAMMA)
    c.upload(x, y, vx, vy, np.full(n, Q0), m)
    return c


@pytest.mark.parametrize("first", ["noalloc", "existing"])
def test_noalloc_kinetic_energy_equals_existing(first):
    for n, seed in ((131072, 1), (1000, 2)):  # the second call on a smaller cloud
        with _cloud(n, seed) as c:
            order = [("noalloc", lib.demon_kinetic_energy_noalloc), ("existing", lib.demon_kinetic_energy)]
            if first == "existing":
                order.reverse()
            got = {name: _energies(c, fn) for name, fn in order}
            a, b = got["noalloc"], got["existing"]
            assert a[0] == 0 and b[0] == 0
            assert a[1] > 0.0 and a[2] > 0.0
            assert a[1] == b[1] and a[2] == b[2], (n, a, b)


def test_noalloc_kinetic_energy_rejects_clouds_beyond_its_scratch():
    n = (1 << 22) + 1
    c = dm.Context(n, shielding=KAPPA, confine=CONFINE, gamma=GAMMA)
    with c:
        x = np.linspace(-1.0, 1.0, n)
        c.upload(x, np.zeros(n), None, None, np.full(n, Q0), np.full(n, M0))
        rc, _, _ = _energies(c, lib.demon_kinetic_energy_noalloc)
        assert rc == -5  # DEMON_ERR_UNSUPPORTED
        assert b"4194304" in lib.demon_last_error(c._h)
        # the existing entry point still serves it
        rc, tot, _ = _energies(c, lib.demon_kinetic_energy)
        assert rc == 0 and tot == 0.0
