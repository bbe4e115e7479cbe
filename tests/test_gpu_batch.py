"""Batched ensemble (demon_batch): every cloud against the oracle, graph and
plain launches, tuned and default tile shapes.  Bar: positions within 1e-9
relative after the run, identical times."""
import numpy as np
import pytest

import demon_simulation_code_amd as dm
from demon_simulation_code_amd.physics import ensemble
from oracle.bind import OracleParams
from tests.helpers import M0, pos_rel_err

pytestmark = pytest.mark.gpu


def _make(clouds, seed, multiple_of=4, n_max=1500, dt_check=True, forces=7):
    ens = ensemble(clouds, seed=seed, n_min=64, n_max=n_max, multiple_of=multiple_of)
    b = dm.Batch([e["n"] for e in ens], [e["kappa"] for e in ens], [e["confine"] for e in ens],
                 [e["gamma"] for e in ens], forces=forces, dt_check=dt_check)
    states = []
    for c, e in enumerate(ens):
        x, y = dm.hex_lattice(e["n"], 500e-6, 50e-6, 100 + c)
        q = np.full(e["n"], e["charge"])
        m = np.full(e["n"], M0)
        b.upload(c, x, y, None, None, q, m)
        states.append((x, y, q, m))
    return b, ens, states


def _check(oracle, b, ens, states, steps, dt_check=True):
    for c, (e, (x, y, q, m)) in enumerate(zip(ens, states)):
        z = np.zeros(e["n"])
        op = OracleParams(e["kappa"], e["confine"], e["gamma"], 0.0, 0, 7, 1 if dt_check else 0)
        o = oracle.rk4_steps(x, y, z, z, q, m, op, 1e-3, steps)
        g = b.download(c)
        st = b.stats(c)
        assert st.steps == steps and st.current_time == o[4], c
        assert pos_rel_err(o[0], o[1], g[0], g[1]) <= 1e-9, (c, e["n"])


def test_batch_matches_oracle_per_cloud(oracle):
    b, ens, states = _make(12, seed=1)
    with b:
        b.steps(20, 1e-3)
        _check(oracle, b, ens, states, 20)


def test_batch_plain_launches_equal_graph():
    b1, ens, states = _make(6, seed=2)
    b2, _, _ = _make(6, seed=2)
    with b1, b2:
        b1.use_graph(True)
        b2.use_graph(False)
        b1.steps(10, 1e-3)
        b2.steps(10, 1e-3)
        for c in range(len(ens)):
            for u, v in zip(b1.download(c), b2.download(c)):
                assert np.array_equal(u, v)


def test_batch_tuned_shapes_agree(oracle):
    b, ens, states = _make(10, seed=3, n_max=4000)
    with b:
        b.tune(reps=1)
        shapes = {bk: b.shape(bk) for bk in range(8, 14)}
        assert any(us > 0 for (_, _, us) in shapes.values())
        b.steps(10, 1e-3)
        _check(oracle, b, ens, states, 10)


def test_batch_any_n_without_dt_check(oracle):
    b, ens, states = _make(8, seed=5, multiple_of=1, dt_check=False)
    with b:
        assert any(e["n"] % 4 for e in ens)
        b.steps(10, 1e-3)
        _check(oracle, b, ens, states, 10, dt_check=False)


def test_batch_full_ensemble_smoke():
    ens = ensemble(256, seed=4)
    b = dm.Batch([e["n"] for e in ens], [e["kappa"] for e in ens], [e["confine"] for e in ens],
                 [e["gamma"] for e in ens])
    with b:
        for c, e in enumerate(ens):
            x, y = dm.hex_lattice(e["n"], 500e-6, 50e-6, 100 + c)
            b.upload(c, x, y, None, None, np.full(e["n"], e["charge"]), np.full(e["n"], M0))
        b.steps(3, 1e-3)
        for c in (0, 100, 255):
            assert all(np.isfinite(a).all() for a in b.download(c))
            assert b.stats(c).steps == 3
