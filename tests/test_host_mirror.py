"""The host C++ mirror (libdemon_host.so): Params.cfg parsing with the
reference's rules, and the Cloud/Force/Runge_Kutta4 classes on the device."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle.bind import OracleParams
from tests.helpers import pair_abs_sum, pos_rel_err

HERE = Path(__file__).resolve().parent
EXE = HERE / "host" / "test_host"
GOLD = HERE / "golden"


def run(*args, check=True):
    if not EXE.exists():
        pytest.fail(f"{EXE} not built: run __graft_entry__.build()")
    r = subprocess.run([str(EXE), *map(str, args)], capture_output=True, text=True, timeout=600)
    if check and r.returncode != 0:
        raise AssertionError(r.stdout + r.stderr)
    return r


def parse(text=None, path=None):
    r = run("parse_text", text) if text is not None else run("parse", path)
    out = {}
    for line in r.stdout.splitlines():
        k, v = line.split("=")
        out[k] = float(v)
    return out


# --- Params.cfg (driver.cpp:626-824) ------------------------------------------

def test_reference_params_cfg():
    """The reference's own Params.cfg: every key read, forceFlags 'B,D,' -> Magnetic|TimeVaryingDrag."""
    p = parse(path=GOLD / "Params.cfg")
    assert p["numParticles"] == 100 and p["simTimeStep"] == 1e-4 and p["endTime"] == 5.0
    assert p["shieldingConstant"] == 20000 and p["dragGamma"] == 10.0 and p["rMean"] == 1.45e-6
    assert p["confinementConstY"] == 1000 and p["massDensity"] == 2200
    assert int(p["usedForces"]) == 1024 | 256


def test_reference_vortex_cfg():
    """example/Params_vortex.cfg ends with an empty forceFlags value: no flags."""
    p = parse(path=GOLD / "Params_vortex.cfg")
    assert p["endTime"] == 10.0 and p["dragGamma"] == 0.0 and p["magneticFieldStrength"] == 0.0
    assert p["voidDecay"] == 0.0 and p["confinementConstY"] == 100
    assert int(p["usedForces"]) == 0


@pytest.mark.parametrize("flags,want", [
    ("T,", 16), ("T", 0), ("B,D", 1024), ("B,D,", 1280), ("G,F,E,", 8 | 16384 | 4096),
    ("R,S,", 8192 | 128), ("w,V,", 64 | 2048),
])
def test_force_flags_need_trailing_comma(flags, want):
    assert int(parse(f"forceFlags {flags}")["usedForces"]) == want


def test_last_duplicate_wins_unknown_ignored_atof():
    p = parse("dragGamma 3 bogusKey 7 dragGamma 4.5 numParticles 12.9 spacing 1e-3x")
    assert p["dragGamma"] == 4.5 and p["numParticles"] == 12 and p["spacing"] == 1e-3


@pytest.mark.parametrize("flags", ["T,L,", "L,v,", "B,B,"])
def test_conflicting_flags_are_errors(flags):
    r = run("parse_text", f"forceFlags {flags}", check=False)
    assert r.returncode == 1 and ("conflicts" in r.stdout or "already set" in r.stdout)


def test_missing_file_is_an_error():
    r = run("parse", "/nonexistent/Params.cfg", check=False)
    assert r.returncode == 1 and "Incorrect parameter filepath" in r.stdout


# --- device classes -------------------------------------------------------------

@pytest.mark.gpu
def test_shielded_coulomb_force_class(oracle):
    n = 512
    a = np.array([[float(v) for v in line.split()] for line in run("forces", n).stdout.splitlines()])
    x, y, q, f1x, f1y, f3x, f3y = a.T
    ox, oy = oracle.yukawa(x, y, q, 2000.0)
    scale = pair_abs_sum(x, y, q, 2000.0)
    assert np.max(np.hypot(ox - f1x, oy - f1y) / scale) <= 1e-12
    xs, ys = x + 1e-5, y - 2e-5
    ox3, oy3 = oracle.yukawa(xs, ys, q, 2000.0)
    assert np.max(np.hypot(ox3 - f3x, oy3 - f3y) / pair_abs_sum(xs, ys, q, 2000.0)) <= 1e-12


@pytest.mark.gpu
def test_runge_kutta4_class(oracle):
    import demon_simulation_code_amd as dm

    n, steps = 256, 40
    lines = run("rk4", n, steps).stdout.splitlines()
    t, total = lines[0].split()
    g = np.array([[float(v) for v in ln.split()] for ln in lines[1:]]).T
    x, y = dm.hex_lattice(n, 500e-6, 50e-6, 1)
    q = np.full(n, 6000 * -1.602e-19)
    m = np.full(n, 2.8e-14)
    z = np.zeros(n)
    op = OracleParams(2000.0, 5e-11 / 9.612e-16, 2.0, 0.0, 0, 7, 1)
    o = oracle.rk4_steps(x, y, z, z, q, m, op, 1e-3, steps)
    assert int(total) == steps and float(t) == o[4]
    assert pos_rel_err(o[0], o[1], g[0], g[1]) <= 1e-9
