"""The reference's own classes with the GPU binding swapped in (oracle/dropin_test.cpp,
built from the reference sources by `make -C oracle dropin`), against the reference
alone on the same inputs: 20 RK4 steps of 1024 particles with the dt check, with
the driver's default forces, with forceFlags B,D (Magnetic, TimeVaryingDrag),
with Runge_Kutta2 (-I), and with the deterministic further plugins (ConfinementVoid,
Driving, RectConfinement, Rotational, Electric, Gravitational, VertElectric).

Bar: positions within 1e-9 relative, velocities within 1e-8 relative, identical time.
"""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

EXE = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "dropin_test"


@pytest.mark.skipif(not EXE.exists(), reason="oracle/_ref/dropin_test not built (needs the reference sources)")
def test_reference_with_gpu_binding():
    r = subprocess.run([str(EXE), "1024", "20"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = {line.split()[0]: line.split()[1:] for line in r.stdout.splitlines()}
    assert set(rows) == {"alacarte", "gpuforces", "gpurk4", "bd_gpuforces", "bd_gpurk4", "rk2_gpurk2",
                         "plug_gpuforces", "plug_gpurk4"}
    for name, (dr, dv, t, tref) in rows.items():
        assert float(dr) <= 1e-9, (name, dr)
        assert float(dv) <= 1e-8, (name, dv)
        assert float(t) == float(tref), name
