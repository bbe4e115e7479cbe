"""The C ABI without a GPU: the library loads, exports every symbol the header
declares, validates arguments before touching the device, and its host-only
helpers (the synthetic lattice) behave as documented."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import demon_simulation_code_amd as dm
from demon_simulation_code_amd import _lib

HEADER = Path(__file__).resolve().parent.parent / "include" / "demon_gpu.h"


def declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_]+\s*\*?\s*(demon_[a-z0-9_]+)\s*\(", text, re.M)))


def test_every_declared_symbol_is_exported_and_bound():
    names = declared_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(dm.lib, name), name
        assert name in _lib.SIGNATURES, name
    assert set(_lib.SIGNATURES) == set(names)


def test_versions():
    assert dm.lib.demon_abi_version() == 3
    assert b"gfx950" in dm.lib.demon_version()


def test_struct_layouts_match_header():
    # demon_params: 4 doubles, 1 uint64, 4 int32, 2 + 19 doubles ; demon_stats: 3 int64, 2 doubles
    assert ctypes.sizeof(_lib.Params) == 4 * 8 + 8 + 4 * 4 + 21 * 8
    assert ctypes.sizeof(_lib.Stats) == 5 * 8
    assert _lib.Params.forces.offset == 40 and _lib.Params.device.offset == 52
    assert _lib.Params.magnetic_field.offset == 56 and _lib.Params.drag_scale.offset == 64
    assert _lib.Params.confine_x.offset == 72 and _lib.Params.heat_offset.offset == 216
    # the header declares the further plugins' constants in this order
    hdr = (Path(__file__).resolve().parents[1] / "include" / "demon_gpu.h").read_text()
    body = hdr[hdr.index("typedef struct demon_params"):hdr.index("} demon_params;")]
    pos = [body.index(name) for name in _lib.PLUGIN_PARAMS]
    assert pos == sorted(pos)


def test_create_rejects_bad_arguments_before_touching_the_device():
    p = _lib.Params(2000.0, 1.0, 2.0, 0.0, 0, 7, 0, 1, 0)
    h = ctypes.c_void_p()
    assert dm.lib.demon_create(ctypes.byref(h), 1, ctypes.byref(p)) == _lib.ERR_ARG
    assert b"n must be" in dm.lib.demon_last_error(None)
    bad = _lib.Params(2000.0, 1.0, 2.0, 0.0, 0, 7 | 32768, 0, 1, 0)  # no such ForceFlag
    assert dm.lib.demon_create(ctypes.byref(h), 16, ctypes.byref(bad)) == _lib.ERR_UNSUPPORTED
    nan = _lib.Params(2000.0, 1.0, 2.0, 0.0, 0, 7 | 2048, 0, 1, 0)
    nan.void_decay = float("nan")
    assert dm.lib.demon_create(ctypes.byref(h), 16, ctypes.byref(nan)) == _lib.ERR_ARG
    neg = _lib.Params(-1.0, 1.0, 2.0, 0.0, 0, 7, 0, 1, 0)
    assert dm.lib.demon_create(ctypes.byref(h), 16, ctypes.byref(neg)) == _lib.ERR_ARG
    prec = _lib.Params(2000.0, 1.0, 2.0, 0.0, 0, 7, 5, 1, 0)
    assert dm.lib.demon_create(ctypes.byref(h), 16, ctypes.byref(prec)) == _lib.ERR_ARG
    assert dm.lib.demon_create(None, 16, ctypes.byref(p)) == _lib.ERR_ARG


def test_null_context_calls_fail_cleanly():
    for rc in (dm.lib.demon_rk4_steps(None, 1, 1e-3, None),
               dm.lib.demon_upload(None, None, None, None, None, None, None),
               dm.lib.demon_synchronize(None),
               dm.lib.demon_use_graph(None, 1)):
        assert rc == _lib.ERR_ARG


def test_hex_lattice_properties():
    n = 4096
    x, y = dm.hex_lattice(n, 500e-6, 0.0, 3)
    assert abs(x.mean()) < 1e-15 and abs(y.mean()) < 1e-15
    # unjittered: nearest-neighbour distance is the spacing; 6 neighbours inside
    from scipy.spatial import cKDTree
    t = cKDTree(np.c_[x, y])
    d, _ = t.query(np.c_[x, y], k=2)
    assert np.allclose(d[:, 1], 500e-6, rtol=1e-12)
    # the n sites nearest the centre: the patch is a disc (radius ~ sqrt(n sqrt3/2pi) a)
    r = np.hypot(x, y)
    assert r.max() < (np.sqrt(n * np.sqrt(3) / (2 * np.pi)) + 1.5) * 500e-6
    xj, yj = dm.hex_lattice(n, 500e-6, 50e-6, 3)
    dev = np.maximum(np.abs(xj - x), np.abs(yj - y))
    assert dev.max() <= 50e-6 * (1 + 1e-9) and dev.max() > 45e-6
    x2, y2 = dm.hex_lattice(n, 500e-6, 50e-6, 3)
    assert np.array_equal(xj, x2) and np.array_equal(yj, y2)
    x3, _ = dm.hex_lattice(n, 500e-6, 50e-6, 4)
    assert not np.array_equal(xj, x3)


def test_hex_lattice_rejects_bad_arguments():
    with pytest.raises(dm.DemonError):
        dm.hex_lattice(0)
    with pytest.raises(dm.DemonError):
        dm.hex_lattice(10, spacing=-1.0)


def test_product_package_does_not_reference_the_oracle():
    pkg = Path(dm.__file__).resolve().parent
    for f in list(pkg.rglob("*.py")) + list(pkg.rglob("*.hip")) + list(pkg.rglob("*.h")) + list(pkg.rglob("*.cpp")):
        text = f.read_text()
        assert "oracle" not in text.replace("the oracle", ""), f
