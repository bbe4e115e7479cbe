"""ctypes binding of libdemon_gpu.so (the C ABI declared in include/demon_gpu.h).

The library is the product: there is no Python or CPU fallback.  If the
shared object is missing or fails to load, importing this module raises.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("DEMON_GPU_LIB", _PKG / "libdemon_gpu.so"))

# return codes
OK = 0
ERR_ARG = -1
ERR_HIP = -2
ERR_ALLOC = -3
ERR_STATE = -4
ERR_UNSUPPORTED = -5

# force flags (Force.h ForceFlag values)
CONFINEMENT = 1
DRAG = 2
SHIELDED_COULOMB = 4
GRAVITATIONAL = 8
THERMAL = 16
THERMAL_LOCALIZED = 32
DRIVING = 64
ROTATIONAL = 128
TIME_VARYING_DRAG = 256
TIME_VARYING_THERMAL = 512
MAGNETIC = 1024
CONFINEMENT_VOID = 2048
ELECTRIC = 4096
RECT_CONFINEMENT = 8192
VERT_ELECTRIC = 16384

# constants of the further force plugins (demon_params fields after drag_scale)
PLUGIN_PARAMS = ("confine_x", "confine_y", "void_decay", "rot_inner", "rot_outer", "rot_const",
                 "drive_const", "drive_amplitude", "drive_shift", "electric_field", "plasma_radius",
                 "vert_electric_field", "vert_decay", "gravity", "heat_in", "heat_out", "heat_radius",
                 "heat_scale", "heat_offset")

FP64 = 0
FP32_PAIR = 1
PAIRS_GATHER = 0
PAIRS_NEWTON3 = 1
BATCH_NEWTON3 = 0  # Batch.set_shape i_per_lane: the Newton-3 form


class DemonError(RuntimeError):
    """A C-ABI call returned a negative code."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [
        ("shielding", ctypes.c_double),
        ("confine", ctypes.c_double),
        ("drag_gamma", ctypes.c_double),
        ("heat_val", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("forces", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("dt_check", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("magnetic_field", ctypes.c_double),
        ("drag_scale", ctypes.c_double),
    ] + [(name, ctypes.c_double) for name in PLUGIN_PARAMS]


class Stats(ctypes.Structure):
    _fields_ = [
        ("steps", ctypes.c_int64),
        ("dt_reductions", ctypes.c_int64),
        ("force_evals", ctypes.c_int64),
        ("current_time", ctypes.c_double),
        ("last_dt", ctypes.c_double),
    ]


_D = ctypes.POINTER(ctypes.c_double)
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_CTX = ctypes.c_void_p

# name -> (restype, argtypes); every entry must be exported (checked by tests)
SIGNATURES = {
    "demon_version": (ctypes.c_char_p, []),
    "demon_abi_version": (ctypes.c_int, []),
    "demon_last_error": (ctypes.c_char_p, [_CTX]),
    "demon_create": (ctypes.c_int, [ctypes.POINTER(_CTX), _I64, ctypes.POINTER(Params)]),
    "demon_destroy": (None, [_CTX]),
    "demon_set_params": (ctypes.c_int, [_CTX, ctypes.POINTER(Params)]),
    "demon_get_params": (ctypes.c_int, [_CTX, ctypes.POINTER(Params)]),
    "demon_upload": (ctypes.c_int, [_CTX, _D, _D, _D, _D, _D, _D]),
    "demon_download": (ctypes.c_int, [_CTX, _D, _D, _D, _D]),
    "demon_force_stage": (ctypes.c_int, [_CTX, _I32, _I32, ctypes.c_uint64, _D, _D, _D, _D, _D, _D]),
    "demon_force_stage_at": (ctypes.c_int, [_CTX, _I32, _I32, ctypes.c_uint64, ctypes.c_double, _D, _D, _D, _D,
                                            _D, _D]),
    "demon_rk4_move": (ctypes.c_int, [_CTX, _D, ctypes.c_double, ctypes.c_double, ctypes.POINTER(_I64)]),
    "demon_rk4_steps": (ctypes.c_int, [_CTX, _I64, ctypes.c_double, _D]),
    "demon_rk2_move": (ctypes.c_int, [_CTX, _D, ctypes.c_double, ctypes.c_double, ctypes.POINTER(_I64)]),
    "demon_rk2_steps": (ctypes.c_int, [_CTX, _I64, ctypes.c_double, _D]),
    "demon_use_graph": (ctypes.c_int, [_CTX, _I32]),
    "demon_get_stats": (ctypes.c_int, [_CTX, ctypes.POINTER(Stats)]),
    "demon_synchronize": (ctypes.c_int, [_CTX]),
    "demon_set_pair_shape": (ctypes.c_int, [_CTX, _I32, _I32]),
    "demon_get_pair_shape": (ctypes.c_int, [_CTX, ctypes.POINTER(_I32), ctypes.POINTER(_I32)]),
    "demon_last_timing": (ctypes.c_int, [_CTX, _D, _D]),
    "demon_set_timing": (ctypes.c_int, [_CTX, _I32]),
    "demon_pair_counts": (ctypes.c_int, [_CTX, ctypes.POINTER(_I64), ctypes.POINTER(_I64)]),
    "demon_pair_tiles": (ctypes.c_int, [_CTX, ctypes.POINTER(_I64), ctypes.POINTER(_I64)]),
    "demon_last_kernel_ms": (ctypes.c_int, [_CTX, _D, ctypes.POINTER(_I64), _D, _D, ctypes.POINTER(_I64)]),
    "demon_set_pair_mode": (ctypes.c_int, [_CTX, _I32]),
    "demon_get_pair_mode": (ctypes.c_int, [_CTX, ctypes.POINTER(_I32)]),
    "demon_hex_lattice": (ctypes.c_int, [_I64, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, _D, _D]),
    "demon_fit_loss_grad": (ctypes.c_int, [_CTX, _I64, ctypes.c_double, _I64, _I64, _D, _D, _D, _D, _D,
                                           _D, _D, _D, _D, _D, _D]),
    "demon_fit_timing": (ctypes.c_int, [_CTX, _D, _D]),
    "demon_fit_adjoint_kernel": (ctypes.c_int, [_CTX, _D, ctypes.POINTER(_I64), ctypes.POINTER(_I64),
                                                ctypes.POINTER(_I64)]),
    "demon_batch_last_error": (ctypes.c_char_p, [_CTX]),
    "demon_batch_create": (ctypes.c_int, [ctypes.POINTER(_CTX), _I32, ctypes.POINTER(_I64), _D, _D, _D,
                                          ctypes.POINTER(Params)]),
    "demon_batch_destroy": (None, [_CTX]),
    "demon_batch_upload": (ctypes.c_int, [_CTX, _I32, _D, _D, _D, _D, _D, _D]),
    "demon_batch_download": (ctypes.c_int, [_CTX, _I32, _D, _D, _D, _D]),
    "demon_batch_steps": (ctypes.c_int, [_CTX, _I64, ctypes.c_double]),
    "demon_batch_cloud_stats": (ctypes.c_int, [_CTX, _I32, ctypes.POINTER(Stats)]),
    "demon_batch_use_graph": (ctypes.c_int, [_CTX, _I32]),
    "demon_batch_last_ms": (ctypes.c_int, [_CTX, _D]),
    "demon_batch_set_timing": (ctypes.c_int, [_CTX, _I32]),
    "demon_batch_kernel_ms": (ctypes.c_int, [_CTX, _D, ctypes.POINTER(_I64), _D, ctypes.POINTER(_I64)]),
    "demon_batch_pair_counts": (ctypes.c_int, [_CTX, ctypes.POINTER(_I64), ctypes.POINTER(_I64)]),
    "demon_batch_set_shape": (ctypes.c_int, [_CTX, _I32, _I32, _I32]),
    "demon_batch_get_shape": (ctypes.c_int, [_CTX, _I32, ctypes.POINTER(_I32), ctypes.POINTER(_I32), _D]),
    "demon_batch_tune": (ctypes.c_int, [_CTX, _I32]),
    "demon_kinetic_energy": (ctypes.c_int, [_CTX, _D, _D]),
    "demon_kinetic_energy_noalloc": (ctypes.c_int, [_CTX, _D, _D]),
    "demon_pair_histogram": (ctypes.c_int, [_CTX, ctypes.c_double, _I32, ctypes.POINTER(ctypes.c_uint64)]),
    "demon_trajectory_spectrum": (ctypes.c_int, [_D, _I64, _I64, _I64, _D, _I32]),
}


def _load() -> ctypes.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python demon_simulation_code_amd/build.py` "
            "(there is no CPU fallback)"
        )
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def _ptr(a):
    if a is None:
        return None
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
        raise TypeError("expected a C-contiguous float64 numpy array")
    return a.ctypes.data_as(_D)


def _check(rc: int, ctx=None):
    if rc != OK:
        msg = lib.demon_last_error(ctx)
        raise DemonError(rc, msg.decode() if msg else "")


def hex_lattice(n: int, spacing: float = 500e-6, jitter: float = 50e-6, seed: int = 0):
    """Synthetic input: n hexagonal-lattice sites nearest the origin, jittered."""
    x = np.empty(n)
    y = np.empty(n)
    _check(lib.demon_hex_lattice(n, spacing, jitter, seed, _ptr(x), _ptr(y)))
    return x, y


def trajectory_spectrum(series, nuse: int, device: int = 0) -> np.ndarray:
    """FFTAnalysis -a: series [rows, stride] of X positions (one row per
    snapshot); amplitude |DFT of the sum of the first nuse columns| / nuse for
    the rows // 2 lowest frequencies."""
    series = np.ascontiguousarray(series, dtype=np.float64)
    rows, stride = series.shape
    amp = np.zeros(rows // 2)
    _check(lib.demon_trajectory_spectrum(_ptr(series), rows, stride, int(nuse), _ptr(amp), device))
    return amp


class Context:
    """One device-resident cloud (demon_ctx)."""

    def __init__(self, n: int, *, shielding: float, confine: float = 0.0, gamma: float = 0.0,
                 heat: float = 0.0, seed: int = 0, forces: int = CONFINEMENT | DRAG | SHIELDED_COULOMB,
                 precision: int = FP64, dt_check: bool = True, device: int = 0,
                 magnetic_field: float = 0.0, drag_scale: float = 0.0, **plugin):
        """plugin: constants of the further force plugins by demon_params
        field name (confine_x, void_decay, rot_const, heat_in, ...)."""
        self.n = int(n)
        self.params = Params(shielding, confine, gamma, heat, seed, forces, precision,
                             1 if dt_check else 0, device, magnetic_field, drag_scale)
        for k, v in plugin.items():
            if k not in PLUGIN_PARAMS:
                raise TypeError(f"unknown force constant {k!r}")
            setattr(self.params, k, v)
        h = _CTX()
        _check(lib.demon_create(ctypes.byref(h), self.n, ctypes.byref(self.params)))
        self._h = h

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            lib.demon_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- calls ------------------------------------------------------------
    def _c(self, rc):
        _check(rc, self._h)

    def set_params(self, **kw):
        for k, v in kw.items():
            setattr(self.params, k, v)
        self._c(lib.demon_set_params(self._h, ctypes.byref(self.params)))

    def upload(self, x, y, vx=None, vy=None, charge=None, mass=None):
        self._c(lib.demon_upload(self._h, _ptr(x), _ptr(y), _ptr(vx), _ptr(vy), _ptr(charge), _ptr(mass)))

    def download(self):
        out = [np.empty(self.n) for _ in range(4)]
        self._c(lib.demon_download(self._h, *[_ptr(a) for a in out]))
        return tuple(out)

    def force_stage(self, forces, stage, step, xs, ys, vxs=None, vys=None, fx=None, fy=None, t=None):
        """Adds the stage force into fx/fy; t is the stage time (required for
        TIME_VARYING_DRAG, TIME_VARYING_THERMAL and DRIVING)."""
        fx = np.zeros(self.n) if fx is None else fx
        fy = np.zeros(self.n) if fy is None else fy
        if t is None:
            self._c(lib.demon_force_stage(self._h, forces, stage, step, _ptr(xs), _ptr(ys), _ptr(vxs),
                                          _ptr(vys), _ptr(fx), _ptr(fy)))
        else:
            self._c(lib.demon_force_stage_at(self._h, forces, stage, step, t, _ptr(xs), _ptr(ys), _ptr(vxs),
                                             _ptr(vys), _ptr(fx), _ptr(fy)))
        return fx, fy

    def kinetic_energy(self):
        """(total, thermal) kinetic energy of the current state [J]; the 2D
        kinetic temperature is thermal / (n k_B)."""
        tot, th = ctypes.c_double(), ctypes.c_double()
        self._c(lib.demon_kinetic_energy(self._h, ctypes.byref(tot), ctypes.byref(th)))
        return tot.value, th.value

    def pair_histogram(self, r_max: float, nbins: int) -> np.ndarray:
        """Unordered pair counts of the current positions in nbins bins of
        squared distance below r_max (uint64)."""
        out = np.zeros(int(nbins), dtype=np.uint64)
        self._c(lib.demon_pair_histogram(self._h, float(r_max), int(nbins),
                                         out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
        return out

    def rk4_steps(self, steps: int, dt: float, t: float = 0.0) -> float:
        tt = ctypes.c_double(t)
        self._c(lib.demon_rk4_steps(self._h, int(steps), dt, ctypes.byref(tt)))
        return tt.value

    def rk4_move(self, t: float, end_time: float, dt: float):
        tt = ctypes.c_double(t)
        st = _I64(0)
        self._c(lib.demon_rk4_move(self._h, ctypes.byref(tt), end_time, dt, ctypes.byref(st)))
        return tt.value, st.value

    def rk2_steps(self, steps: int, dt: float, t: float = 0.0) -> float:
        """Runge_Kutta2 (midpoint) steps; returns the time."""
        tt = ctypes.c_double(t)
        self._c(lib.demon_rk2_steps(self._h, int(steps), dt, ctypes.byref(tt)))
        return tt.value

    def rk2_move(self, t: float, end_time: float, dt: float):
        tt = ctypes.c_double(t)
        st = _I64(0)
        self._c(lib.demon_rk2_move(self._h, ctypes.byref(tt), end_time, dt, ctypes.byref(st)))
        return tt.value, st.value

    def use_graph(self, on: bool = True):
        self._c(lib.demon_use_graph(self._h, 1 if on else 0))

    def stats(self) -> Stats:
        s = Stats()
        self._c(lib.demon_get_stats(self._h, ctypes.byref(s)))
        return s

    def synchronize(self):
        self._c(lib.demon_synchronize(self._h))

    def set_pair_shape(self, i_per_lane: int, j_slices: int = 0):
        self._c(lib.demon_set_pair_shape(self._h, i_per_lane, j_slices))

    def pair_shape(self):
        r, s = _I32(0), _I32(0)
        self._c(lib.demon_get_pair_shape(self._h, ctypes.byref(r), ctypes.byref(s)))
        return r.value, s.value

    def fit_loss_grad(self, steps, dt, obs_every, ckpt_every, target_x, target_y, grad=True,
                      state_grad=False):
        """Loss of the trajectory fit and (optionally) its gradient.

        Returns (loss, grads) with grads = dict(q, kappa, confine, gamma[, x0, y0, vx0, vy0]).
        target_x/target_y: arrays [steps // obs_every, n] in caller order."""
        tx = np.ascontiguousarray(target_x, dtype=np.float64)
        ty = np.ascontiguousarray(target_y, dtype=np.float64)
        loss = ctypes.c_double(0)
        if not grad:
            self._c(lib.demon_fit_loss_grad(self._h, int(steps), dt, int(obs_every), int(ckpt_every),
                                            _ptr(tx), _ptr(ty), ctypes.byref(loss), *([None] * 8)))
            return loss.value, {}
        gq = np.zeros(self.n)
        g3 = [ctypes.c_double(0) for _ in range(3)]
        gs = [np.zeros(self.n) for _ in range(4)] if state_grad else [None] * 4
        self._c(lib.demon_fit_loss_grad(self._h, int(steps), dt, int(obs_every), int(ckpt_every),
                                        _ptr(tx), _ptr(ty), ctypes.byref(loss), _ptr(gq),
                                        ctypes.byref(g3[0]), ctypes.byref(g3[1]), ctypes.byref(g3[2]),
                                        *[_ptr(a) for a in gs]))
        out = dict(q=gq, kappa=g3[0].value, confine=g3[1].value, gamma=g3[2].value)
        if state_grad:
            out.update(x0=gs[0], y0=gs[1], vx0=gs[2], vy0=gs[3])
        return loss.value, out

    def fit_timing(self):
        a, b = ctypes.c_double(0), ctypes.c_double(0)
        self._c(lib.demon_fit_timing(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_timing(self, on: bool):
        """Per-launch HIP events around plain launches (on by default)."""
        self._c(lib.demon_set_timing(self._h, 1 if on else 0))

    def fit_adjoint_kernel(self):
        """The last fit's adjoint pair kernel: dict(pair_ms, launches,
        near_pairs, tested_pairs); the time comes from HIP events (0 with
        timing off)."""
        ms, nl, nn, nt = ctypes.c_double(0), _I64(0), _I64(0), _I64(0)
        self._c(lib.demon_fit_adjoint_kernel(self._h, ctypes.byref(ms), ctypes.byref(nl), ctypes.byref(nn),
                                             ctypes.byref(nt)))
        return dict(pair_ms=ms.value, launches=nl.value, near_pairs=nn.value, tested_pairs=nt.value)

    def last_timing(self):
        a, b = ctypes.c_double(0), ctypes.c_double(0)
        self._c(lib.demon_last_timing(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def kernel_timing(self):
        """Device ms of the last rk call by kernel, from HIP events around every
        plain launch: dict(pair_ms, pair_launches, list_ms, update_ms,
        update_launches)."""
        pm, lm, um = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_double(0)
        pl, ul = _I64(0), _I64(0)
        self._c(lib.demon_last_kernel_ms(self._h, ctypes.byref(pm), ctypes.byref(pl), ctypes.byref(lm),
                                         ctypes.byref(um), ctypes.byref(ul)))
        return dict(pair_ms=pm.value, pair_launches=pl.value, list_ms=lm.value, update_ms=um.value,
                    update_launches=ul.value)

    def pair_tiles(self):
        """(Newton-3 superblock tiles evaluated off the diagonal, work items)
        since upload."""
        a, b = _I64(0), _I64(0)
        self._c(lib.demon_pair_tiles(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_pair_mode(self, mode: int):
        """PAIRS_NEWTON3 (each unordered pair once) or PAIRS_GATHER (each ordered pair)."""
        self._c(lib.demon_set_pair_mode(self._h, int(mode)))

    def pair_mode(self) -> int:
        m = _I32(0)
        self._c(lib.demon_get_pair_mode(self._h, ctypes.byref(m)))
        return m.value

    def pair_counts(self):
        """(ordered pairs inside the cutoff, ordered pairs distance-tested) since upload."""
        a, b = _I64(0), _I64(0)
        self._c(lib.demon_pair_counts(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value


class Batch:
    """Many independent clouds stepped together (demon_batch)."""

    def __init__(self, sizes, kappa, confine, gamma, *, forces=CONFINEMENT | DRAG | SHIELDED_COULOMB,
                 precision=FP64, dt_check=True, heat=0.0, seed=0, device=0):
        self.sizes = [int(v) for v in sizes]
        k = len(self.sizes)
        n_arr = (_I64 * k)(*self.sizes)
        kap = np.ascontiguousarray(np.broadcast_to(kappa, (k,)), dtype=np.float64)
        con = np.ascontiguousarray(np.broadcast_to(confine, (k,)), dtype=np.float64)
        gam = np.ascontiguousarray(np.broadcast_to(gamma, (k,)), dtype=np.float64)
        p = Params(1.0, 0.0, 0.0, heat, seed, forces, precision, 1 if dt_check else 0, device)
        h = _CTX()
        rc = lib.demon_batch_create(ctypes.byref(h), k, n_arr, _ptr(kap), _ptr(con), _ptr(gam), ctypes.byref(p))
        if rc != OK:
            raise DemonError(rc, lib.demon_batch_last_error(None).decode())
        self._h = h

    def _c(self, rc):
        if rc != OK:
            raise DemonError(rc, lib.demon_batch_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib.demon_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, cloud, x, y, vx=None, vy=None, charge=None, mass=None):
        self._c(lib.demon_batch_upload(self._h, cloud, _ptr(x), _ptr(y), _ptr(vx), _ptr(vy), _ptr(charge), _ptr(mass)))

    def download(self, cloud):
        n = self.sizes[cloud]
        out = [np.empty(n) for _ in range(4)]
        self._c(lib.demon_batch_download(self._h, cloud, *[_ptr(a) for a in out]))
        return tuple(out)

    def steps(self, steps, dt):
        self._c(lib.demon_batch_steps(self._h, int(steps), dt))

    def stats(self, cloud) -> Stats:
        s = Stats()
        self._c(lib.demon_batch_cloud_stats(self._h, cloud, ctypes.byref(s)))
        return s

    def use_graph(self, on=True):
        self._c(lib.demon_batch_use_graph(self._h, 1 if on else 0))

    def last_ms(self):
        v = ctypes.c_double(0)
        self._c(lib.demon_batch_last_ms(self._h, ctypes.byref(v)))
        return v.value

    def set_timing(self, on: bool):
        """HIP events around every pair pass and stage update of plain
        (uncaptured) steps."""
        self._c(lib.demon_batch_set_timing(self._h, 1 if on else 0))

    def kernel_timing(self):
        """Device ms of the last steps call by kernel: dict(pair_ms,
        pair_launches, update_ms, update_launches)."""
        pm, um, pl, ul = ctypes.c_double(0), ctypes.c_double(0), _I64(0), _I64(0)
        self._c(lib.demon_batch_kernel_ms(self._h, ctypes.byref(pm), ctypes.byref(pl), ctypes.byref(um),
                                          ctypes.byref(ul)))
        return dict(pair_ms=pm.value, pair_launches=pl.value, update_ms=um.value, update_launches=ul.value)

    def pair_counts(self):
        """(ordered pairs inside the cutoff, ordered pairs distance-tested),
        summed over the clouds since their uploads."""
        a, b = _I64(0), _I64(0)
        self._c(lib.demon_batch_pair_counts(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_shape(self, bucket, i_per_lane, jslice):
        """i_per_lane 1, 2, 4: gather form, j slices of jslice particles;
        BATCH_NEWTON3: Newton-3 form, jslice superblock offsets per item."""
        self._c(lib.demon_batch_set_shape(self._h, bucket, i_per_lane, jslice))

    def shape(self, bucket):
        r, j, us = _I32(0), _I32(0), ctypes.c_double(0)
        self._c(lib.demon_batch_get_shape(self._h, bucket, ctypes.byref(r), ctypes.byref(j), ctypes.byref(us)))
        return r.value, j.value, us.value

    def tune(self, reps=3):
        self._c(lib.demon_batch_tune(self._h, reps))
