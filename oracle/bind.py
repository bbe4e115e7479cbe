"""ctypes binding of the CPU oracle (liboracle.so) and, where it was built, of
the reference's own hot path compiled in place (_ref/libdemon_ref_*.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py -- never by the product package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle.so"
REF_SERIAL_SO = HERE / "_ref" / "libdemon_ref_serial.so"
REF_OMP_SO = HERE / "_ref" / "libdemon_ref_omp.so"

_D = ctypes.POINTER(ctypes.c_double)
_I64 = ctypes.c_int64


class OracleParams(ctypes.Structure):
    _fields_ = [
        ("shielding", ctypes.c_double),
        ("confine", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("heat_val", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("forces", ctypes.c_int32),
        ("dt_check", ctypes.c_int32),
        ("gather", ctypes.c_int32),
        ("magnetic_field", ctypes.c_double),
        ("drag_scale", ctypes.c_double),
    ] + [(name, ctypes.c_double) for name in (
        "confine_x", "confine_y", "void_decay", "rot_inner", "rot_outer", "rot_const",
        "drive_const", "drive_amplitude", "drive_shift", "electric_field", "plasma_radius",
        "vert_electric_field", "vert_decay", "gravity", "heat_in", "heat_out", "heat_radius",
        "heat_scale", "heat_offset")]


# Order of the parameter vector of ref_rk_all / ref_force_all (ref_harness.cpp RefPar).
REF_PAR = ("shielding", "confine", "gamma", "heat_val", "magnetic_field", "drag_scale", "confine_x",
           "confine_y", "void_decay", "rot_inner", "rot_outer", "rot_const", "drive_const",
           "drive_amplitude", "drive_shift", "electric_field", "plasma_radius", "vert_electric_field",
           "vert_decay", "gravity", "heat_in", "heat_out", "heat_radius", "heat_scale", "heat_offset")


def ref_par(op: "OracleParams") -> np.ndarray:
    """The OracleParams constants as the reference harness's parameter vector."""
    return np.array([getattr(op, k) for k in REF_PAR], dtype=np.float64)


def build(ref: bool = True) -> None:
    """make -C oracle (the _ref targets only where the reference sources exist)."""
    targets = ["oracle"]
    if ref and Path("/root/reference/Cloud.cpp").exists():
        targets.append("ref")
        targets.append("driver")
        if (HERE.parent / "demon_simulation_code_amd" / "libdemon_gpu.so").exists():
            targets.append("dropin")
    subprocess.run(["make", "-s", "-C", str(HERE), *targets], check=True)


def p(a):
    return None if a is None else a.ctypes.data_as(_D)


def _f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a


class Oracle:
    def __init__(self, path: Path = ORACLE_SO):
        if not path.exists():
            build(ref=False)
        L = ctypes.CDLL(str(path))
        self.L = L
        L.oracle_coulomb_const.restype = ctypes.c_double
        L.oracle_yukawa_add.argtypes = [_I64, _D, _D, _D, ctypes.c_double, _D, _D]
        L.oracle_yukawa_add.restype = ctypes.c_int
        for name in ("oracle_yukawa_gather_add", "oracle_yukawa_longdouble"):
            getattr(L, name).argtypes = [_I64, _D, _D, _D, ctypes.c_double, _D, _D]
            getattr(L, name).restype = None
        L.oracle_yukawa_rows.argtypes = [_I64, _D, _D, _D, ctypes.c_double,
                                         ctypes.POINTER(ctypes.c_int64), _I64, _D, _D]
        L.oracle_yukawa_rows.restype = None
        L.oracle_yukawa_cells_add.argtypes = [_I64, _D, _D, _D, ctypes.c_double, _D, _D]
        L.oracle_yukawa_cells_add.restype = ctypes.c_int
        L.oracle_dtcheck_min_s_cells.argtypes = [_I64, _D, _D, ctypes.c_double]
        L.oracle_dtcheck_min_s_cells.restype = ctypes.c_float
        L.oracle_modify_time_step_cells.argtypes = [_I64, _D, _D, ctypes.c_float, ctypes.c_double,
                                                    ctypes.c_double]
        L.oracle_modify_time_step_cells.restype = ctypes.c_double
        L.oracle_near_pairs.argtypes = [_I64, _D, _D, ctypes.c_double]
        L.oracle_near_pairs.restype = ctypes.c_int64
        L.oracle_confinement_add.argtypes = [_I64, _D, _D, _D, ctypes.c_double, _D, _D]
        L.oracle_drag_add.argtypes = [_I64, _D, _D, _D, ctypes.c_double, _D, _D]
        L.oracle_philox4x32.argtypes = [ctypes.POINTER(ctypes.c_uint32)] * 3
        L.oracle_thermal_add.argtypes = [_I64, ctypes.c_double, ctypes.c_uint64, ctypes.c_uint64,
                                         ctypes.c_int, _D, _D]
        L.oracle_modify_time_step.argtypes = [_I64, _D, _D, ctypes.c_float, ctypes.c_double]
        L.oracle_modify_time_step.restype = ctypes.c_double
        L.oracle_dtcheck_min_s.argtypes = [_I64, _D, _D]
        L.oracle_dtcheck_min_s.restype = ctypes.c_float
        L.oracle_dtcheck_covers.argtypes = [_I64, _I64]
        L.oracle_dtcheck_covers.restype = ctypes.c_int
        L.oracle_rk4_steps.argtypes = [_I64, _D, _D, _D, _D, _D, _D, ctypes.POINTER(OracleParams),
                                       ctypes.c_double, _I64, _D, ctypes.POINTER(ctypes.c_uint64)]
        L.oracle_rk2_steps.argtypes = L.oracle_rk4_steps.argtypes
        L.oracle_rk2_steps.restype = ctypes.c_int
        L.oracle_rk4_move.argtypes = [_I64, _D, _D, _D, _D, _D, _D, ctypes.POINTER(OracleParams),
                                      ctypes.c_double, ctypes.c_double, _D,
                                      ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(_I64)]
        L.oracle_stage_force.argtypes = [_I64, _D, _D, _D, _D, _D, _D, ctypes.POINTER(OracleParams),
                                         ctypes.c_uint64, ctypes.c_int, ctypes.c_double, _D, _D]
        L.oracle_fit_loss_grad.argtypes = [_I64, _D, _D, _D, _D, _D, _D, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_int, _I64, _I64, _D, _D, _D, _D, _D, _D, _D,
                                           _D, _D, _D, _D]
        L.oracle_fit_loss_grad.restype = ctypes.c_int
        L.oracle_average_fft.argtypes = [_D, _I64, _I64, _I64, _D]
        L.oracle_average_fft.restype = None
        self.coulomb = L.oracle_coulomb_const()

    # --- forces -----------------------------------------------------------
    def yukawa(self, x, y, q, kappa, fx=None, fy=None):
        n = len(x)
        fx = np.zeros(n) if fx is None else fx
        fy = np.zeros(n) if fy is None else fy
        x, y, q = _f64(x), _f64(y), _f64(q)
        if n % 2 == 0:
            rc = self.L.oracle_yukawa_add(n, p(x), p(y), p(q), kappa, p(fx), p(fy))
            assert rc == 0
        else:
            self.L.oracle_yukawa_gather_add(n, p(x), p(y), p(q), kappa, p(fx), p(fy))
        return fx, fy

    def yukawa_cells(self, x, y, q, kappa, fx=None, fy=None):
        """oracle_yukawa_add's result from a cell list (pairs in range only)."""
        n = len(x)
        fx = np.zeros(n) if fx is None else fx
        fy = np.zeros(n) if fy is None else fy
        rc = self.L.oracle_yukawa_cells_add(n, p(_f64(x)), p(_f64(y)), p(_f64(q)), kappa, p(fx), p(fy))
        assert rc == 0, "oracle_yukawa_cells_add rejected the input (odd n?)"
        return fx, fy

    def yukawa_gather(self, x, y, q, kappa):
        n = len(x)
        fx, fy = np.zeros(n), np.zeros(n)
        self.L.oracle_yukawa_gather_add(n, p(_f64(x)), p(_f64(y)), p(_f64(q)), kappa, p(fx), p(fy))
        return fx, fy

    def yukawa_longdouble(self, x, y, q, kappa):
        n = len(x)
        fx, fy = np.zeros(n), np.zeros(n)
        self.L.oracle_yukawa_longdouble(n, p(_f64(x)), p(_f64(y)), p(_f64(q)), kappa, p(fx), p(fy))
        return fx, fy

    def yukawa_rows(self, x, y, q, kappa, rows):
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        fx, fy = np.zeros(len(rows)), np.zeros(len(rows))
        self.L.oracle_yukawa_rows(len(x), p(_f64(x)), p(_f64(y)), p(_f64(q)), kappa,
                                  rows.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(rows),
                                  p(fx), p(fy))
        return fx, fy

    def near_pairs(self, x, y, kappa) -> int:
        return int(self.L.oracle_near_pairs(len(x), p(_f64(x)), p(_f64(y)), kappa))

    def stage_force(self, xs, ys, vxs, vys, q, m, params: OracleParams, step=0, stage=1, t=0.0):
        n = len(xs)
        fx, fy = np.zeros(n), np.zeros(n)
        rc = self.L.oracle_stage_force(n, p(_f64(xs)), p(_f64(ys)), p(_f64(vxs)), p(_f64(vys)),
                                       p(_f64(q)), p(_f64(m)), ctypes.byref(params), step, stage, t,
                                       p(fx), p(fy))
        assert rc == 0
        return fx, fy

    def thermal(self, n, heat, seed, step, stage):
        fx, fy = np.zeros(n), np.zeros(n)
        self.L.oracle_thermal_add(n, heat, seed, step, stage, p(fx), p(fy))
        return fx, fy

    def average_fft(self, series, nuse):
        """FFTAnalysis -a on series [rows, stride]: amplitudes for k < rows // 2."""
        series = np.ascontiguousarray(series, dtype=np.float64)
        rows, stride = series.shape
        amp = np.zeros(rows // 2)
        self.L.oracle_average_fft(p(series), rows, stride, nuse, p(amp))
        return amp

    def philox(self, ctr, key):
        c = (ctypes.c_uint32 * 4)(*ctr)
        k = (ctypes.c_uint32 * 2)(*key)
        o = (ctypes.c_uint32 * 4)()
        self.L.oracle_philox4x32(c, k, o)
        return list(o)

    def fit_loss_grad(self, x0, y0, vx0, vy0, q, m, kappa, confine, gamma, dt, steps, obs_every,
                      tx, ty, state_grad=False, dt_check=False):
        """(loss, dq, dkappa, dconfine, dgamma[, (dx0, dy0, dvx0, dvy0)])"""
        n = len(x0)
        loss = np.zeros(1)
        gq = np.zeros(n)
        g3 = np.zeros(3)
        gs = [np.zeros(n) for _ in range(4)] if state_grad else [None] * 4
        tx = np.ascontiguousarray(tx, dtype=np.float64)
        ty = np.ascontiguousarray(ty, dtype=np.float64)
        rc = self.L.oracle_fit_loss_grad(n, p(_f64(x0)), p(_f64(y0)), p(_f64(vx0)), p(_f64(vy0)),
                                         p(_f64(q)), p(_f64(m)), kappa, confine, gamma, dt,
                                         int(bool(dt_check)), steps,
                                         obs_every, p(tx), p(ty), p(loss), p(gq), p(g3[0:]),
                                         p(g3[1:]), p(g3[2:]), *[p(a) for a in gs])
        assert rc == 0
        out = (float(loss[0]), gq, float(g3[0]), float(g3[1]), float(g3[2]))
        return out + (tuple(gs),) if state_grad else out

    # --- dt check ---------------------------------------------------------
    def modify_time_step(self, x, y, dist=1.0e-4, dt=1.0e-3) -> float:
        return self.L.oracle_modify_time_step(len(x), p(_f64(x)), p(_f64(y)), dist, dt)

    def dtcheck_min_s(self, x, y) -> float:
        return self.L.oracle_dtcheck_min_s(len(x), p(_f64(x)), p(_f64(y)))

    def dtcheck_min_s_cells(self, x, y, cell=5e-3) -> float:
        return self.L.oracle_dtcheck_min_s_cells(len(x), p(_f64(x)), p(_f64(y)), cell)

    def modify_time_step_cells(self, x, y, dist=1.0e-4, dt=1.0e-3, cell=5e-3) -> float:
        return self.L.oracle_modify_time_step_cells(len(x), p(_f64(x)), p(_f64(y)), dist, dt, cell)

    def covers(self, i, j) -> bool:
        return bool(self.L.oracle_dtcheck_covers(i, j))

    # --- RK4 ----------------------------------------------------------------
    def rk4_steps(self, x, y, vx, vy, q, m, params: OracleParams, dt, steps, t=0.0, step0=0):
        """Returns (x, y, vx, vy, t, step) after `steps` steps (copies)."""
        s = [_f64(a).copy() for a in (x, y, vx, vy)]
        tt = np.array([t])
        sc = ctypes.c_uint64(step0)
        rc = self.L.oracle_rk4_steps(len(s[0]), *[p(a) for a in s], p(_f64(q)), p(_f64(m)),
                                     ctypes.byref(params), dt, steps, p(tt), ctypes.byref(sc))
        assert rc == 0, "oracle_rk4_steps rejected the input"
        return (*s, float(tt[0]), sc.value)

    def rk2_steps(self, x, y, vx, vy, q, m, params: OracleParams, dt, steps, t=0.0, step0=0):
        """Runge_Kutta2 steps; returns (x, y, vx, vy, t, step) (copies)."""
        s = [_f64(a).copy() for a in (x, y, vx, vy)]
        tt = np.array([t])
        sc = ctypes.c_uint64(step0)
        rc = self.L.oracle_rk2_steps(len(s[0]), *[p(a) for a in s], p(_f64(q)), p(_f64(m)),
                                     ctypes.byref(params), dt, steps, p(tt), ctypes.byref(sc))
        assert rc == 0, "oracle_rk2_steps rejected the input"
        return (*s, float(tt[0]), sc.value)

    def rk4_move(self, x, y, vx, vy, q, m, params: OracleParams, dt, end_time, t=0.0):
        s = [_f64(a).copy() for a in (x, y, vx, vy)]
        tt = np.array([t])
        sc = ctypes.c_uint64(0)
        nst = _I64(0)
        rc = self.L.oracle_rk4_move(len(s[0]), *[p(a) for a in s], p(_f64(q)), p(_f64(m)),
                                    ctypes.byref(params), dt, end_time, p(tt), ctypes.byref(sc),
                                    ctypes.byref(nst))
        assert rc == 0
        return (*s, float(tt[0]), nst.value)


def default_threads() -> int:
    """OpenMP threads for the reference: the CPUs this process may run on,
    at most 16 (the reference itself would start omp_get_num_procs() threads,
    Cloud.cpp:34-36, which oversubscribes a CPU-limited container)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


class Reference:
    """The reference's own classes (compiled in place) behind ref_harness.cpp."""

    def __init__(self, openmp: bool = False):
        path = REF_OMP_SO if openmp else REF_SERIAL_SO
        if not path.exists():
            raise FileNotFoundError(f"{path} not built (needs the reference sources)")
        L = ctypes.CDLL(str(path))
        self.L = L
        self.openmp = bool(L.ref_is_openmp())
        self.threads = default_threads() if self.openmp else 1
        L.ref_yukawa_stage.argtypes = [ctypes.c_int, _I64, _D, _D, _D, ctypes.c_double, _D, _D, ctypes.c_int]
        L.ref_modify_time_step.argtypes = [_I64, _D, _D, ctypes.c_float, ctypes.c_double]
        L.ref_modify_time_step.restype = ctypes.c_double
        L.ref_rk4_record.argtypes = [_I64, _D, _D, _D, _D, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                     _I64, _I64, _D, ctypes.c_int]
        L.ref_rk4.argtypes = [_I64, _D, _D, _D, _D, _D, _D, ctypes.c_double, ctypes.c_double,
                              ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                              ctypes.c_int, _I64, ctypes.c_double, _D, ctypes.c_int]
        L.ref_rk4_ext.argtypes = [_I64, _D, _D, _D, _D, _D, _D, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_int, ctypes.c_double, _I64, _D, ctypes.c_int]
        L.ref_rk2_ext.argtypes = L.ref_rk4_ext.argtypes
        L.ref_fits_write.argtypes = [ctypes.c_char_p, _I64, _D, _D, _D, _D, _D, _D, ctypes.c_int, _D, _I64]
        L.ref_fits_forces_roundtrip.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
        L.ref_num_par.restype = ctypes.c_int
        assert L.ref_num_par() == len(REF_PAR)
        L.ref_rk_all.argtypes = [ctypes.c_int, _I64, _D, _D, _D, _D, _D, _D, _D, ctypes.c_int,
                                 ctypes.c_double, _I64, _D, ctypes.c_int]
        L.ref_force_all.argtypes = [_I64, _D, _D, _D, _D, _D, _D, _D, ctypes.c_int, ctypes.c_double,
                                    _I64, _D, _D]
        L.ref_fits_read.argtypes = [ctypes.c_char_p, _I64, _D, _D, _D, _D, _D, _D,
                                    ctypes.POINTER(ctypes.c_long), _D, _D]

    @staticmethod
    def available(openmp: bool = False) -> bool:
        return (REF_OMP_SO if openmp else REF_SERIAL_SO).exists()

    def yukawa(self, x, y, q, kappa, stage=1, threads=0):
        n = len(x)
        fx, fy = np.zeros(n), np.zeros(n)
        rc = self.L.ref_yukawa_stage(stage, n, p(_f64(x)), p(_f64(y)), p(_f64(q)), kappa, p(fx), p(fy),
                                     threads or self.threads)
        assert rc == 0
        return fx, fy

    def modify_time_step(self, x, y, dist=1.0e-4, dt=1.0e-3):
        return self.L.ref_modify_time_step(len(x), p(_f64(x)), p(_f64(y)), dist, dt)

    def rk4_steps(self, x, y, vx, vy, q, m, kappa, confine, gamma, heat, forces, dt, steps,
                  t=0.0, threads=0):
        s = [_f64(a).copy() for a in (x, y, vx, vy)]
        tt = np.array([t])
        rc = self.L.ref_rk4(len(s[0]), *[p(a) for a in s], p(_f64(q)), p(_f64(m)), kappa, confine,
                            gamma, heat, forces, dt, 0, steps, 0.0, p(tt), threads or self.threads)
        assert rc == 0
        return (*s, float(tt[0]))

    def rk4_steps_ext(self, x, y, vx, vy, q, m, kappa, confine, gamma, heat, bfield, drag_scale, forces,
                      dt, steps, t=0.0, threads=0, order=4):
        """rk4_steps with the Magnetic and TimeVaryingDrag forces available;
        order=2 runs Runge_Kutta2 instead."""
        s = [_f64(a).copy() for a in (x, y, vx, vy)]
        tt = np.array([t])
        fn = self.L.ref_rk4_ext if order == 4 else self.L.ref_rk2_ext
        rc = fn(len(s[0]), *[p(a) for a in s], p(_f64(q)), p(_f64(m)), kappa, confine,
                gamma, heat, bfield, drag_scale, forces, dt, steps, p(tt), threads or self.threads)
        assert rc == 0
        return (*s, float(tt[0]))

    def rk_all(self, x, y, vx, vy, q, m, op, dt, steps, t=0.0, order=4, threads=0):
        """`steps` steps of Runge_Kutta4 / Runge_Kutta2 with every force of op.forces,
        constructed as the driver does from the OracleParams constants."""
        s = [_f64(a).copy() for a in (x, y, vx, vy)]
        tt = np.array([t])
        rc = self.L.ref_rk_all(order, len(s[0]), *[p(a) for a in s], p(_f64(q)), p(_f64(m)),
                               p(ref_par(op)), op.forces, dt, steps, p(tt), threads or self.threads)
        assert rc == 0
        return (*s, float(tt[0]))

    def force_all(self, x, y, vx, vy, q, m, op, t=0.0, reps=1):
        """The forces of op.forces on one state at time t, `reps` evaluations
        (force1..force4 in turn): fx, fy of shape [reps, n]."""
        n = len(x)
        fx, fy = np.zeros((reps, n)), np.zeros((reps, n))
        rc = self.L.ref_force_all(n, *[p(_f64(a)) for a in (x, y, vx, vy, q, m)], p(ref_par(op)),
                                  op.forces, t, reps, p(fx), p(fy))
        assert rc == 0
        return fx, fy

    def fits_write(self, path, x, y, vx, vy, q, m, forces, par, rows):
        """The reference's own writer (driver + Cloud::writeCloudSetup/writeTimeStep);
        par = (confine, gamma, B, kappa, heat, TV scale); row k has x + 1e-6 k at t = 0.01 k."""
        rc = self.L.ref_fits_write(str(path).encode(), len(x), *[p(_f64(a)) for a in (x, y, vx, vy, q, m)],
                                   forces, p(_f64(np.asarray(par, dtype=np.float64))), rows)
        assert rc == 0, rc

    def fits_forces_roundtrip(self, src, dst):
        """The reference's readForce of every force in src's FORCES, then its
        writeForce of the same objects into a new file dst (primary header only)."""
        rc = self.L.ref_fits_forces_roundtrip(str(src).encode(), str(dst).encode())
        assert rc == 0, f"cfitsio status {rc}"

    def fits_read(self, path, n):
        """The reference's reader: (x, y, vx, vy, q, m, forces, par, t) of the last row."""
        out = [np.zeros(n) for _ in range(6)]
        par = np.zeros(6)
        t = np.zeros(1)
        fl = ctypes.c_long(0)
        rc = self.L.ref_fits_read(str(path).encode(), n, *[p(a) for a in out], ctypes.byref(fl), p(par), p(t))
        assert rc == 0, rc
        return (*out, fl.value, par, float(t[0]))

    def rk4_record(self, x, y, q, m, kappa, confine, gamma, heat, forces, dt, steps, every, threads=0):
        """Snapshots [steps//every, 4, n] (x, y, vx, vy) of one integrator run."""
        n = len(x)
        rec = np.zeros((steps // every, 4, n))
        rc = self.L.ref_rk4_record(n, p(_f64(x)), p(_f64(y)), p(_f64(q)), p(_f64(m)), kappa, confine,
                                   gamma, heat, forces, dt, steps, every, p(rec), threads or self.threads)
        assert rc == 0
        return rec

    def rk4_move(self, x, y, vx, vy, q, m, kappa, confine, gamma, heat, forces, dt, end_time,
                 t=0.0, threads=0):
        s = [_f64(a).copy() for a in (x, y, vx, vy)]
        tt = np.array([t])
        rc = self.L.ref_rk4(len(s[0]), *[p(a) for a in s], p(_f64(q)), p(_f64(m)), kappa, confine,
                            gamma, heat, forces, dt, 1, 0, end_time, p(tt), threads or self.threads)
        assert rc == 0
        return (*s, float(tt[0]))
