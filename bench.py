#!/usr/bin/env python3
"""Benchmark of the Yukawa all-pairs + RK4 hot path on one MI355X.

Workload (the headline metric's configuration): one cloud of N = 131072
particles on a hexagonal lattice (500 um spacing, +-50 um jitter), charge
6000 e, mass 2.8e-14 kg, lambda = 500 um (kappa = 2000 /m), confinement
k = 5e-11 N/m, drag gamma = 2 /s, thermal off, RK4 with dt = 1e-3 s and the
reference's dt check, FP64.  A "step" is one RK4 step (4 all-pairs force
evaluations + the dt check) from the lattice; the timed `--steps K` are the
first K steps of the 200-step trajectory (after a warm-up and a re-upload),
each step replayed as the context's captured step graph (`--no-graphs`:
plain launches).

Prints ONE JSON line:
  value        pair interactions/s = 4 N (N-1) K / t   (ordered pairs)
  rk4_steps_per_s, ms_per_step
  roofline     the all-pairs kernel over the timed launches: the FP64 FLOP of
               the pairs it actually evaluated -- 7 per ordered pair given a
               distance test, 20 per one inside the cutoff (counted exactly in
               the kernel) -- over its mean launch time (HIP events around
               every launch on the kernel's stream, in a pass of the same
               steps from the same state with plain launches, checked to end
               in the timed run's state), against the FP64 vector peak of
               78.6 TFLOP/s.  `effective` prices every ordered pair of
               N (N-1), the box-culled ones included, at 7 FLOP (the rate the
               culling buys, not a hardware throughput).
  update       the fused stage-update kernel: bytes per launch (its loads and
               stores per particle, plus the all-pairs partial sums and
               Newton-3 reactions it reads) over its mean launch time
  c3_200       the whole 200-step trajectory (the cloud implodes: after about
               100 steps every pair is inside the cutoff), FP64 and FP32-pair
  fp32_pair    the timed run with FP32 pair arithmetic (speed and error)
  cpu_baseline the reference's own SSE/OpenMP code (compiled from its sources
               into oracle/_ref) on the host cores: whole Runge_Kutta4 steps
               (moveParticles with modifyTimeStep) from the same lattice
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK = 78.6e12       # MI355X FP64 vector, FLOP/s
FP32_PEAK = 157.3e12      # MI355X FP32 vector, FLOP/s
HBM_PEAK = 8.0e12         # MI355X HBM3E, B/s
FLOP_FAR, FLOP_NEAR = 7, 20
# The adjoint pair sum per ordered pair inside the cutoff, in the same
# arithmetic (+ - x / sqrt exp each 1, FMA 2): the 7 of the distance test
# (which forms r and u = kappa r), y = 1/r (1), y^2, y^3 (2), C q_j (1),
# exp (1), phi/q_i = C q_j e^-u (1 + u) y^3 (4), C q_i q_j e^-u (2), phi (1),
# phi'/r (6), u_i - u_j (2), its dot with d (3), w (1), the two
# force-adjoint sums (8), the charge sum (2), the kappa sum (7).
FLOP_ADJ_NEAR = 48
# HBM traffic of the FP64 all-pairs kernel per launch: FETCH_SIZE and
# WRITE_SIZE from two rocprofv3 --pmc passes over a run of the same
# command, scaled by what they read for a plain copy of known size (the
# summary file gives the commands, the counts and the scale).
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_pairs_n3_hex131072_fp64_traffic.md")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(workload, steps, warmup, start, mode):
    """Bytes per launch from the committed counter summary, when a run on the
    same workload, step count, warm-up, start step and pair kernel was
    counted there; None otherwise."""
    try:
        text = open(TRAFFIC_FILE).read()
    except OSError:
        return None
    want = {"workload": workload, "steps": str(steps), "warmup": str(warmup), "start_step": str(start),
            "pair_kernel": mode}
    for line in text.splitlines():
        if not line.startswith("traffic:"):
            continue
        fields = dict(kv.partition("=")[::2] for kv in line[len("traffic:"):].split())
        if all(fields.get(k) == v for k, v in want.items()) and "bytes_per_launch" in fields:
            return float(fields["bytes_per_launch"])
    return None


def setup(wl):
    import demon_simulation_code_amd as dm

    x, y = dm.hex_lattice(wl.n, wl.spacing, wl.jitter, wl.seed)
    q = np.full(wl.n, wl.charge)
    m = np.full(wl.n, wl.mass)
    return x, y, q, m


# Every per-particle force plugin with zero strength: the trajectory is the
# headline one bit for bit, so a profile of the stage-update kernels with and
# without them measures what the plugins cost.
ZERO_PLUGINS = dict(confine_x=0.0, confine_y=0.0, void_decay=0.0, rot_inner=1e-3, rot_outer=1.0, rot_const=0.0,
                    drive_const=1e-5, drive_amplitude=0.0, drive_shift=7e-3, electric_field=0.0, plasma_radius=1.0,
                    vert_electric_field=0.0, vert_decay=1.0, gravity=0.0, heat_in=0.0, heat_out=0.0,
                    heat_radius=1e-3, heat_scale=0.0, heat_offset=0.0)


def run_gpu(wl, steps, warmup, precision, shape=None, snapshots=(), start=0, mode=None, plugins_zero=False,
            graph=True):
    """Warm up on the workload, re-upload the initial state, then time exactly
    `steps` RK4 steps from it (one demon_rk4_steps call, each step replayed
    as the captured step graph when `graph`); with `start`, from the state
    after that many untimed steps instead.  An instrumented pass then runs
    the same steps from the same state with plain launches and HIP events
    around every launch (the kernel times and pair counts of the roofline:
    the same work, the steps being deterministic).  With `snapshots` (step
    numbers), an untimed pass records the state at those steps."""
    import demon_simulation_code_amd as dm

    x, y, q, m = setup(wl)
    n = wl.n
    forces = dm.CONFINEMENT | dm.DRAG | dm.SHIELDED_COULOMB
    extra = {}
    if plugins_zero:
        # ConfinementVoid with decay 0 is the confinement force plus exactly 0
        forces = (dm.CONFINEMENT_VOID | dm.DRAG | dm.DRIVING | dm.MAGNETIC | dm.RECT_CONFINEMENT | dm.ROTATIONAL
                  | dm.SHIELDED_COULOMB | dm.THERMAL_LOCALIZED | dm.TIME_VARYING_THERMAL | dm.ELECTRIC
                  | dm.GRAVITATIONAL | dm.VERT_ELECTRIC)
        extra = ZERO_PLUGINS
    c = dm.Context(n, shielding=wl.kappa, confine=wl.confine, gamma=wl.gamma, forces=forces,
                   precision=precision, **extra)
    if mode is not None:
        c.set_pair_mode(mode)
    if shape:
        c.set_pair_shape(*shape)
    c.upload(x, y, None, None, q, m)
    c.use_graph(graph)
    if warmup:
        c.rk4_steps(warmup, wl.dt, 0.0)
    c.upload(x, y, None, None, q, m)
    t_start = c.rk4_steps(start, wl.dt, 0.0) if start else 0.0
    n0, t0_ = c.pair_counts()
    tiles0, items0 = c.pair_tiles()
    c.synchronize()
    t0 = time.perf_counter()
    t = c.rk4_steps(steps, wl.dt, t_start)
    c.synchronize()
    wall = time.perf_counter() - t0
    if graph:
        # the instrumented pass, from the same state
        timed_final = c.download()
        c.use_graph(False)
        c.upload(x, y, None, None, q, m)
        if start:
            c.rk4_steps(start, wl.dt, 0.0)
        n0, t0_ = c.pair_counts()
        tiles0, items0 = c.pair_tiles()
        c.rk4_steps(steps, wl.dt, t_start)
        if not all(np.array_equal(a, b) for a, b in zip(timed_final, c.download())):
            raise RuntimeError("the instrumented pass did not reproduce the timed steps")
    pair_ms, total_ms = c.last_timing()
    kt = c.kernel_timing()
    near, tested = c.pair_counts()
    near, tested = near - n0, tested - t0_
    tiles, items = c.pair_tiles()
    tiles, items = tiles - tiles0, items - items0
    xs1, ys1, vx1, vy1 = c.download()
    st = c.stats()
    shape_used = c.pair_shape()
    mode_used = c.pair_mode()
    slices = shape_used[1]
    snaps = []
    if snapshots:
        c.upload(x, y, None, None, q, m)
        done, tt = 0, 0.0
        for s in sorted(snapshots):
            if s > done:
                tt = c.rk4_steps(s - done, wl.dt, tt)
                done = s
            snaps.append(c.download()[:2])
    c.close()
    return dict(wall=wall, graph=graph, pair_ms=pair_ms, total_ms=total_ms, kt=kt, t=t, stats=st, near=near,
                tested=tested,
                tiles=tiles, items=items, final=(xs1, ys1, vx1, vy1), shape=shape_used, mode=mode_used,
                slices=slices, snaps=snaps)


def roofline(r, n, launches, peak, kernel):
    """The all-pairs kernel's FLOP rate over the run's launches: executed
    (the pairs it distance-tested or evaluated) and effective (every ordered
    pair of the cloud, culled ones at 7 FLOP)."""
    ordered = n * (n - 1)
    near, tested = r["near"], r["tested"]
    pair_s = r["kt"]["pair_ms"] / 1e3
    executed = FLOP_FAR * (tested - near) + FLOP_NEAR * near
    effective = FLOP_FAR * (launches * ordered - near) + FLOP_NEAR * near
    return {
        "bound": "valu",
        "achieved": executed / pair_s / 1e12,
        "peak": peak / 1e12,
        "unit": "TFLOP/s",
        "frac": executed / pair_s / peak,
        "kernel": kernel,
        "launches": r["kt"]["pair_launches"],
        "mean_launch_ms": r["kt"]["pair_ms"] / max(1, r["kt"]["pair_launches"]),
        "flop_per_launch": executed / launches,
        "flop_counting": "executed: 7 per ordered pair distance-tested, 20 per ordered pair inside the cutoff",
        "near_fraction": near / (launches * ordered),
        "tested_fraction": tested / (launches * ordered),
        "effective": {"achieved": effective / pair_s / 1e12, "frac": effective / pair_s / peak,
                      "flop_counting": "every ordered pair of N(N-1), box-culled ones at 7 FLOP"},
        "list_kernel_ms_per_launch": r["kt"]["list_ms"] / max(1, r["kt"]["pair_launches"]),
        "pair_kernel_share": r["kt"]["pair_ms"] / r["total_ms"] if r["total_ms"] else None,
    }


def counts_roofline(near, tested, ordered, pair_ms, launches, peak, kernel, flop_near=FLOP_NEAR):
    """A pair kernel's executed FP64 FLOP rate from its in-kernel counts and
    its HIP-event time: FLOP_FAR per ordered pair distance-tested, flop_near
    per ordered pair inside the cutoff; `ordered` is the ordered pairs one
    launch covers (culled or not)."""
    pair_s = pair_ms / 1e3
    executed = FLOP_FAR * (tested - near) + flop_near * near
    return {
        "bound": "valu",
        "achieved": executed / pair_s / 1e12 if pair_s else None,
        "peak": peak / 1e12,
        "unit": "TFLOP/s",
        "frac": executed / pair_s / peak if pair_s else None,
        "kernel": kernel,
        "launches": launches,
        "mean_launch_ms": pair_ms / max(1, launches),
        "flop_per_launch": executed / max(1, launches),
        "flop_counting": f"executed: {FLOP_FAR} per ordered pair distance-tested, {flop_near} per ordered pair "
                         "inside the cutoff",
        "near_fraction": near / (launches * ordered) if launches else None,
        "tested_fraction": tested / (launches * ordered) if launches else None,
    }


# Bytes per particle of one stage update that do not depend on the pair
# kernel: stage record in (x, y, vx, vy), charge, mass, caller index, base
# state, stage record out (x, y, vx, vy, nc) and its box table, and the RK
# increments (written at stages 1-3, read at 3 and 4, base state written at 4;
# averaged over the four stages).
UPDATE_FIXED_B = 32 + 16 + 4 + 32 + 40 + 0.5 + (0 + 0 + 32 + 64) / 4 + (32 + 32 + 32 + 0) / 4 + 32 / 4


def update_traffic(r, n):
    """The stage-update kernel's bytes per launch and GB/s: the fixed bytes
    per particle, plus the pair kernel's i-side sums (one 16-byte pair per
    particle per j slice or Newton-3 work item) and reactions (256 per
    evaluated tile) it reads."""
    import demon_simulation_code_amd as dm

    kt = r["kt"]
    launches = max(1, kt["update_launches"])
    if r["mode"] == dm.PAIRS_NEWTON3:
        partial_b = 16.0 * 256 * (r["items"] + r["tiles"]) / launches
    else:
        partial_b = 16.0 * n * r["slices"]
    total_b = n * UPDATE_FIXED_B + partial_b
    ms = kt["update_ms"] / launches
    return {"kernel": "k_update<1..4>", "bytes_per_launch": total_b, "partial_sum_bytes_per_launch": partial_b,
            "mean_launch_ms": ms, "achieved_GBps": total_b / (ms / 1e3) / 1e9 if ms else None,
            "peak_GBps": HBM_PEAK / 1e9, "frac": total_b / (ms / 1e3) / HBM_PEAK if ms else None,
            "bytes_counting": f"{UPDATE_FIXED_B:g} B per particle of state, records and increments, plus 16 B per "
                              "particle per partial sum and per reaction read"}


def run_ensemble(steps, warmup, graph=True):
    """The batched ensemble: 256 clouds, N log-uniform in [64, 8192] (multiples
    of 4 for the dt check), lambda, charge and k per cloud, gamma = 2/s, hex
    lattice (500 um, +-50 um), FP64, dt check on, 500 RK4 steps of 1 ms; the
    launch sequence of four steps captured as one graph and replayed, tile
    shapes from the autotuner."""
    import demon_simulation_code_amd as dm
    from demon_simulation_code_amd.physics import ensemble

    ens = ensemble(256, seed=4, multiple_of=4)
    sizes = [e["n"] for e in ens]
    b = dm.Batch(sizes, [e["kappa"] for e in ens], [e["confine"] for e in ens], [e["gamma"] for e in ens])

    def upload():
        for c, e in enumerate(ens):
            x, y = dm.hex_lattice(e["n"], 500e-6, 50e-6, 100 + c)
            b.upload(c, x, y, None, None, np.full(e["n"], e["charge"]), np.full(e["n"], 2.8e-14))

    upload()
    b.tune(2)
    b.use_graph(graph)
    if warmup:
        b.steps(warmup, 1e-3)
    upload()
    t0 = time.perf_counter()
    b.steps(steps, 1e-3)
    wall = time.perf_counter() - t0
    pairs = sum(4.0 * n * (n - 1) for n in sizes) * steps
    # the N buckets the clouds occupy (floor(log2 npad), npad = N rounded up
    # to 256) and the form and shape the autotuner chose for each
    buckets = {}
    for n in sizes:
        npad = -(-n // 256) * 256
        bk = npad.bit_length() - 1
        buckets[bk] = buckets.get(bk, 0) + 1
    shapes = {}
    for bk in sorted(buckets):
        R, js, us = b.shape(bk)
        shapes[f"npad_{1 << bk}_{(1 << (bk + 1)) - 1}"] = {
            "clouds": buckets[bk], "form": "newton3" if R == 0 else "gather",
            "i_per_lane": None if R == 0 else R,
            ("superblock_offsets_per_item" if R == 0 else "j_slice"): js, "tuned_us_per_pass": us}
    out = {"value": pairs / wall, "unit": "pair interactions/s", "clouds": len(sizes), "steps": steps,
           "n_total": int(sum(sizes)), "wall_s": wall, "device_ms": b.last_ms(), "graph": graph, "shapes": shapes}
    # an instrumented pass over the same steps: plain launches with HIP
    # events around every all-pairs pass and stage update, and the pair
    # kernels' own counts
    upload()
    b.use_graph(False)
    b.set_timing(True)
    b.steps(steps, 1e-3)
    kt = b.kernel_timing()
    near, tested = b.pair_counts()
    b.set_timing(False)
    ordered = sum(n * (n - 1) for n in sizes)
    rl = counts_roofline(near, tested, ordered, kt["pair_ms"], kt["pair_launches"], FP64_PEAK,
                         "k_bpairs / k_bpairs_n3 (FP64 all-pairs, every cloud in one launch)")
    rl["instrumented_pass_ms"] = b.last_ms()
    rl["pair_kernel_share"] = kt["pair_ms"] / b.last_ms() if b.last_ms() else None
    rl["update_mean_launch_ms"] = kt["update_ms"] / max(1, kt["update_launches"])
    out["roofline"] = rl
    b.close()
    return out


def run_small(wl, steps, heat=0.0, graph=True):
    """A smaller single-cloud configuration: the hex lattice of wl, FP64, dt
    check on, `steps` RK4 steps of wl.dt from the lattice (after a warm-up and
    a re-upload), the step captured as a hipGraph; heat > 0 adds the thermal
    force (counter-based draws)."""
    import demon_simulation_code_amd as dm

    x, y, q, m = setup(wl)
    n = wl.n
    forces = dm.CONFINEMENT | dm.DRAG | dm.SHIELDED_COULOMB | (dm.THERMAL if heat else 0)
    with dm.Context(n, shielding=wl.kappa, confine=wl.confine, gamma=wl.gamma, heat=heat, seed=3,
                    forces=forces) as c:
        c.upload(x, y, None, None, q, m)
        c.use_graph(graph)
        c.rk4_steps(5, wl.dt, 0.0)
        c.upload(x, y, None, None, q, m)
        c.synchronize()
        t0 = time.perf_counter()
        c.rk4_steps(steps, wl.dt, 0.0)
        c.synchronize()
        wall = time.perf_counter() - t0
    return {"value": 4.0 * n * (n - 1) * steps / wall, "unit": "pair interactions/s", "n": n, "steps": steps,
            "rk4_steps_per_s": steps / wall, "wall_s": wall, "thermal_heat_N": heat, "graph": graph}


def run_adjoint(reps=3):
    """The differentiable fit: N = 4096 on the 500 um lattice, 200 RK4 steps
    of 1 ms with the dt check, loss on positions every 10th step, checkpoints
    every 20; gradients of the 4096 charges, kappa, C and gamma (and of the
    initial state).  One full gradient = one demon_fit_loss_grad call."""
    import demon_simulation_code_amd as dm
    from demon_simulation_code_amd.physics import ADJOINT_4K as wl

    n, steps, obs, ckpt = wl.n, wl.steps, 10, 20
    x, y = dm.hex_lattice(n, wl.spacing, wl.jitter, wl.seed)
    rng = np.random.default_rng(11)
    q_true = wl.charge * (1 + 0.1 * rng.uniform(-1, 1, n))
    m = np.full(n, wl.mass)
    with dm.Context(n, shielding=wl.kappa, confine=wl.confine, gamma=wl.gamma) as c:
        c.upload(x, y, None, None, q_true, m)
        tx, ty, t = [], [], 0.0
        for _ in range(steps // obs):
            t = c.rk4_steps(obs, wl.dt, t)
            g = c.download()
            tx.append(g[0])
            ty.append(g[1])
    tx, ty = np.array(tx), np.array(ty)
    q = wl.charge * (1 + 0.1 * np.random.default_rng(12).uniform(-1, 1, n))
    with dm.Context(n, shielding=wl.kappa * 1.02, confine=wl.confine * 0.97, gamma=wl.gamma * 1.05) as c:
        c.set_timing(False)  # the timed gradients run without per-launch events
        c.upload(x, y, None, None, q, m)
        c.fit_loss_grad(steps, wl.dt, obs, ckpt, tx, ty)  # warm-up (allocates the fit arena)
        fwd = bwd = 0.0
        t0 = time.perf_counter()
        for _ in range(reps):
            c.fit_loss_grad(steps, wl.dt, obs, ckpt, tx, ty)
            f, b = c.fit_timing()
            fwd += f
            bwd += b
        wall = (time.perf_counter() - t0) / reps
        t0 = time.perf_counter()
        for _ in range(reps):
            c.fit_loss_grad(steps, wl.dt, obs, ckpt, tx, ty, grad=False)
        loss_only = (time.perf_counter() - t0) / reps
        # one instrumented gradient: HIP events around every adjoint pair
        # launch, and the kernel's own pair counts
        c.set_timing(True)
        c.fit_loss_grad(steps, wl.dt, obs, ckpt, tx, ty)
        ak = c.fit_adjoint_kernel()
    rl = counts_roofline(ak["near_pairs"], ak["tested_pairs"], n * (n - 1), ak["pair_ms"], ak["launches"],
                         FP64_PEAK, "k_adj_pairs (FP64 adjoint pair sums)", flop_near=FLOP_ADJ_NEAR)
    rl["share_of_recompute_adjoint"] = ak["pair_ms"] / (bwd / reps) if bwd else None
    return {"grads_per_s": 1.0 / wall, "unit": "full gradients/s", "n": n, "steps": steps,
            "params": n + 3, "forward_ms": fwd / reps, "recompute_adjoint_ms": bwd / reps,
            "gradient_over_forward": (fwd + bwd) / fwd, "loss_only_s": loss_only, "wall_s": wall,
            "roofline": rl}


def fp32_force_error(wl, state):
    """FP32-pair against FP64 Yukawa forces on one state (the FP64 path is
    within 1e-12 of the oracle): max over 256 sampled particles of
    |dF_i| / sum_j |F_ij| (condition-aware)."""
    import demon_simulation_code_amd as dm
    from demon_simulation_code_amd.physics import COULOMB

    x, y = state
    n = wl.n
    q = np.full(n, wl.charge)
    m = np.full(n, wl.mass)
    out = []
    for prec in (dm.FP64, dm.FP32_PAIR):
        with dm.Context(n, shielding=wl.kappa, confine=wl.confine, gamma=wl.gamma,
                        forces=dm.SHIELDED_COULOMB, precision=prec) as c:
            c.upload(x, y, None, None, q, m)
            out.append(c.force_stage(dm.SHIELDED_COULOMB, 1, 0, x, y))
    rows = np.random.default_rng(5).choice(n, 256, replace=False)
    dx = x[rows, None] - x[None, :]
    dy = y[rows, None] - y[None, :]
    d = np.hypot(dx, dy)
    u = d * wl.kappa
    with np.errstate(divide="ignore", invalid="ignore"):
        f = np.abs(COULOMB * wl.charge * wl.charge * (1 + u) * np.exp(-u) / d ** 2)
    f[(u >= 10.0) | (d == 0)] = 0.0
    scale = f.sum(axis=1)
    err = np.hypot(out[0][0][rows] - out[1][0][rows], out[0][1][rows] - out[1][1][rows]) / scale
    return float(err.max())


def cpu_info():
    model = platform.processor() or ""
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count(), "affinity_cpus": affinity,
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


CPU_STEPS = 3
CPU_SUBSET = 16384


def cpu_baseline(wl, snaps, snap_steps, threads):
    """The reference's own code (OpenMP build) on the host cores, on a bounded
    sample of the same workload: whole Runge_Kutta4::moveParticles steps
    (the dt check and four force evaluations of Confinement, Drag and
    ShieldedCoulomb) from the same 131072-particle lattice as the timed GPU
    steps.  Beside it, for the imploded phase of the 200-step trajectory,
    where one force evaluation of the full cloud takes minutes,
    ShieldedCoulombForce::force1 on a random CPU_SUBSET-particle subset of
    the GPU trajectory's state, scaled by the pair count N(N-1)/(M(M-1)): the
    subset's pairs are a uniform sample of the state's pairs."""
    from oracle.bind import Oracle, Reference

    n = wl.n
    x, y, q, m = setup(wl)
    info = cpu_info()
    if Reference.available(openmp=True):
        ref, kind = Reference(openmp=True), "reference"
        z = np.zeros(n)
        forces = 1 | 2 | 4  # Confinement, Drag, ShieldedCoulomb (Force.h:75-91)
        t0 = time.perf_counter()
        ref.rk4_steps(x, y, z, z, q, m, wl.kappa, wl.confine, wl.gamma, 0.0, forces, wl.dt, CPU_STEPS,
                      threads=threads)
        step_s = (time.perf_counter() - t0) / CPU_STEPS
        log(f"CPU reference: {CPU_STEPS} RK4 steps of N={n} from the lattice on {threads} threads: "
            f"{step_s:.2f} s per step")

        def evaluate(xx, yy, qq):
            ref.yukawa(xx, yy, qq, wl.kappa, 1, threads=threads)
    else:
        orc, kind, threads = Oracle(), "port", 1
        step_s = None

        def evaluate(xx, yy, qq):
            orc.yukawa(xx, yy, qq, wl.kappa)

    rng = np.random.default_rng(0)
    m_sub = min(n, CPU_SUBSET)
    scale = n * (n - 1) / (m_sub * (m_sub - 1))
    imploded = []
    for (xx, yy), s in zip(snaps, snap_steps):
        idx = np.sort(rng.choice(n, m_sub, replace=False))
        xs, ys, qs = np.ascontiguousarray(xx[idx]), np.ascontiguousarray(yy[idx]), q[idx]
        t0 = time.perf_counter()
        evaluate(xs, ys, qs)
        imploded.append((time.perf_counter() - t0) * scale)
        log(f"CPU force1 at step {s} on {m_sub} particles: {imploded[-1] / scale:.2f} s "
            f"(scaled to N={n}: {imploded[-1]:.1f} s)")
    out = {
        "unit": "pair interactions/s",
        "cores": threads,
        "kind": kind,
        "cpu": info,
        "threads_note": "OpenMP threads = OMP_NUM_THREADS when set (the CPU share given to this GPU), else the "
                        "CPUs this process may run on; the reference itself starts omp_get_num_procs() threads "
                        "(Cloud.cpp:34-36)",
    }
    if step_s is not None:
        out.update({
            "value": 4.0 * n * (n - 1) / step_s,
            "rk4_steps_per_s": 1.0 / step_s,
            "s_per_step": step_s,
            "sample": (f"{CPU_STEPS} whole Runge_Kutta4 steps (modifyTimeStep + 4 x Confinement, Drag, "
                       f"ShieldedCoulomb) of the N={n} lattice from step 0, the phase the timed GPU steps run"),
        })
    if imploded:
        t_eval = sum(imploded) / len(imploded)
        out["imploded"] = {
            "value": n * (n - 1) / t_eval,
            "rk4_steps_per_s": 1.0 / (4.0 * t_eval),
            "sample": (f"ShieldedCoulombForce::force1 on a random {m_sub}-particle subset of the GPU FP64 "
                       f"trajectory's state after steps {list(snap_steps)}, scaled by the pair count to "
                       f"N={n}: " + ", ".join(f"{t:.1f} s" for t in imploded) + " per evaluation"),
        }
    return out


def cpu_threads():
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.strip().isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=131072)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-c3", action="store_true", help="skip the whole-trajectory (200-step) leg")
    ap.add_argument("--shape", type=str, default="", help="i_per_lane,j_slices")
    ap.add_argument("--start-step", type=int, default=0,
                    help="time the steps that follow this many untimed ones (to profile a later "
                         "phase of the trajectory; the headline workload is 0)")
    ap.add_argument("--mode", type=str, default="newton3", choices=["newton3", "gather"],
                    help="all-pairs kernel")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the N=1024 and N=16384 thermal legs")
    ap.add_argument("--no-graphs", action="store_true",
                    help="time plain launches instead of replaying captured step graphs (every leg)")
    ap.add_argument("--plugins-zero", action="store_true",
                    help="add every per-particle force plugin at zero strength (same trajectory; "
                         "profiles what the plugins cost in the stage updates)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the single-cloud kernel would shard by superblock rows; that is design only).", file=sys.stderr)
        sys.exit(2)

    import demon_simulation_code_amd as dm
    from demon_simulation_code_amd.physics import LARGE_131K, Workload

    wl = LARGE_131K if args.n == LARGE_131K.n else Workload(f"hex{args.n}_fp64", n=args.n, steps=args.steps)
    shape = tuple(int(v) for v in args.shape.split(",")) if args.shape else None
    n, K, W = wl.n, args.steps, args.warmup
    full = wl.steps  # the whole trajectory of the configuration (200 steps)
    mode = dm.PAIRS_NEWTON3 if args.mode == "newton3" else dm.PAIRS_GATHER
    kernel = "k_pairs_n3" if mode == dm.PAIRS_NEWTON3 else "k_pairs"
    run_c3 = not args.no_c3 and (K != full or args.start_step)
    # states of the imploded phase for the CPU sample (from the whole run)
    snap_steps = [] if args.no_cpu else [full // 2, 7 * full // 8]

    log(f"FP64 run: N={n} warmup={W} steps={K}")
    graph = not args.no_graphs
    r = run_gpu(wl, K, W, dm.FP64, shape, snapshots=[] if run_c3 else snap_steps, start=args.start_step, graph=graph,
                mode=mode, plugins_zero=args.plugins_zero)
    pairs_total = 4.0 * n * (n - 1) * K
    value = pairs_total / r["wall"]
    log(f"FP64 done: {r['wall']:.3f} s, {value:.4g} pair interactions/s, shape={r['shape']}")
    launches = 4 * K
    rl = roofline(r, n, launches, FP64_PEAK, f"{kernel}<false> (FP64 all-pairs)")
    rl["traffic"] = measured_traffic(wl.name + "_rk4", K, W, args.start_step, args.mode)
    rl["traffic_unit"] = "bytes per launch (HBM, rocprofv3 FETCH_SIZE + WRITE_SIZE)"

    out = {
        "metric": "pair interactions/s + RK4 steps/s, N=131k, FP64 vs FP32-pair; adjoint grads/s",
        "value": value,
        "unit": "pair interactions/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": r["wall"] / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": wl.name + "_rk4",
            "n": n,
            "dt": wl.dt,
            "lambda_m": wl.lam,
            "charge_e": wl.charge_e,
            "k_spring": wl.k_spring,
            "gamma": wl.gamma,
            "thermal": False,
            "dt_check": True,
            "pair_kernel": args.mode,
            "pair_shape": list(r["shape"]),
            "start_step": args.start_step,
            "plugins_zero": args.plugins_zero,
            "step_graph": graph,
        },
        "timing_note": ("the timed steps replay the captured step graph; kernel times and pair counts come "
                        "from a pass of the same steps from the same state with plain launches and HIP events "
                        "around each (checked to end in the same state)") if graph else
                       "plain launches with HIP events around each",
        "rk4_steps_per_s": K / r["wall"],
        "roofline": rl,
        "update": update_traffic(r, n),
        "dt_reductions": r["stats"].dt_reductions,
    }

    if not args.no_fp32:
        log("FP32-pair run")
        r32 = run_gpu(wl, K, W, dm.FP32_PAIR, shape, start=args.start_step, mode=mode, graph=graph)
        x64, y64 = r["final"][0], r["final"][1]
        x32, y32 = r32["final"][0], r32["final"][1]
        perr = float(np.max(np.hypot(x64 - x32, y64 - y32)) / np.max(np.hypot(x64, y64)))
        out["fp32_pair"] = {
            "value": pairs_total / r32["wall"],
            "unit": "pair interactions/s",
            "rk4_steps_per_s": K / r32["wall"],
            "speedup_vs_fp64": r["wall"] / r32["wall"],
            "pos_rel_err_vs_fp64_after_steps": perr,
            "steps": K,
            "roofline": roofline(r32, n, launches, FP32_PEAK, f"{kernel}<true> (FP32-pair)"),
        }
        log(f"FP32-pair done: {r32['wall']:.3f} s")

    c3 = r if not run_c3 else None
    if run_c3:
        log(f"whole trajectory: {full} steps, FP64")
        c3 = run_gpu(wl, full, W, dm.FP64, shape, snapshots=snap_steps, mode=mode, graph=graph)
    if not args.no_c3:
        v3 = 4.0 * n * (n - 1) * full / c3["wall"]
        out["c3_200"] = {
            "value": v3, "unit": "pair interactions/s", "steps": full, "ms_per_step": c3["wall"] / full * 1e3,
            "rk4_steps_per_s": full / c3["wall"], "dt_reductions": c3["stats"].dt_reductions,
            "roofline": roofline(c3, n, 4 * full, FP64_PEAK, f"{kernel}<false> (FP64 all-pairs)"),
            "update": update_traffic(c3, n),
        }
        out["c3_200"]["roofline"]["traffic"] = measured_traffic(wl.name + "_rk4", full, W, 0, args.mode)
        log(f"whole trajectory done: {c3['wall']:.3f} s, {v3:.4g} pair interactions/s")
        if not args.no_fp32:
            log(f"whole trajectory: {full} steps, FP32-pair")
            c32 = run_gpu(wl, full, W, dm.FP32_PAIR, shape, mode=mode, graph=graph)
            x64, y64 = c3["final"][0], c3["final"][1]
            out["c3_200"]["fp32_pair"] = {
                "value": 4.0 * n * (n - 1) * full / c32["wall"],
                "speedup_vs_fp64": c3["wall"] / c32["wall"],
                "roofline": roofline(c32, n, 4 * full, FP32_PEAK, f"{kernel}<true> (FP32-pair)"),
                "pos_rel_err_vs_fp64_after_200_steps": float(
                    np.max(np.hypot(x64 - c32["final"][0], y64 - c32["final"][1])) / np.max(np.hypot(x64, y64))),
                "pos_note": "the imploded cloud is chaotic: FP32 and FP64 trajectories separate",
                "force_rel_err_vs_fp64_on_final_state": fp32_force_error(wl, (x64, y64)),
            }
            log(f"FP32-pair whole trajectory done: {c32['wall']:.3f} s")

    if not args.no_configs:
        from demon_simulation_code_amd.physics import PLAIN_1K, THERMAL_16K
        # the thermal magnitude for about 300 K with the reference's per-stage
        # model: heatVal = sqrt(43.2 gamma k_B T m / dt)
        heat = float(np.sqrt(43.2 * THERMAL_16K.gamma * 1.380649e-23 * 300.0 * THERMAL_16K.mass / THERMAL_16K.dt))
        log("N=1024, 1000 steps; N=16384 thermal, 5000 steps")
        out["configs"] = {"hex1024_fp64_rk4": run_small(PLAIN_1K, 1000, graph=graph),
                          "hex16384_fp64_rk4_thermal": run_small(THERMAL_16K, 5000, heat=heat, graph=graph)}
        log("configs done")
    if not args.no_ensemble:
        log("batched ensemble (256 clouds, 500 steps)")
        out["ensemble"] = run_ensemble(500, 5, graph=graph)
        log(f"ensemble done: {out['ensemble']['wall_s']:.3f} s")
    if not args.no_adjoint:
        log("trajectory-fit gradient (N=4096, 200 steps)")
        out["adjoint"] = run_adjoint()
        log(f"adjoint done: {out['adjoint']['grads_per_s']:.3f} gradients/s")

    if not args.no_cpu:
        log("CPU baseline (reference code on host cores)")
        try:
            out["cpu_baseline"] = cpu_baseline(wl, c3["snaps"] if c3 else [], snap_steps, cpu_threads())
            if "value" in out["cpu_baseline"]:
                out["gpu_vs_cpu"] = value / out["cpu_baseline"]["value"]
            if "imploded" in out["cpu_baseline"] and "c3_200" in out:
                out["c3_200"]["gpu_vs_cpu_imploded_sample"] = (
                    out["c3_200"]["value"] / out["cpu_baseline"]["imploded"]["value"])
        except Exception as e:  # the GPU result stands without it
            out["cpu_baseline"] = {"error": repr(e)}
        log("CPU baseline done")

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
