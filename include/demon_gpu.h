/*
 * demon_gpu.h -- C ABI of the MI355X (gfx950) implementation of DEMON's hot
 * path: the O(N^2) shielded-Coulomb (Yukawa) force with its 10-lambda cutoff,
 * the reference's per-particle force plugins (confinement, drag, thermal and
 * the rest), and Runge-Kutta-4 / -2 stepping with the reference's adaptive
 * time-step check.
 *
 * Plain C types only: host arrays are double* of length n in the caller's
 * particle order; nothing here exposes HIP or torch types.  Every call returns
 * DEMON_OK (0) or a negative DEMON_ERR_* code; demon_last_error() gives the
 * message.  A context is not thread-safe; use one per thread / per cloud.
 *
 * Reference interfaces replaced (AU-PSL/demon_simulation_code):
 *   Cloud (Cloud.h:16-87)                  -> demon_create / demon_upload /
 *                                             demon_download
 *   Force::force1..force4 (Force.h:15-69)  -> demon_force_stage_at /
 *                                             demon_force_stage
 *     ShieldedCoulombForce (ShieldedCoulombForce.cpp:37-237)
 *     ConfinementForce     (ConfinementForce.cpp:13-54)
 *     DragForce            (DragForce.cpp:13-55)
 *     ThermalForce         (ThermalForce.cpp:42-122)
 *     MagneticForce        (MagneticForce.cpp:13-49)
 *     TimeVaryingDragForce (TimeVaryingDragForce.cpp:12-39)
 *     RectConfinementForce  (RectConfinementForce.cpp:13-59)
 *     ConfinementForceVoid  (ConfinementForceVoid.cpp:13-55)
 *     RotationalForce       (RotationalForce.cpp:13-66)
 *     DrivingForce          (DrivingForce.cpp:14-61)
 *     ElectricForce         (ElectricForce.cpp:13-62)
 *     VertElectricForce     (VertElectricForce.cpp:13-53)
 *     GravitationalForce    (GravitationalForce.cpp:13-50)
 *     ThermalForceLocalized (ThermalForceLocalized.cpp:39-133)
 *     TimeVaryingThermalForce (TimeVaryingThermalForce.cpp:14-42)
 *   Force::readForce / constructor args    -> demon_set_params
 *   Runge_Kutta4::moveParticles (Runge_Kutta4.cpp:23-137), including
 *     Integrator::modifyTimeStep (Integrator.cpp:68-182)
 *                                          -> demon_rk4_move / demon_rk4_steps
 *   Runge_Kutta2::moveParticles (Runge_Kutta2.cpp:23-78)
 *                                          -> demon_rk2_move / demon_rk2_steps
 */
#ifndef DEMON_GPU_H
#define DEMON_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DEMON_ABI_VERSION 3

/* Return codes. */
#define DEMON_OK 0
#define DEMON_ERR_ARG (-1)         /* bad argument (null pointer, bad n, bad flag) */
#define DEMON_ERR_HIP (-2)         /* HIP runtime error (message has the HIP string) */
#define DEMON_ERR_ALLOC (-3)       /* device allocation failed */
#define DEMON_ERR_STATE (-4)       /* state the reference cannot step (see message) */
#define DEMON_ERR_UNSUPPORTED (-5) /* option not available in this build */

/* Force selection: the reference's ForceFlag bit values (Force.h:75-91). */
#define DEMON_FORCE_CONFINEMENT 1
#define DEMON_FORCE_DRAG 2
#define DEMON_FORCE_SHIELDED_COULOMB 4
#define DEMON_FORCE_THERMAL 16
#define DEMON_FORCE_GRAVITATIONAL 8
#define DEMON_FORCE_THERMAL_LOCALIZED 32
#define DEMON_FORCE_DRIVING 64
#define DEMON_FORCE_ROTATIONAL 128
#define DEMON_FORCE_TIME_VARYING_DRAG 256 /* the reference's driver then leaves out DRAG */
#define DEMON_FORCE_TIME_VARYING_THERMAL 512
#define DEMON_FORCE_MAGNETIC 1024
#define DEMON_FORCE_CONFINEMENT_VOID 2048 /* the driver then leaves out CONFINEMENT */
#define DEMON_FORCE_ELECTRIC 4096
#define DEMON_FORCE_RECT_CONFINEMENT 8192 /* the driver then leaves out CONFINEMENT */
#define DEMON_FORCE_VERT_ELECTRIC 16384

/* Pair-force arithmetic. */
#define DEMON_PRECISION_FP64 0      /* FP64 throughout */
#define DEMON_PRECISION_FP32_PAIR 1 /* FP32 distance/exp/force per pair, FP64 sums and state */

typedef struct demon_params {
    double shielding;  /* kappa = 1/lambda [1/m]; ShieldedCoulombForce(cloud, shieldingConstant) */
    double confine;    /* C [V/m^2]; ConfinementForce(cloud, confineConst): F += q*C*r */
    double drag_gamma; /* gamma [1/s]; DragForce(cloud, gamma): F += m*(-gamma)*v */
    double heat_val;   /* [N]; ThermalForce(cloud, heatVal): F += heatVal*U*(cos t, sin t) */
    uint64_t seed;     /* key of the counter-based (Philox4x32-10) thermal RNG */
    int32_t forces;    /* OR of DEMON_FORCE_* ; the reference always adds the dt check */
    int32_t precision; /* DEMON_PRECISION_* */
    int32_t dt_check;  /* 1: Integrator::modifyTimeStep each step (reference behaviour) */
    int32_t device;    /* HIP device ordinal */
    double magnetic_field; /* B [T]; MagneticForce(cloud, B): F += q*B*(vy, -vx) */
    double drag_scale;     /* [Hz/s]; TimeVaryingDragForce(cloud, scale, drag_gamma):
                              F += m*(-(scale*t + drag_gamma))*v at the stage time t */
    /* the further force plugins, with the driver's constructor arguments
       (driver.cpp:373-402); each is used only when its flag is set */
    double confine_x, confine_y;   /* RectConfinementForce(cloud, X, Y) [V/m^2] */
    double void_decay;             /* ConfinementForceVoid(cloud, confine, decay) [1/m] */
    double rot_inner, rot_outer;   /* RotationalForce(cloud, rmin, rmax, rot_const) [m] */
    double rot_const;              /* [N] */
    double drive_const;            /* DrivingForce(cloud, driveConst, amplitude, shift) [m^2] */
    double drive_amplitude;        /* [N] */
    double drive_shift;            /* [m] */
    double electric_field;         /* ElectricForce(cloud, E, plasma_radius) [V/m^2] */
    double plasma_radius;          /* [m] */
    double vert_electric_field;    /* VertElectricForce(cloud, E, vert_decay) [V/m^2] */
    double vert_decay;             /* [m] */
    double gravity;                /* GravitationalForce(cloud, g) [m/s^2] */
    double heat_in, heat_out;      /* ThermalForceLocalized(cloud, heat_in, heat_out, heat_radius):
                                      the thermal magnitude inside / outside heat_radius [N] */
    double heat_radius;            /* [m] */
    double heat_scale, heat_offset;/* TimeVaryingThermalForce(cloud, heat_scale, heat_offset):
                                      the thermal magnitude heat_scale*t + heat_offset
                                      at the stage time t [N/s, N] */
} demon_params;

/* Counters of the work done since demon_upload (or the last reset). */
typedef struct demon_stats {
    int64_t steps;           /* RK4 steps taken */
    int64_t dt_reductions;   /* steps whose dt was divided by 10 at least once */
    int64_t force_evals;     /* all-pairs force evaluations (4 per step) */
    double current_time;     /* integrator time */
    double last_dt;          /* dt of the last step */
} demon_stats;

typedef struct demon_ctx demon_ctx;

/* Library version string and ABI version. */
const char *demon_version(void);
int demon_abi_version(void);

/* Message of the last failed call on ctx (or of the calling thread when ctx is
 * NULL, e.g. after a failed demon_create).  Never NULL. */
const char *demon_last_error(const demon_ctx *ctx);

/* Allocates an n-particle cloud on device p->device (Cloud::Cloud, Cloud.cpp:23-37).
 * n >= 2.  With dt_check the reference requires n % 4 == 0 (Integrator.cpp:90);
 * this build accepts any n and tests the partial last group's pairs. */
int demon_create(demon_ctx **ctx, int64_t n, const demon_params *p);
void demon_destroy(demon_ctx *ctx);

/* Replaces the force parameters (ShieldedCoulombForce/ConfinementForce/
 * DragForce/ThermalForce::readForce, e.g. ShieldedCoulombForce.cpp:268-277). */
int demon_set_params(demon_ctx *ctx, const demon_params *p);
int demon_get_params(const demon_ctx *ctx, demon_params *p);

/* Copies the host Cloud arrays (x, y, Vx, Vy, charge, mass; Cloud.h:21-23) to
 * the device and resets the step counters.  Velocities may be NULL (zero). */
int demon_upload(demon_ctx *ctx, const double *x, const double *y, const double *vx,
                 const double *vy, const double *charge, const double *mass);

/* Copies the current state back (any pointer may be NULL). */
int demon_download(demon_ctx *ctx, double *x, double *y, double *vx, double *vy);

/* Force::force<stage>(t) for the forces in `forces` (a subset of the
 * context's flags) evaluated on the given stage state at time t, ADDED into
 * fx/fy (host arrays, caller's order), as Force::forceS adds into
 * Cloud::forceX/Y, in the driver's construction order (Confinement,
 * ConfinementVoid, Drag, Driving, Magnetic, RectConfinement, Rotational,
 * ShieldedCoulomb, Thermal, ThermalLocalized, TimeVaryingDrag,
 * TimeVaryingThermal, Electric, Gravitational, VertElectric;
 * driver.cpp:372-402); each flag is one plugin object, so CONFINEMENT |
 * CONFINEMENT_VOID adds the confinement term twice, as two such objects do.
 * stage (1..4) keys the thermal RNGs (with `step`); t is the stage time the
 * integrator passes (Runge_Kutta4.cpp:31,50,69,88), read by TimeVaryingDrag,
 * TimeVaryingThermal and Driving.  vxs/vys may be NULL when no
 * velocity-dependent force is requested.
 * Uses the context's charge and mass. */
int demon_force_stage_at(demon_ctx *ctx, int32_t forces, int32_t stage, uint64_t step, double t,
                         const double *xs, const double *ys, const double *vxs,
                         const double *vys, double *fx, double *fy);

/* The same without the time; rejects TIME_VARYING_DRAG, TIME_VARYING_THERMAL
 * and DRIVING. */
int demon_force_stage(demon_ctx *ctx, int32_t forces, int32_t stage, uint64_t step,
                      const double *xs, const double *ys, const double *vxs,
                      const double *vys, double *fx, double *fy);

/* Runge_Kutta4::moveParticles(end_time): while (*current_time < end_time)
 * take a step of dt = modifyTimeStep(1e-4f, init_dt) (Runge_Kutta4.cpp:24-25).
 * *current_time is read on entry and updated.  *steps_done may be NULL. */
int demon_rk4_move(demon_ctx *ctx, double *current_time, double end_time, double init_dt,
                   int64_t *steps_done);

/* Exactly `steps` iterations of the same loop body (no end-time test). */
int demon_rk4_steps(demon_ctx *ctx, int64_t steps, double init_dt, double *current_time);

/* Runge_Kutta2::moveParticles(end_time) (Runge_Kutta2.cpp:23-78, the
 * driver's -I): the same loop and dt check with the midpoint rule -- force at
 * (t, x), then at (t + dt/2, x + l1/2), then x += l2, v += k2. */
int demon_rk2_move(demon_ctx *ctx, double *current_time, double end_time, double init_dt,
                   int64_t *steps_done);
int demon_rk2_steps(demon_ctx *ctx, int64_t steps, double init_dt, double *current_time);

/* Captures one step into a HIP graph and replays it for later demon_rk4_* /
 * demon_rk2_* calls (1 = on, 0 = plain launches). */
int demon_use_graph(demon_ctx *ctx, int32_t enable);

int demon_get_stats(const demon_ctx *ctx, demon_stats *s);

/* Waits for all work queued on the context's stream. */
int demon_synchronize(demon_ctx *ctx);

/* Kernel shape of the all-pairs kernel: i particles per lane (1, 2 or 4) and
 * number of j slices (0 = choose from n).  Exposed for tuning. */
int demon_set_pair_shape(demon_ctx *ctx, int32_t i_per_lane, int32_t j_slices);
int demon_get_pair_shape(const demon_ctx *ctx, int32_t *i_per_lane, int32_t *j_slices);

/* Per-launch HIP events around the plain (uncaptured) launches of the
 * context: on by default; off, the kernel splits below read zero
 * (instrumentation; no counterpart in the reference). */
int demon_set_timing(demon_ctx *ctx, int32_t enable);

/* Milliseconds of device time of the last demon_rk4_* call, split into the
 * all-pairs kernels and everything else (HIP events on the context stream). */
int demon_last_timing(const demon_ctx *ctx, double *pair_ms, double *total_ms);

/* The same split further (instrumentation; no counterpart in the reference):
 * device milliseconds of the all-pairs kernel and its launch count, of the
 * Newton-3 tile-list kernel, and of the stage-update kernels (the fused
 * CacheOperator / force epilogue / RK increment pass) and their launch
 * count, from HIP events around each launch of the last demon_rk4_* /
 * demon_rk2_* call.  Steps replayed from a captured graph are not split
 * (zero). */
int demon_last_kernel_ms(const demon_ctx *ctx, double *pair_ms, int64_t *pair_launches, double *list_ms,
                         double *update_ms, int64_t *update_launches);

/* All-pairs kernel (no counterpart in the reference's interface, whose
 * ShieldedCoulombForce always uses Newton's third law,
 * ShieldedCoulombForce.cpp:51-56).  NEWTON3 evaluates every unordered pair
 * once and applies it to both particles (the default from npad = 6144 on,
 * while its reaction buffers, about 8 * npad^2 / 256 bytes, stay under
 * 16 GiB); GATHER evaluates every ordered pair and keeps no reaction
 * buffers (the default for smaller clouds, where it is faster).  Both give
 * the reference's forces to the same tolerance. */
#define DEMON_PAIRS_GATHER 0
#define DEMON_PAIRS_NEWTON3 1
int demon_set_pair_mode(demon_ctx *ctx, int32_t mode);
int demon_get_pair_mode(const demon_ctx *ctx, int32_t *mode);

/* Pair counts of the all-pairs kernels since the last upload (no counterpart
 * in the reference; instrumentation for the throughput report): ordered
 * pairs that contributed a Yukawa term (inside the cutoff), and ordered
 * pairs that reached a distance test at all (pairs in 64-particle blocks
 * whose bounding boxes are farther apart than the cutoff are never tested). */
int demon_pair_counts(const demon_ctx *ctx, int64_t *near_pairs, int64_t *tested_pairs);

/* Newton-3 superblock tiles (256 x 256 particles, off the diagonal) whose
 * bounding boxes came within the cutoff, and the work items (runs of live
 * tiles of one superblock row) they were cut into, summed over the
 * evaluations since the last upload (instrumentation; no counterpart in the
 * reference).  Each tile leaves 256 reactions and each item 256 partial sums
 * for the stage update.  Either pointer may be NULL. */
int demon_pair_tiles(const demon_ctx *ctx, int64_t *tiles, int64_t *items);

/* Differentiable trajectory fit (no counterpart in the reference, which has
 * no differentiation; it differentiates Runge_Kutta4::moveParticles with
 * Confinement, Drag and ShieldedCoulomb forces).
 * From the uploaded state, runs `steps` RK4 steps of dt (each reduced by
 * the dt check when the context's dt_check is on, as moveParticles does;
 * the adjoint replays every step's dt) and returns the loss
 *   L = sum_{o=1..steps/obs_every} sum_i |r_i(o obs_every dt) - r*_{o,i}|^2 / (2 n n_obs)
 * (target_x/target_y: [steps/obs_every][n] host arrays in caller order) and,
 * by the discrete adjoint of the RK4 map with checkpoints every ckpt_every
 * steps and recompute, its gradient with respect to the charges (n values),
 * the shielding constant, the confinement constant, gamma and the initial
 * state.  Gradient pointers may be NULL (all NULL: loss only).  The context's
 * state is the initial state again on return.  FP64 only.
 * The forward sweep keeps the stage records of the last checkpoint segments,
 * as many as fit in 16 GiB (or DEMON_FIT_KEEP_BYTES, in bytes) and a quarter
 * of the free device memory, at least one; only the segments before them are
 * recomputed.  The result does not depend on how many are kept. */
int demon_fit_loss_grad(demon_ctx *ctx, int64_t steps, double dt, int64_t obs_every,
                        int64_t ckpt_every, const double *target_x, const double *target_y,
                        double *loss, double *grad_q, double *grad_kappa, double *grad_confine,
                        double *grad_gamma, double *grad_x0, double *grad_y0, double *grad_vx0,
                        double *grad_vy0);

/* Device milliseconds of the last demon_fit_loss_grad: forward sweep, and
 * recompute + adjoint sweep. */
int demon_fit_timing(const demon_ctx *ctx, double *forward_ms, double *backward_ms);

/* The adjoint pair kernel of the last demon_fit_loss_grad (instrumentation;
 * no counterpart in the reference): its device milliseconds and launches from
 * HIP events (0 when timing is off), and the ordered pairs it found inside
 * the cutoff and gave a distance test (not box-culled).  NULL pointers are
 * skipped. */
int demon_fit_adjoint_kernel(const demon_ctx *ctx, double *pair_ms, int64_t *launches,
                             int64_t *near_pairs, int64_t *tested_pairs);

/* ---- Batched ensemble: many independent clouds in one device arena ----
 * Each cloud is its own Cloud + Runge_Kutta4 system (own kappa, C, gamma,
 * dt check and time); all clouds step together: per stage one persistent
 * work-queue all-pairs launch per tile shape and one update launch, a step
 * captured as a HIP graph.  p supplies forces, precision, dt_check, heat_val,
 * seed (cloud c uses seed + c * 0x9E3779B97F4A7C15) and device. */
typedef struct demon_batch demon_batch;
const char *demon_batch_last_error(const demon_batch *b);
int demon_batch_create(demon_batch **b, int32_t nclouds, const int64_t *n, const double *kappa,
                       const double *confine, const double *gamma, const demon_params *p);
void demon_batch_destroy(demon_batch *b);
int demon_batch_upload(demon_batch *b, int32_t cloud, const double *x, const double *y,
                       const double *vx, const double *vy, const double *charge,
                       const double *mass);
int demon_batch_download(demon_batch *b, int32_t cloud, double *x, double *y, double *vx,
                         double *vy);
/* `steps` RK4 steps of every cloud (each with its own dt check). */
int demon_batch_steps(demon_batch *b, int64_t steps, double init_dt);
int demon_batch_cloud_stats(const demon_batch *b, int32_t cloud, demon_stats *s);
int demon_batch_use_graph(demon_batch *b, int32_t enable);
int demon_batch_last_ms(const demon_batch *b, double *ms);
/* Instrumentation (no counterpart in the reference).  With timing on, plain
 * (uncaptured) steps record HIP events around every all-pairs pass and every
 * stage update; demon_batch_kernel_ms returns their device milliseconds and
 * launch counts for the last demon_batch_steps (zero launches when the step
 * ran as a captured graph).  demon_batch_pair_counts sums, over the clouds,
 * the ordered pairs found inside the cutoff and given a distance test since
 * each cloud's upload. */
int demon_batch_set_timing(demon_batch *b, int32_t enable);
int demon_batch_kernel_ms(const demon_batch *b, double *pair_ms, int64_t *pair_launches,
                          double *update_ms, int64_t *update_launches);
int demon_batch_pair_counts(const demon_batch *b, int64_t *near_pairs, int64_t *tested_pairs);
/* Form and tile shape of the clouds whose padded size falls in
 * [2^bucket, 2^(bucket+1)): i_per_lane 1, 2 or 4 is the gather form (each
 * ordered pair evaluated for its i; j slices of jslice particles, a multiple
 * of 64); i_per_lane DEMON_BATCH_NEWTON3 is the Newton-3 form (each unordered
 * pair once, as ShieldedCoulombForce.cpp:51-56; work items of one 256-particle
 * row against jslice superblock offsets). */
#define DEMON_BATCH_NEWTON3 0
int demon_batch_set_shape(demon_batch *b, int32_t bucket, int32_t i_per_lane, int32_t jslice);
int demon_batch_get_shape(const demon_batch *b, int32_t bucket, int32_t *i_per_lane,
                          int32_t *jslice, double *tuned_us);
/* Autotuner: times every candidate form and shape per N bucket present and
 * keeps the fastest (reps timed runs each, after one warm-up). */
int demon_batch_tune(demon_batch *b, int32_t reps);

/* Analysis of the current state (the reference has these only as offline
 * tools or not at all; no device state is changed).
 *
 * Kinetic energy: *total = sum m |v|^2 / 2 and *thermal = sum m |v - v_cm|^2 / 2
 * with the mass-weighted mean velocity v_cm; the 2D kinetic temperature is
 * thermal / (n k_B).  Fixed summation order, so repeatable.  Either pointer
 * may be NULL. */
int demon_kinetic_energy(demon_ctx *ctx, double *total, double *thermal);

/* demon_kinetic_energy without any device allocation (no counterpart in the
 * reference; for callers that cannot allocate: code under HIP graph capture,
 * tight serving loops).  Same kernels and the same results, bit for bit;
 * the block sums live in a fixed 512 KB device array that every GPU the
 * library runs on holds until the process ends, instead of in scratch the
 * context allocates on first use.  Limit: n <= 4194304 (larger clouds get
 * DEMON_ERR_UNSUPPORTED).  Calls that overlap on one GPU (several contexts
 * or threads at once) share the array and race; serialise them. */
int demon_kinetic_energy_noalloc(demon_ctx *ctx, double *total, double *thermal);

/* Pair-distance histogram of the current positions (the counts behind g(r)):
 * counts[b], b < nbins, is the number of unordered pairs i < j with
 * e_b^2 <= d^2 < e_{b+1}^2, where d^2 = dx*dx + dy*dy and e_k = k*(r_max/nbins)
 * (e_nbins = r_max), all in double; pairs with d >= r_max are not counted.
 * nbins in [1, 2048].  Integer counts, exact. */
int demon_pair_histogram(demon_ctx *ctx, double r_max, int32_t nbins, uint64_t *counts);

/* FFTAnalysis -a (FFTAnalysis.cpp:213-283) on device `device`: series is a
 * row-major [rows][stride] host array (row t = one snapshot's X positions,
 * the TIME_STEP column X_POSITION), of which the first nuse <= stride
 * particles are used.  amp[k] = |sum_p sum_t series[t][p] e^{-2 pi i k t/rows}|
 * / nuse for k < rows/2 (the reference sums the particles' transforms; the
 * transform of the sum is the same).  rows <= 2^22 (a direct transform,
 * O(rows^2)).  No context needed. */
int demon_trajectory_spectrum(const double *series, int64_t rows, int64_t stride, int64_t nuse, double *amp,
                              int32_t device);

/* Host-only helper (no device needed): the synthetic hexagonal-lattice input
 * used by the benchmarks -- the n sites of a triangular lattice of spacing
 * `spacing` nearest the origin (ordered by lattice radius, then row, then
 * column), centred, each coordinate jittered by U(-jitter, +jitter) drawn from
 * Philox4x32-10 keyed by `seed`. */
int demon_hex_lattice(int64_t n, double spacing, double jitter, uint64_t seed, double *x,
                      double *y);

#ifdef __cplusplus
}
#endif

#endif /* DEMON_GPU_H */
