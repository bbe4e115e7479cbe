/*
 * demon_oracle.h -- CPU oracle for the DEMON Yukawa + RK4 hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP product is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product (libdemon_gpu.so and the
 * demon_simulation_code_amd package) never links, imports or calls it.
 *
 * Each function restates one piece of the reference algorithm
 * (AU-PSL/demon_simulation_code) in plain C, citing the file:line it follows.
 * Arithmetic is kept unfused (the build uses -ffp-contract=off) and in the
 * reference's evaluation order, so that a serial build of the reference and
 * this oracle agree bit for bit where libm agrees.
 *
 * How each function is pinned (beyond "retyped formula"):
 *   oracle_yukawa_add        bitwise vs the compiled reference (serial build),
 *                            two-body closed form, Newton-3, cutoff at u=10,
 *                            long-double brute force (tests/test_oracle.py)
 *   oracle_yukawa_gather_add long-double brute force, equality with
 *                            oracle_yukawa_add to rounding
 *   oracle_confinement_add,
 *   oracle_drag_add,
 *   oracle_magnetic_add,
 *   oracle_tvdrag_add        linear closed forms (cyclotron rotation for the
 *                            magnetic force); bitwise vs reference trajectories
 *   oracle_rk4_steps         bitwise vs compiled reference trajectories;
 *                            RK4 amplification polynomial on the linear
 *                            (confinement + drag) system
 *   oracle_rk2_steps         bitwise vs compiled reference trajectories;
 *                            the midpoint rule's polynomial 1 + hA + (hA)^2/2
 *   oracle_modify_time_step  constructed cases incl. the coverage hole;
 *                            bitwise vs compiled reference
 *   oracle_dtcheck_min_s     brute force over the coverage predicate
 *   oracle_yukawa_cells_add  bitwise equal to oracle_yukawa_add (sparse,
 *                            dense and mixed-charge clouds, pairs at +-1 ulp
 *                            of the cutoff); its RK4 trajectories (gather = 2)
 *                            bitwise equal to the all-pairs ones
 *   oracle_dtcheck_min_s_cells, oracle_modify_time_step_cells
 *                            equal to oracle_dtcheck_min_s and
 *                            oracle_modify_time_step on constructed close
 *                            pairs, the coverage hole included
 *   oracle_philox4x32        published Random123 known-answer vectors
 *   oracle_thermal_add       moments of the published uniform/angle model
 *   oracle_rect_add, oracle_void_add, oracle_rotational_add,
 *   oracle_driving_add, oracle_electric_add, oracle_vert_electric_add,
 *   oracle_gravity_add       bitwise vs the compiled reference's force1 and
 *                            its RK4 / RK2 trajectories; closed forms (linear
 *                            rect field, radial void field, the tangential
 *                            rotation of the shear layer, gravity = m g)
 *   oracle_average_fft       numpy's FFT (a textbook transform) and the
 *                            closed-form spectrum of a pure cosine
 *   oracle_thermal_localized_add, oracle_tvthermal_add
 *                            the reference's own draws: per-particle magnitude
 *                            bounds inside / outside the radius, and |F|^2
 *                            moments at the stage time
 */
#ifndef DEMON_ORACLE_H
#define DEMON_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 1/(4*pi*eps0) with eps0 = 8.8541878e-12 (Cloud.cpp:16, ShieldedCoulombForce.cpp:21). */
double oracle_coulomb_const(void);

/* Force flags, same bit values as the reference's ForceFlag (Force.h:75-91). */
enum {
    ORACLE_CONFINEMENT = 1,
    ORACLE_DRAG = 2,
    ORACLE_SHIELDED_COULOMB = 4,
    ORACLE_GRAVITATIONAL = 8,
    ORACLE_THERMAL = 16,
    ORACLE_THERMAL_LOCALIZED = 32,
    ORACLE_DRIVING = 64,
    ORACLE_ROTATIONAL = 128,
    ORACLE_TIME_VARYING_DRAG = 256,
    ORACLE_TIME_VARYING_THERMAL = 512,
    ORACLE_MAGNETIC = 1024,
    ORACLE_CONFINEMENT_VOID = 2048,
    ORACLE_ELECTRIC = 4096,
    ORACLE_RECT_CONFINEMENT = 8192,
    ORACLE_VERT_ELECTRIC = 16384
};

typedef struct oracle_params {
    double shielding;   /* kappa = 1/lambda [1/m]            (ShieldedCoulombForce.h:29) */
    double confine;     /* confinement constant C [V/m^2]    (ConfinementForce.h:20) */
    double gamma;       /* drag frequency; dragConst = -gamma (DragForce.h:17) */
    double heat_val;    /* thermal force scale [N]           (ThermalForce.h:25) */
    uint64_t seed;      /* key of the counter-based thermal RNG */
    int32_t forces;     /* bitmask of ORACLE_* flags */
    int32_t dt_check;   /* 1: run Integrator::modifyTimeStep each step */
    int32_t gather;     /* 1: Yukawa in gather form (j ascending per i), OpenMP over i;
                           deterministic, differs from the reference order by rounding;
                           2: oracle_yukawa_cells_add and the cell-list dt check (the
                           reference's serial results, for clouds whose pairs are
                           mostly beyond the cutoff) */
    double magnetic_field; /* B [T]                             (MagneticForce.h:29) */
    double drag_scale;  /* TimeVaryingDragForce scaleConst [Hz/s]; its offset is gamma
                           (driver.cpp:394, TimeVaryingDragForce.h:16-17) */
    /* constructor arguments of the further plugins (driver.cpp:373-402) */
    double confine_x, confine_y;      /* RectConfinementForce.h confineX, confineY */
    double void_decay;                /* ConfinementForceVoid decay [1/m] */
    double rot_inner, rot_outer, rot_const; /* RotationalForce.h:20-21 */
    double drive_const;               /* DrivingForce(C, dampConst, ...): the member
                                         driveConst is -dampConst (DrivingForce.h:17-18) */
    double drive_amplitude, drive_shift;
    double electric_field, plasma_radius;         /* ElectricForce.h:16-17 */
    double vert_electric_field, vert_decay;       /* VertElectricForce.h:16-17 */
    double gravity;                               /* GravitationalForce */
    double heat_in, heat_out, heat_radius;        /* ThermalForceLocalized.h:16-17 */
    double heat_scale, heat_offset;               /* TimeVaryingThermalForce.h:16-17 */
} oracle_params;

/* ShieldedCoulombForce::forceS (ShieldedCoulombForce.cpp:37-237), serial
 * evaluation order.  Adds into fx/fy.  n must be even (the reference loads
 * particles in pairs).  Returns 0, or -1 if n is odd. */
int oracle_yukawa_add(int64_t n, const double *xs, const double *ys,
                      const double *q, double kappa, double *fx, double *fy);

/* oracle_yukawa_add's result, bit for bit, from the pairs inside the
 * cutoff only (a pair with u >= 10 adds exactly +-0 in the reference,
 * ShieldedCoulombForce.cpp:279-292): candidates from a cell list of cells
 * at least 10/kappa wide, the reference's exact test u = sqrt(dX*dX+dY*dY)*kappa
 * < 10, and each particle's contributions added in the order the
 * reference's serial loop adds them (ShieldedCoulombForce.cpp:43-56: by
 * partner pair index, the pair's own intra term at its place, the straight
 * lane before the swapped one), with the reference's intra-pair or
 * cross-pair formula.  OpenMP over particles.  n even; returns -1 if odd or
 * out of memory. */
int oracle_yukawa_cells_add(int64_t n, const double *xs, const double *ys,
                            const double *q, double kappa, double *fx, double *fy);

/* oracle_dtcheck_min_s from a cell list: pairs within about 1e-4 m only
 * (the only ones the dt check can act on), +inf if there are none. */
float oracle_dtcheck_min_s_cells(int64_t n, const double *x, const double *y, double cell);

/* oracle_modify_time_step's result from the smallest tested squared
 * distance: while sqrtf(s_min) <= dist, dist /= 10 and dt /= 10 -- the
 * divisions the reference makes (Integrator.cpp:144-155) in any pair order. */
double oracle_modify_time_step_cells(int64_t n, const double *x, const double *y, float dist, double dt,
                                     double cell);

/* Same physics for any n as a per-particle gather, j ascending, with the
 * cross-pair formula (ShieldedCoulombForce.cpp:183-184) for every pair. */
void oracle_yukawa_gather_add(int64_t n, const double *xs, const double *ys,
                              const double *q, double kappa, double *fx, double *fy);

/* Gather form restricted to the particles listed in rows[0..nrows): writes
 * (not adds) their Yukawa force into fx[k], fy[k]; cross-pair formula,
 * j ascending.  For sampled checks at sizes where the full loop is slow. */
void oracle_yukawa_rows(int64_t n, const double *xs, const double *ys, const double *q,
                        double kappa, const int64_t *rows, int64_t nrows, double *fx, double *fy);

/* Same, evaluated in long double (reference-free brute force for pins). */
void oracle_yukawa_longdouble(int64_t n, const double *xs, const double *ys,
                              const double *q, double kappa, double *fx, double *fy);

/* Number of ordered pairs (i != j) with sqrt(dx^2+dy^2)*kappa < 10 under the
 * reference's rounding (ShieldedCoulombForce.cpp:175-178). */
int64_t oracle_near_pairs(int64_t n, const double *xs, const double *ys, double kappa);

/* ConfinementForce::force (ConfinementForce.cpp:48-54): F += (q*C)*r_s. */
void oracle_confinement_add(int64_t n, const double *xs, const double *ys,
                            const double *q, double confine, double *fx, double *fy);

/* DragForce::force (DragForce.cpp:143-149): F += (m*(-gamma))*v_s. */
void oracle_drag_add(int64_t n, const double *vxs, const double *vys,
                     const double *m, double gamma, double *fx, double *fy);

/* MagneticForce::force (MagneticForce.cpp:44-49): qB = q*B;
 * Fx += qB*vy, Fy -= qB*vx. */
void oracle_magnetic_add(int64_t n, const double *vxs, const double *vys,
                         const double *q, double bfield, double *fx, double *fy);

/* TimeVaryingDragForce::forceS(t) (TimeVaryingDragForce.cpp:12-31, 37-39):
 * dragConst = -(scale*t + offset), then DragForce::force: F += (m*dragConst)*v. */
void oracle_tvdrag_add(int64_t n, const double *vxs, const double *vys, const double *m,
                       double scale, double offset, double t, double *fx, double *fy);

/* Philox4x32-10 (Salmon et al., SC'11).  ctr[4], key[2] -> out[4]. */
void oracle_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ThermalForce::force (ThermalForce.cpp:118-122) with the random draws of
 * RandCache (RandomNumbers.h:34-61) replaced by Philox: for particle i at
 * (step, stage), U in [0,1) and theta in [0,2pi); F += heat*U*(cos, sin). */
void oracle_thermal_add(int64_t n, double heat, uint64_t seed, uint64_t step,
                        int stage, double *fx, double *fy);

/* RectConfinementForce::force (RectConfinementForce.cpp:50-59):
 * Fx += (q*cX)*x, Fy += (q*cY)*y. */
void oracle_rect_add(int64_t n, const double *xs, const double *ys, const double *q,
                     double cx, double cy, double *fx, double *fy);

/* ConfinementForceVoid::forceS (ConfinementForceVoid.cpp:13-55): the
 * ConfinementForce term, then with r = sqrt(x^2+y^2):
 * F += ((q*(-decay)) * (exp(r*(-decay))/r)) * (x, y). */
void oracle_void_add(int64_t n, const double *xs, const double *ys, const double *q,
                     double confine, double decay, double *fx, double *fy);

/* RotationalForce::force (RotationalForce.cpp:50-66): for rmin < r < rmax,
 * Fx -= (c*y)/r, Fy += (c*x)/r. */
void oracle_rotational_add(int64_t n, const double *xs, const double *ys, double rmin,
                           double rmax, double c, double *fx, double *fy);

/* DrivingForce::force (DrivingForce.cpp:14-15, 52-61) at stage time t:
 * Fx += (sin(x*k - t*w) * exp(((x-shift)-(x-shift))/dampConst)) * amplitude,
 * k = 2 pi/0.002, w = 2 pi 10. */
void oracle_driving_add(int64_t n, const double *xs, double t, double damp_const,
                        double amplitude, double shift, double *fx);

/* ElectricForce::force (ElectricForce.cpp:48-62):
 * F += ((q*E) * exp(sqrt(x^2+y^2)/(-radius))) * (x, y). */
void oracle_electric_add(int64_t n, const double *xs, const double *ys, const double *q,
                         double e_field, double radius, double *fx, double *fy);

/* VertElectricForce::force (VertElectricForce.cpp:47-53):
 * Fy += (q*E) * exp(y/decay). */
void oracle_vert_electric_add(int64_t n, const double *ys, const double *q, double e_field,
                              double decay, double *fy);

/* GravitationalForce::force (GravitationalForce.cpp:46-50): Fy += m*g. */
void oracle_gravity_add(int64_t n, const double *m, double g, double *fy);

/* ThermalForceLocalized::force (ThermalForceLocalized.cpp:124-133) with
 * Philox draws of counter domain 3: heat = r < radius ? heat_in : heat_out
 * on the stage position, F += (heat*U)*(cos, sin). */
void oracle_thermal_localized_add(int64_t n, const double *xs, const double *ys,
                                  double heat_in, double heat_out, double radius,
                                  uint64_t seed, uint64_t step, int stage, double *fx,
                                  double *fy);

/* TimeVaryingThermalForce::forceS(t) (TimeVaryingThermalForce.cpp:14-42):
 * heat = scale*t + offset, then the thermal force with Philox draws of
 * counter domain 4. */
void oracle_tvthermal_add(int64_t n, double scale, double offset, double t, uint64_t seed,
                          uint64_t step, int stage, double *fx, double *fy);

/* FFTAnalysis averageFFT (FFTAnalysis.cpp:213-283): for each of the first
 * nuse particles the discrete Fourier transform of its X_POSITION series
 * (rows snapshots; series is row-major [rows][stride]), the transforms
 * summed, and amp[i] = |sum_i| / nuse for i < rows/2.  The transform is the
 * definition sum_t x_t e^{-2 pi i k t / rows} (FFTW's FFTW_FORWARD), with the
 * angle reduced exactly through (k t) mod rows.  O(nuse rows^2): small
 * cases only. */
void oracle_average_fft(const double *series, int64_t rows, int64_t stride, int64_t nuse, double *amp);

/* Integrator::modifyTimeStep (Integrator.cpp:68-134, 144-182), serial, with
 * the reference's exact pair coverage.  n must be a multiple of 4. */
double oracle_modify_time_step(int64_t n, const double *x, const double *y,
                               float dist, double dt);

/* Minimum over the pairs the reference tests of the FP32 squared distance
 * RN(RN(dx*dx)+RN(dy*dy)), +inf if none.  Any n (groups of 4, partial last
 * group).  Used to check the GPU's reduction. */
float oracle_dtcheck_min_s(int64_t n, const double *x, const double *y);

/* 1 if the reference's modifyTimeStep tests the pair (i, j), else 0. */
int oracle_dtcheck_covers(int64_t i, int64_t j);

/* Runge_Kutta4::moveParticles (Runge_Kutta4.cpp:23-137) driven for exactly
 * `steps` steps (the while-loop body, with the dt check when enabled).
 * State arrays are updated in place; *t is advanced by each step's dt.
 * Forces are summed in the driver's construction order (driver.cpp:372-402):
 * Confinement, ConfinementVoid, Drag, Driving, Magnetic, RectConfinement,
 * Rotational, ShieldedCoulomb, Thermal, ThermalLocalized, TimeVaryingDrag,
 * TimeVaryingThermal, Electric, Gravitational, VertElectric; stage s
 * runs at t, t + dt/2, t + dt/2, t + dt (Runge_Kutta4.cpp:31,50,69,88).
 * Returns 0 or -1 (bad n). */
int oracle_rk4_steps(int64_t n, double *x, double *y, double *vx, double *vy,
                     const double *q, const double *m, const oracle_params *p,
                     double init_dt, int64_t steps, double *t, uint64_t *step_counter);

/* Runge_Kutta2::moveParticles (Runge_Kutta2.cpp:23-78) driven for exactly
 * `steps` steps: the midpoint rule with the same dt check, stage times t and
 * t + dt/2, and the same force order as oracle_rk4_steps. */
int oracle_rk2_steps(int64_t n, double *x, double *y, double *vx, double *vy,
                     const double *q, const double *m, const oracle_params *p,
                     double init_dt, int64_t steps, double *t, uint64_t *step_counter);

/* Runge_Kutta4::moveParticles(endTime) with its while (t < end) loop. */
int oracle_rk4_move(int64_t n, double *x, double *y, double *vx, double *vy,
                    const double *q, const double *m, const oracle_params *p,
                    double init_dt, double end_time, double *t,
                    uint64_t *step_counter, int64_t *steps_done);

/* The stage force of the RK4 loop on a given stage state (all enabled
 * forces) at the stage time t_stage. */
int oracle_stage_force(int64_t n, const double *xs, const double *ys,
                       const double *vxs, const double *vys, const double *q,
                       const double *m, const oracle_params *p, uint64_t step,
                       int stage, double t_stage, double *fx, double *fy);

/* Trajectory-fit loss and its exact gradient (discrete adjoint of the RK4
 * map of oracle_rk4_steps with Confinement + Drag + ShieldedCoulomb,
 * gather-form Yukawa with the cross formula).  With dt_check each step's dt
 * is dt0 reduced by oracle_modify_time_step on the step's start positions
 * (n a multiple of 4), and the adjoint replays those dts; dt is piecewise
 * constant in the state, so it carries no gradient.  The reference has no
 * differentiation; this is the CPU reference gradient for the device adjoint.
 *   L = sum_{o=1..steps/obs_every} sum_i |r_i(o*obs_every*dt) - r*_{o,i}|^2 / (2 n n_obs)
 * tx, ty: [n_obs][n] targets.  Outputs: *loss, dL/dq (n), dL/dkappa,
 * dL/dconfine, dL/dgamma; dL/d(initial state) into gx0..gvy0 when non-NULL.
 * Stores every stage state: memory 16 n steps doubles.  Returns 0 or -1. */
int oracle_fit_loss_grad(int64_t n, const double *x0, const double *y0, const double *vx0,
                         const double *vy0, const double *q, const double *m, double kappa,
                         double confine, double gamma, double dt0, int dt_check, int64_t steps,
                         int64_t obs_every, const double *tx, const double *ty, double *loss,
                         double *gq, double *gkappa, double *gconfine, double *ggamma,
                         double *gx0, double *gy0, double *gvx0, double *gvy0);

#ifdef __cplusplus
}
#endif

#endif /* DEMON_ORACLE_H */
