// ref_harness.cpp -- extern "C" entry points around the reference's own hot-path
// classes, compiled together with the reference sources where they lie
// (AU-PSL/demon_simulation_code) into oracle/_ref/.  TEST INFRASTRUCTURE ONLY:
// used to pin the oracle and as the CPU baseline; never linked by the product.
//
// The reference sources are not copied: oracle/Makefile compiles them in place.
#include "Cloud.h"
#include "ConfinementForce.h"
#include "ConfinementForceVoid.h"
#include "DragForce.h"
#include "DrivingForce.h"
#include "ElectricForce.h"
#include "GravitationalForce.h"
#include "MagneticForce.h"
#include "RectConfinementForce.h"
#include "RotationalForce.h"
#include "Runge_Kutta2.h"
#include "Runge_Kutta4.h"
#include "ShieldedCoulombForce.h"
#include "ThermalForce.h"
#include "ThermalForceLocalized.h"
#include "TimeVaryingDragForce.h"
#include "TimeVaryingThermalForce.h"
#include "VertElectricForce.h"

#include <cmath>
#include <cstdint>
#include <string>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

// Static Cloud members are defined by the reference's driver (driver.cpp:122-127),
// which is not part of this harness; give them the driver's defaults.
double Cloud::interParticleSpacing = 0.003;
double Cloud::dustParticleMassDensity = 2200;
double Cloud::justX = 0.0;
double Cloud::justY = 0.0;
double Cloud::velX = 0.0;
double Cloud::velY = 0.0;

namespace {

void set_threads(int nthreads) {
#ifdef _OPENMP
    // Cloud::Cloud forces omp_get_num_procs() threads (Cloud.cpp:34-36); the
    // harness re-applies the caller's count afterwards.
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

Cloud *make_cloud(int64_t n, const double *x, const double *y, const double *vx,
                  const double *vy, const double *q, const double *m) {
    Cloud *c = new Cloud((cloud_index)n);
    for (int64_t i = 0; i < n; ++i) {
        c->x[i] = x ? x[i] : 0.0;
        c->y[i] = y ? y[i] : 0.0;
        c->Vx[i] = vx ? vx[i] : 0.0;
        c->Vy[i] = vy ? vy[i] : 0.0;
        c->charge[i] = q ? q[i] : 0.0;
        c->mass[i] = m ? m[i] : 1.0;
        // Cloud never zeroes the force arrays (Cloud.cpp:31); the harness does.
        c->forceX[i] = 0.0;
        c->forceY[i] = 0.0;
    }
    return c;
}

// Exposes the protected Integrator::modifyTimeStep.
struct RK4Probe : public Runge_Kutta4 {
    RK4Probe(Cloud *C, const ForceArray &FA, double dt, double t0) : Runge_Kutta4(C, FA, dt, t0) {}
    double check(float dist, double dt) const { return modifyTimeStep(dist, dt); }
};

// Parameter vector of ref_rk_all / ref_force_all.
enum RefPar {
    kKappa, kConfine, kGamma, kHeat, kBField, kDragScale, kConfineX, kConfineY, kVoidDecay,
    kRotInner, kRotOuter, kRotConst, kDriveConst, kDriveAmplitude, kDriveShift, kElectric,
    kPlasmaRadius, kVertElectric, kVertDecay, kGravity, kHeatIn, kHeatOut, kHeatRadius,
    kHeatScale, kHeatOffset, kNumPar
};

// Every force of `forces` with the driver's constructors, in its
// construction order (driver.cpp:372-402).
void make_forces(Cloud *c, int forces, const double *p, ForceArray &fa) {
    if (forces & ConfinementForceFlag) fa.push_back(new ConfinementForce(c, p[kConfine]));
    if (forces & ConfinementForceVoidFlag) fa.push_back(new ConfinementForceVoid(c, p[kConfine], p[kVoidDecay]));
    if (forces & DragForceFlag) fa.push_back(new DragForce(c, p[kGamma]));
    if (forces & DrivingForceFlag)
        fa.push_back(new DrivingForce(c, p[kDriveConst], p[kDriveAmplitude], p[kDriveShift]));
    if (forces & MagneticForceFlag) fa.push_back(new MagneticForce(c, p[kBField]));
    if (forces & RectConfinementForceFlag) fa.push_back(new RectConfinementForce(c, p[kConfineX], p[kConfineY]));
    if (forces & RotationalForceFlag) fa.push_back(new RotationalForce(c, p[kRotInner], p[kRotOuter], p[kRotConst]));
    if (forces & ShieldedCoulombForceFlag) fa.push_back(new ShieldedCoulombForce(c, p[kKappa]));
    if (forces & ThermalForceFlag) fa.push_back(new ThermalForce(c, p[kHeat]));
    if (forces & ThermalForceLocalizedFlag)
        fa.push_back(new ThermalForceLocalized(c, p[kHeatIn], p[kHeatOut], p[kHeatRadius]));
    if (forces & TimeVaryingDragForceFlag) fa.push_back(new TimeVaryingDragForce(c, p[kDragScale], p[kGamma]));
    if (forces & TimeVaryingThermalForceFlag)
        fa.push_back(new TimeVaryingThermalForce(c, p[kHeatScale], p[kHeatOffset]));
    if (forces & ElectricForceFlag) fa.push_back(new ElectricForce(c, p[kElectric], p[kPlasmaRadius]));
    if (forces & GravitationalForceFlag) fa.push_back(new GravitationalForce(c, p[kGravity]));
    if (forces & VertElectricForceFlag) fa.push_back(new VertElectricForce(c, p[kVertElectric], p[kVertDecay]));
}

}  // namespace

extern "C" {

int ref_num_par(void) { return kNumPar; }

// `steps` steps of Runge_Kutta4 (order 4) or Runge_Kutta2 (order 2) with
// every force of `forces` (par: RefPar order).  State updated in place.
int ref_rk_all(int order, int64_t n, double *x, double *y, double *vx, double *vy, const double *q,
               const double *m, const double *par, int forces, double dt, int64_t steps, double *t,
               int nthreads) {
    if (n < 4 || (n & 3) || (order != 2 && order != 4)) return -1;
    Cloud *c = make_cloud(n, x, y, vx, vy, q, m);
    set_threads(nthreads);
    ForceArray fa;
    make_forces(c, forces, par, fa);
    if (order == 4) {
        Runge_Kutta4 I(c, fa, dt, *t);
        for (int64_t s = 0; s < steps; ++s) I.moveParticles(std::nextafter(I.currentTime, INFINITY));
        *t = I.currentTime;
    } else {
        Runge_Kutta2 I(c, fa, dt, *t);
        for (int64_t s = 0; s < steps; ++s) I.moveParticles(std::nextafter(I.currentTime, INFINITY));
        *t = I.currentTime;
    }
    for (Force *f : fa) delete f;
    for (int64_t i = 0; i < n; ++i) {
        x[i] = c->x[i];
        y[i] = c->y[i];
        vx[i] = c->Vx[i];
        vy[i] = c->Vy[i];
    }
    delete c;
    return 0;
}

// The forces of `forces` evaluated `reps` times on one state at time t:
// repetition k calls force<k % 4 + 1>(t) on the same force objects (every
// stage reads the given state; the thermal forces draw afresh each call) and
// writes the sum into fx/fy[k][n].
int ref_force_all(int64_t n, const double *xs, const double *ys, const double *vxs,
                  const double *vys, const double *q, const double *m, const double *par,
                  int forces, double t, int64_t reps, double *fx, double *fy) {
    if (n < 2 || (n & 1) || reps < 1) return -1;
    Cloud *c = make_cloud(n, xs, ys, vxs, vys, q, m);
    for (int64_t i = 0; i < n; i += 2) {
        c->xCache[i / 2] = _mm_set_pd(xs[i + 1], xs[i]);
        c->yCache[i / 2] = _mm_set_pd(ys[i + 1], ys[i]);
        c->VxCache[i / 2] = _mm_set_pd(vxs ? vxs[i + 1] : 0.0, vxs ? vxs[i] : 0.0);
        c->VyCache[i / 2] = _mm_set_pd(vys ? vys[i + 1] : 0.0, vys ? vys[i] : 0.0);
    }
    ForceArray fa;
    make_forces(c, forces, par, fa);
    for (int64_t k = 0; k < reps; ++k) {
        for (int64_t i = 0; i < n; ++i) {
            c->forceX[i] = 0.0;
            c->forceY[i] = 0.0;
        }
        for (Force *f : fa) {
            switch (k % 4) {
                case 0: f->force1(t); break;
                case 1: f->force2(t); break;
                case 2: f->force3(t); break;
                default: f->force4(t); break;
            }
        }
        for (int64_t i = 0; i < n; ++i) {
            fx[(size_t)k * n + i] = c->forceX[i];
            fy[(size_t)k * n + i] = c->forceY[i];
        }
    }
    for (Force *f : fa) delete f;
    delete c;
    return 0;
}

int ref_is_openmp(void) {
#ifdef _OPENMP
    return 1;
#else
    return 0;
#endif
}

// ShieldedCoulombForce::force<stage>(t) on the given stage positions; adds the
// result into fx/fy.  Stage 1 reads x/y, stages 2-4 read the stage caches.
int ref_yukawa_stage(int stage, int64_t n, const double *xs, const double *ys, const double *q,
                     double kappa, double *fx, double *fy, int nthreads) {
    if (n < 2 || (n & 1) || stage < 1 || stage > 4) return -1;
    Cloud *c = make_cloud(n, xs, ys, nullptr, nullptr, q, nullptr);
    set_threads(nthreads);
    if (stage > 1) {
        for (int64_t i = 0; i < n; i += 2) {
            c->xCache[i / 2] = _mm_set_pd(xs[i + 1], xs[i]);
            c->yCache[i / 2] = _mm_set_pd(ys[i + 1], ys[i]);
        }
    }
    {
        ShieldedCoulombForce f(c, kappa);
        switch (stage) {
            case 1: f.force1(0.0); break;
            case 2: f.force2(0.0); break;
            case 3: f.force3(0.0); break;
            default: f.force4(0.0); break;
        }
    }
    for (int64_t i = 0; i < n; ++i) {
        fx[i] += c->forceX[i];
        fy[i] += c->forceY[i];
    }
    delete c;
    return 0;
}

// Integrator::modifyTimeStep(dist, dt) on positions x/y.
double ref_modify_time_step(int64_t n, const double *x, const double *y, float dist, double dt) {
    if (n < 4 || (n & 3)) return -1.0;
    Cloud *c = make_cloud(n, x, y, nullptr, nullptr, nullptr, nullptr);
    ForceArray none;
    double r;
    {
        RK4Probe probe(c, none, dt, 0.0);
        r = probe.check(dist, dt);
    }
    delete c;
    return r;
}

// Runge_Kutta4 with the driver's force set and order (driver.cpp:322-326,372-392):
// Confinement, Drag, ShieldedCoulomb, Thermal (flags as in Force.h:75-91).
// mode 0: run exactly `steps` steps (each moveParticles call is given an end
//         time just above the current time, so the while loop runs once);
// mode 1: a single moveParticles(end_time) call.
// State is updated in place; *t receives the integrator's currentTime.
int ref_rk4(int64_t n, double *x, double *y, double *vx, double *vy, const double *q,
            const double *m, double kappa, double confine, double gamma, double heat,
            int forces, double dt, int mode, int64_t steps, double end_time, double *t,
            int nthreads) {
    if (n < 4 || (n & 3)) return -1;
    Cloud *c = make_cloud(n, x, y, vx, vy, q, m);
    set_threads(nthreads);
    ForceArray fa;
    if (forces & ConfinementForceFlag) fa.push_back(new ConfinementForce(c, confine));
    if (forces & DragForceFlag) fa.push_back(new DragForce(c, gamma));
    if (forces & ShieldedCoulombForceFlag) fa.push_back(new ShieldedCoulombForce(c, kappa));
    if (forces & ThermalForceFlag) fa.push_back(new ThermalForce(c, heat));
    {
        Runge_Kutta4 I(c, fa, dt, *t);
        if (mode == 0) {
            for (int64_t s = 0; s < steps; ++s)
                I.moveParticles(std::nextafter(I.currentTime, INFINITY));
        } else {
            I.moveParticles(end_time);
        }
        *t = I.currentTime;
    }
    for (Force *f : fa) delete f;
    for (int64_t i = 0; i < n; ++i) {
        x[i] = c->x[i];
        y[i] = c->y[i];
        vx[i] = c->Vx[i];
        vy[i] = c->Vy[i];
    }
    delete c;
    return 0;
}

// ref_rk4 (mode 0) with two more of the driver's forces in its order
// (driver.cpp:373-394): Confinement, Drag, Magnetic(bfield), ShieldedCoulomb,
// Thermal, TimeVaryingDrag(drag_scale, gamma).  Flags as in Force.h:75-91.
int ref_rk4_ext(int64_t n, double *x, double *y, double *vx, double *vy, const double *q,
                const double *m, double kappa, double confine, double gamma, double heat,
                double bfield, double drag_scale, int forces, double dt, int64_t steps, double *t,
                int nthreads) {
    if (n < 4 || (n & 3)) return -1;
    Cloud *c = make_cloud(n, x, y, vx, vy, q, m);
    set_threads(nthreads);
    ForceArray fa;
    if (forces & ConfinementForceFlag) fa.push_back(new ConfinementForce(c, confine));
    if (forces & DragForceFlag) fa.push_back(new DragForce(c, gamma));
    if (forces & MagneticForceFlag) fa.push_back(new MagneticForce(c, bfield));
    if (forces & ShieldedCoulombForceFlag) fa.push_back(new ShieldedCoulombForce(c, kappa));
    if (forces & ThermalForceFlag) fa.push_back(new ThermalForce(c, heat));
    if (forces & TimeVaryingDragForceFlag) fa.push_back(new TimeVaryingDragForce(c, drag_scale, gamma));
    {
        Runge_Kutta4 I(c, fa, dt, *t);
        for (int64_t s = 0; s < steps; ++s) I.moveParticles(std::nextafter(I.currentTime, INFINITY));
        *t = I.currentTime;
    }
    for (Force *f : fa) delete f;
    for (int64_t i = 0; i < n; ++i) {
        x[i] = c->x[i];
        y[i] = c->y[i];
        vx[i] = c->Vx[i];
        vy[i] = c->Vy[i];
    }
    delete c;
    return 0;
}

// ref_rk4_ext with Runge_Kutta2 (the driver's -I, driver.cpp:437-438).
int ref_rk2_ext(int64_t n, double *x, double *y, double *vx, double *vy, const double *q,
                const double *m, double kappa, double confine, double gamma, double heat,
                double bfield, double drag_scale, int forces, double dt, int64_t steps, double *t,
                int nthreads) {
    if (n < 4 || (n & 3)) return -1;
    Cloud *c = make_cloud(n, x, y, vx, vy, q, m);
    set_threads(nthreads);
    ForceArray fa;
    if (forces & ConfinementForceFlag) fa.push_back(new ConfinementForce(c, confine));
    if (forces & DragForceFlag) fa.push_back(new DragForce(c, gamma));
    if (forces & MagneticForceFlag) fa.push_back(new MagneticForce(c, bfield));
    if (forces & ShieldedCoulombForceFlag) fa.push_back(new ShieldedCoulombForce(c, kappa));
    if (forces & ThermalForceFlag) fa.push_back(new ThermalForce(c, heat));
    if (forces & TimeVaryingDragForceFlag) fa.push_back(new TimeVaryingDragForce(c, drag_scale, gamma));
    {
        Runge_Kutta2 I(c, fa, dt, *t);
        for (int64_t s = 0; s < steps; ++s) I.moveParticles(std::nextafter(I.currentTime, INFINITY));
        *t = I.currentTime;
    }
    for (Force *f : fa) delete f;
    for (int64_t i = 0; i < n; ++i) {
        x[i] = c->x[i];
        y[i] = c->y[i];
        vx[i] = c->Vx[i];
        vy[i] = c->Vy[i];
    }
    delete c;
    return 0;
}

// A data.fits as the reference's driver writes it (driver.cpp:364-420,
// Cloud.cpp:198-271): the primary image, each force's writeForce in the
// driver's order (flags as Force.h:75-91; TimeVaryingDrag instead of Drag),
// writeCloudSetup, then rows - 1 writeTimeStep calls (row k: x + 1e-6 k,
// t = 0.01 k).  Parameters: confine, gamma, B, kappa, heat, TV scale.
int ref_fits_write(const char *path, int64_t n, const double *x, const double *y, const double *vx,
                   const double *vy, const double *q, const double *m, int forces, const double *par,
                   int64_t rows) {
    Cloud *c = make_cloud(n, x, y, vx, vy, q, m);
    ForceArray fa;
    if (forces & ConfinementForceFlag) fa.push_back(new ConfinementForce(c, par[0]));
    if ((forces & DragForceFlag) && !(forces & TimeVaryingDragForceFlag)) fa.push_back(new DragForce(c, par[1]));
    if (forces & MagneticForceFlag) fa.push_back(new MagneticForce(c, par[2]));
    if (forces & ShieldedCoulombForceFlag) fa.push_back(new ShieldedCoulombForce(c, par[3]));
    if (forces & ThermalForceFlag) fa.push_back(new ThermalForce(c, par[4]));
    if (forces & TimeVaryingDragForceFlag) fa.push_back(new TimeVaryingDragForce(c, par[5], par[1]));
    int error = 0;
    fitsfile *file = nullptr;
    const std::string name = std::string("!") + path;  // '!' replaces an existing file
    fits_create_file(&file, name.c_str(), &error);
    fits_create_img(file, 16, 0, NULL, &error);
    for (Force *f : fa) f->writeForce(file, &error);
    c->writeCloudSetup(file, error);
    for (int64_t k = 1; k < rows && !error; ++k) {
        for (int64_t i = 0; i < n; ++i) c->x[i] = x[i] + 1e-6 * (double)k;
        c->writeTimeStep(file, error, 0.01 * (double)k);
    }
    fits_close_file(file, &error);
    for (Force *f : fa) delete f;
    delete c;
    return error;
}

// The driver's -c reading (driver.cpp:332-343, 404-408): FORCES, the last
// TIME_STEP row (Cloud::initializeFromFile) and each force's readForce.
// out: x, y, vx, vy, q, m (n each); par as ref_fits_write; *t the row's time.
int ref_fits_read(const char *path, int64_t n, double *x, double *y, double *vx, double *vy, double *q,
                  double *m, long *forces, double *par, double *t) {
    int error = 0;
    fitsfile *file = nullptr;
    fits_open_file(&file, path, READONLY, &error);
    fits_read_key_lng(file, const_cast<char *>("FORCES"), forces, NULL, &error);
    if (error) return error;
    Cloud *c = Cloud::initializeFromFile(file, error, t);
    if (error || (int64_t)c->n != n) {
        delete c;
        return error ? error : -1;
    }
    fits_movabs_hdu(file, 1, NULL, &error);
    ConfinementForce conf(c, 0.0);
    DragForce drag(c, 0.0);
    MagneticForce mag(c, 0.0);
    ShieldedCoulombForce yuk(c, 1.0);
    ThermalForce therm(c, 0.0);
    TimeVaryingDragForce tvd(c, 0.0, 0.0);
    if (*forces & ConfinementForceFlag) conf.readForce(file, &error);
    if ((*forces & DragForceFlag) && !(*forces & TimeVaryingDragForceFlag)) drag.readForce(file, &error);
    if (*forces & MagneticForceFlag) mag.readForce(file, &error);
    if (*forces & ShieldedCoulombForceFlag) yuk.readForce(file, &error);
    if (*forces & ThermalForceFlag) therm.readForce(file, &error);
    if (*forces & TimeVaryingDragForceFlag) tvd.readForce(file, &error);
    // the members those readForce calls filled are private: report the same
    // keywords, read the same way (fits_read_key_dbl by name)
    const char *names[6] = {"confineConst", "dragConst", "magneticField", "shieldingConstant", "heatingValue",
                            "TVDragScaleConst"};
    const long bits[6] = {ConfinementForceFlag, DragForceFlag, MagneticForceFlag, ShieldedCoulombForceFlag,
                          ThermalForceFlag, TimeVaryingDragForceFlag};
    for (int k = 0; k < 6; ++k) {
        par[k] = 0.0;
        if (*forces & bits[k]) fits_read_key_dbl(file, const_cast<char *>(names[k]), &par[k], NULL, &error);
    }
    fits_close_file(file, &error);
    for (int64_t i = 0; i < n; ++i) {
        x[i] = c->x[i];
        y[i] = c->y[i];
        vx[i] = c->Vx[i];
        vy[i] = c->Vy[i];
        q[i] = c->charge[i];
        m[i] = c->mass[i];
    }
    delete c;
    return error;
}

// The reference's readForce of every force in in_path's FORCES (constructed as
// the driver's -c does, driver.cpp:337-408), then the same objects' writeForce
// into a new out_path: a file whose keywords the reference reads back
// unchanged.  Returns the cfitsio status (nonzero when a keyword is missing).
int ref_fits_forces_roundtrip(const char *in_path, const char *out_path) {
    int error = 0;
    fitsfile *file = nullptr;
    long forces = 0;
    fits_open_file(&file, in_path, READONLY, &error);
    fits_read_key_lng(file, const_cast<char *>("FORCES"), &forces, NULL, &error);
    if (error) return error;
    double t = 0.0;
    Cloud *c = Cloud::initializeFromFile(file, error, &t);
    if (error) {
        delete c;
        return error;
    }
    fits_movabs_hdu(file, 1, NULL, &error);
    double par[kNumPar] = {};
    par[kKappa] = 1.0;
    ForceArray fa;
    make_forces(c, (int)forces, par, fa);
    for (Force *f : fa) f->readForce(file, &error);
    fits_close_file(file, &error);
    fitsfile *out = nullptr;
    const std::string name = std::string("!") + out_path;
    if (!error) fits_create_file(&out, name.c_str(), &error);
    if (!error) fits_create_img(out, 16, 0, NULL, &error);
    for (Force *f : fa) f->writeForce(out, &error);
    if (out) fits_close_file(out, &error);
    for (Force *f : fa) delete f;
    delete c;
    return error;
}

// One Runge_Kutta4 instance (so one RNG stream) run for `steps` steps,
// recording x, y, Vx, Vy every `every` steps into rec[k][4][n] (k = 0.. steps/every-1).
int ref_rk4_record(int64_t n, const double *x, const double *y, const double *q, const double *m,
                   double kappa, double confine, double gamma, double heat, int forces, double dt,
                   int64_t steps, int64_t every, double *rec, int nthreads) {
    if (n < 4 || (n & 3) || every < 1) return -1;
    Cloud *c = make_cloud(n, x, y, nullptr, nullptr, q, m);
    set_threads(nthreads);
    ForceArray fa;
    if (forces & ConfinementForceFlag) fa.push_back(new ConfinementForce(c, confine));
    if (forces & DragForceFlag) fa.push_back(new DragForce(c, gamma));
    if (forces & ShieldedCoulombForceFlag) fa.push_back(new ShieldedCoulombForce(c, kappa));
    if (forces & ThermalForceFlag) fa.push_back(new ThermalForce(c, heat));
    {
        Runge_Kutta4 I(c, fa, dt, 0.0);
        int64_t k = 0;
        for (int64_t s = 1; s <= steps; ++s) {
            I.moveParticles(std::nextafter(I.currentTime, INFINITY));
            if (s % every == 0) {
                double *r = rec + (size_t)k * 4 * n;
                for (int64_t i = 0; i < n; ++i) {
                    r[i] = c->x[i];
                    r[n + i] = c->y[i];
                    r[2 * n + i] = c->Vx[i];
                    r[3 * n + i] = c->Vy[i];
                }
                ++k;
            }
        }
    }
    for (Force *f : fa) delete f;
    delete c;
    return 0;
}

}  // extern "C"
