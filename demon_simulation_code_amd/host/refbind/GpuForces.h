// GpuForces.h -- the binding a maintainer of AU-PSL/demon_simulation_code adds
// to the reference tree to run its hot path on an MI355X through the C ABI of
// libdemon_gpu.so (include/demon_gpu.h).  Header-only; it is compiled against
// the reference's own headers (Force.h, Integrator.h, Cloud.h, cfitsio).
//
//   GpuShieldedCoulombForce  drop-in for ShieldedCoulombForce
//                            (ShieldedCoulombForce.h:15-39): force1..force4
//                            read the same stage getters' data (x / xCache)
//                            and add into Cloud::forceX/forceY.
//   GpuConfinementForce, GpuDragForce, GpuThermalForce, GpuMagneticForce,
//   GpuTimeVaryingDragForce
//                            the same for ConfinementForce, DragForce,
//                            ThermalForce, MagneticForce and
//                            TimeVaryingDragForce, carrying their parameters
//                            (the stage time reaches the device for the
//                            time-varying drag).
//   GpuRectConfinementForce, GpuConfinementForceVoid, GpuRotationalForce,
//   GpuDrivingForce, GpuElectricForce, GpuVertElectricForce,
//   GpuGravitationalForce, GpuThermalForceLocalized,
//   GpuTimeVaryingThermalForce
//                            the same for the reference's further force
//                            plugins, with their constructor arguments and
//                            FITS keywords.
//   GpuRungeKutta4           drop-in for Runge_Kutta4 (Runge_Kutta4.h:14-31):
//                            moveParticles(endTime) uploads the Cloud, runs
//                            the whole while (t < end) loop -- dt check,
//                            four force stages, RK4 combine -- on the device
//                            and downloads x, y, Vx, Vy.  Every force in its
//                            ForceArray must be one of the Gpu* forces.
//   GpuRungeKutta2           the same for Runge_Kutta2 (Runge_Kutta2.h), the
//                            driver's -I.
// Errors of the C ABI are reported on stderr and end the program with
// exit(1), the reference's convention for FITS errors (driver.cpp:202-213).
#ifndef DEMON_GPU_FORCES_H
#define DEMON_GPU_FORCES_H

#include <cstdio>
#include <cstdlib>

#include "Force.h"
#include "Integrator.h"
#include "demon_gpu.h"

namespace demon_gpu_binding {

inline void require(int rc, const demon_ctx *ctx, const char *what) {
    if (rc != DEMON_OK) {
        std::fprintf(stderr, "Error: %s: %s\n", what, demon_last_error(ctx));
        std::exit(1);
    }
}

// Stage arrays of the reference's Cloud as contiguous doubles: stage 1 is
// x / Vx; stages 2-4 are the doubleV caches (Cloud.cpp:277-475), whose
// memory holds the same particle order.
inline const double *stage_x(const Cloud *c, int s) { return s == 1 ? c->x : reinterpret_cast<const double *>(c->xCache); }
inline const double *stage_y(const Cloud *c, int s) { return s == 1 ? c->y : reinterpret_cast<const double *>(c->yCache); }
inline const double *stage_vx(const Cloud *c, int s) { return s == 1 ? c->Vx : reinterpret_cast<const double *>(c->VxCache); }
inline const double *stage_vy(const Cloud *c, int s) { return s == 1 ? c->Vy : reinterpret_cast<const double *>(c->VyCache); }

// Common part of the Gpu* forces: a device context of the cloud's size.
class GpuForceBase : public Force {
public:
    GpuForceBase(Cloud *C, int flag) : Force(C), flag_(flag) {}
    ~GpuForceBase() { demon_destroy(ctx_); }
    int flag() const { return flag_; }
    virtual void contribute(demon_params &p) const = 0;

protected:
    void stage(int s, double t) {
        demon_params p{};
        p.shielding = 1.0;
        p.forces = flag_;
        contribute(p);
        if (!ctx_) require(demon_create(&ctx_, (int64_t)cloud->n, &p), nullptr, "demon_create");
        else require(demon_set_params(ctx_, &p), ctx_, "demon_set_params");
        require(demon_upload(ctx_, cloud->x, cloud->y, cloud->Vx, cloud->Vy, cloud->charge, cloud->mass),
                ctx_, "demon_upload");
        require(demon_force_stage_at(ctx_, flag_, s, step_, t, stage_x(cloud, s), stage_y(cloud, s),
                                     stage_vx(cloud, s), stage_vy(cloud, s), cloud->forceX, cloud->forceY),
                ctx_, "demon_force_stage_at");
        if (s == 4) ++step_;
    }

private:
    int flag_;
    demon_ctx *ctx_ = nullptr;
    uint64_t step_ = 0;
};

#define DEMON_GPU_FORCE_STAGES                        \
    void force1(const double t) { stage(1, t); }      \
    void force2(const double t) { stage(2, t); }      \
    void force3(const double t) { stage(3, t); }      \
    void force4(const double t) { stage(4, t); }

// FITS keywords as the CPU classes write them (e.g. ShieldedCoulombForce.cpp:239-277).
inline void write_key(fitsfile *file, int *error, long bit, const char *key, double value, const char *comment) {
    if (!*error) fits_movabs_hdu(file, 1, IMAGE_HDU, error);
    if (!*error) {
        long flags = 0;
        fits_read_key_lng(file, const_cast<char *>("FORCES"), &flags, NULL, error);
        flags |= bit;
        if (*error == KEY_NO_EXIST || *error == VALUE_UNDEFINED) *error = 0;
        if (!*error)
            fits_update_key(file, TLONG, const_cast<char *>("FORCES"), &flags,
                            const_cast<char *>("Force configuration."), error);
    }
    if (!*error) fits_write_key_dbl(file, const_cast<char *>(key), value, 6, const_cast<char *>(comment), error);
}

inline void read_key(fitsfile *file, int *error, const char *key, double *value) {
    if (!*error) fits_movabs_hdu(file, 1, IMAGE_HDU, error);
    if (!*error) fits_read_key_dbl(file, const_cast<char *>(key), value, NULL, error);
}

}  // namespace demon_gpu_binding

class GpuShieldedCoulombForce : public demon_gpu_binding::GpuForceBase {
public:
    GpuShieldedCoulombForce(Cloud *C, double shieldingConstant)
        : GpuForceBase(C, DEMON_FORCE_SHIELDED_COULOMB), shielding(shieldingConstant) {}
    DEMON_GPU_FORCE_STAGES
    void contribute(demon_params &p) const { p.shielding = shielding; }
    void writeForce(fitsfile *const f, int *const e) const {
        demon_gpu_binding::write_key(f, e, ShieldedCoulombForceFlag, "shieldingConstant", shielding,
                                     "[m^-1] (ShieldedCoulombForce)");
    }
    void readForce(fitsfile *const f, int *const e) { demon_gpu_binding::read_key(f, e, "shieldingConstant", &shielding); }
    double shielding;
};

class GpuConfinementForce : public demon_gpu_binding::GpuForceBase {
public:
    GpuConfinementForce(Cloud *C, double confineConst) : GpuForceBase(C, DEMON_FORCE_CONFINEMENT), confine(confineConst) {}
    DEMON_GPU_FORCE_STAGES
    void contribute(demon_params &p) const { p.confine = confine; }
    void writeForce(fitsfile *const f, int *const e) const {
        demon_gpu_binding::write_key(f, e, ConfinementForceFlag, "confineConst", confine, "[V/m^2] (ConfinementForce)");
    }
    void readForce(fitsfile *const f, int *const e) { demon_gpu_binding::read_key(f, e, "confineConst", &confine); }
    double confine;
};

class GpuDragForce : public demon_gpu_binding::GpuForceBase {
public:
    GpuDragForce(Cloud *C, double gamma) : GpuForceBase(C, DEMON_FORCE_DRAG), dragConst(-gamma) {}
    DEMON_GPU_FORCE_STAGES
    void contribute(demon_params &p) const { p.drag_gamma = -dragConst; }
    void writeForce(fitsfile *const f, int *const e) const {
        demon_gpu_binding::write_key(f, e, DragForceFlag, "dragConst", dragConst, "[Hz] (DragForce)");
    }
    void readForce(fitsfile *const f, int *const e) { demon_gpu_binding::read_key(f, e, "dragConst", &dragConst); }
    double dragConst;
};

class GpuThermalForce : public demon_gpu_binding::GpuForceBase {
public:
    GpuThermalForce(Cloud *C, double redFactor, uint64_t seed = 0)
        : GpuForceBase(C, DEMON_FORCE_THERMAL), heatVal(redFactor), seed(seed) {}
    DEMON_GPU_FORCE_STAGES
    void contribute(demon_params &p) const {
        p.heat_val = heatVal;
        p.seed = seed;
    }
    void writeForce(fitsfile *const f, int *const e) const {
        demon_gpu_binding::write_key(f, e, ThermalForceFlag, "heatingValue", heatVal, "[N] (ThermalForce)");
    }
    void readForce(fitsfile *const f, int *const e) { demon_gpu_binding::read_key(f, e, "heatingValue", &heatVal); }
    double heatVal;
    uint64_t seed;
};

class GpuMagneticForce : public demon_gpu_binding::GpuForceBase {
public:
    GpuMagneticForce(Cloud *C, double magneticField) : GpuForceBase(C, DEMON_FORCE_MAGNETIC), BField(magneticField) {}
    DEMON_GPU_FORCE_STAGES
    void contribute(demon_params &p) const { p.magnetic_field = BField; }
    void writeForce(fitsfile *const f, int *const e) const {
        demon_gpu_binding::write_key(f, e, MagneticForceFlag, "magneticField", BField, "[T] (MagneticForce)");
    }
    void readForce(fitsfile *const f, int *const e) { demon_gpu_binding::read_key(f, e, "magneticField", &BField); }
    double BField;
};

// TimeVaryingDragForce(C, scale, offset) (TimeVaryingDragForce.h:16-17): the
// drag constant at stage time t is -(scale t + offset).
class GpuTimeVaryingDragForce : public demon_gpu_binding::GpuForceBase {
public:
    GpuTimeVaryingDragForce(Cloud *C, double scale, double offset)
        : GpuForceBase(C, DEMON_FORCE_TIME_VARYING_DRAG), scaleConst(scale), offsetConst(offset) {}
    DEMON_GPU_FORCE_STAGES
    void contribute(demon_params &p) const {
        p.drag_scale = scaleConst;
        p.drag_gamma = offsetConst;
    }
    void writeForce(fitsfile *const f, int *const e) const {
        // as the CPU class: DragForce's keyword and bit (its dragConst is +offset between stages:
        // TimeVaryingDragForce.h:16-17 builds DragForce(C, -offset)), then its own
        // (TimeVaryingDragForce.cpp:41-64)
        demon_gpu_binding::write_key(f, e, DragForceFlag, "dragConst", offsetConst, "[Hz] (DragForce)");
        demon_gpu_binding::write_key(f, e, TimeVaryingDragForceFlag, "TVDragScaleConst", scaleConst,
                                     "[Hz/s] (TimeVaryingDragForce)");
        demon_gpu_binding::write_key(f, e, TimeVaryingDragForceFlag, "TVDragOffsetConst", offsetConst,
                                     "[Hz] (TimeVaryingDragForce)");
    }
    void readForce(fitsfile *const f, int *const e) {
        demon_gpu_binding::read_key(f, e, "TVDragScaleConst", &scaleConst);
        demon_gpu_binding::read_key(f, e, "TVDragOffsetConst", &offsetConst);
    }
    double scaleConst, offsetConst;
};

class GpuRectConfinementForce : public demon_gpu_binding::GpuForceBase {
public:
    GpuRectConfinementForce(Cloud *C, double confineConstX, double confineConstY)
        : GpuForceBase(C, DEMON_FORCE_RECT_CONFINEMENT), confineX(confineConstX), confineY(confineConstY) {}
    DEMON_GPU_FORCE_STAGES
    void contribute(demon_params &p) const {
        p.confine_x = confineX;
        p.confine_y = confineY;
    }
    void writeForce(fitsfile *const f, int *const e) const {
        demon_gpu_binding::write_key(f, e, RectConfinementForceFlag, "confineConstX", confineX,
                                     "[V/m^2] (RectConfinementForce)");
        demon_gpu_binding::write_key(f, e, RectConfinementForceFlag, "confineConstY", confineY,
                                     "[V/m^2] (RectConfinementForce)");
    }
    void readForce(fitsfile *const f, int *const e) {
        demon_gpu_binding::read_key(f, e, "confineConstX", &confineX);
        demon_gpu_binding::read_key(f, e, "confineConstY", &confineY);
    }
    double confineX, confineY;
};

// ConfinementForceVoid(C, confineConst, decay): the confinement term plus the
// radial void term, as one plugin (ConfinementForceVoid.cpp:13-55).
class GpuConfinementForceVoid : public demon_gpu_binding::GpuForceBase {
public:
    GpuConfinementForceVoid(Cloud *C, double confineConst, double voidDecay)
        : GpuForceBase(C, DEMON_FORCE_CONFINEMENT_VOID), confine(confineConst), decay(voidDecay) {}
    DEMON_GPU_FORCE_STAGES
    void contribute(demon_params &p) const {
        p.confine = confine;
        p.void_decay = decay;
    }
    void writeForce(fitsfile *const f, int *const e) const {
        demon_gpu_binding::write_key(f, e, ConfinementForceFlag, "confineConst", confine, "[V/m^2] (ConfinementForce)");
        demon_gpu_binding::write_key(f, e, ConfinementForceVoidFlag, "decay", decay, "[m^-1] (ConfinementForceVoid)");
    }
    void readForce(fitsfile *const f, int *const e) {
        demon_gpu_binding::read_key(f, e, "confineConst", &confine);
        demon_gpu_binding::read_key(f, e, "decay", &decay);
    }
    double confine, decay;
};

class GpuRotationalForce : public demon_gpu_binding::GpuForceBase {
public:
    GpuRotationalForce(Cloud *C, double rmin, double rmax, double rotConst)
        : GpuForceBase(C, DEMON_FORCE_ROTATIONAL), innerRad(rmin), outerRad(rmax), rotationalConst(rotConst) {}
    DEMON_GPU_FORCE_STAGES
    void contribute(demon_params &p) const {
        p.rot_inner = innerRad;
        p.rot_outer = outerRad;
        p.rot_const = rotationalConst;
    }
    void writeForce(fitsfile *const f, int *const e) const {
        demon_gpu_binding::write_key(f, e, RotationalForceFlag, "rotationalConst", rotationalConst, "[N] (RotationalForce)");
        demon_gpu_binding::write_key(f, e, RotationalForceFlag, "innerRadius", innerRad, "[m] (RotationalForce)");
        demon_gpu_binding::write_key(f, e, RotationalForceFlag, "outerRadius", outerRad, "[m] (RotationalForce)");
    }
    void readForce(fitsfile *const f, int *const e) {
        demon_gpu_binding::read_key(f, e, "rotationalConst", &rotationalConst);
        demon_gpu_binding::read_key(f, e, "innerRadius", &innerRad);
        demon_gpu_binding::read_key(f, e, "outerRadius", &outerRad);
    }
    double innerRad, outerRad, rotationalConst;
};

// DrivingForce(C, dampConst, amplitude, shift): the member driveConst is
// -dampConst (DrivingForce.h:17-18) and is what the FITS keyword holds.
class GpuDrivingForce : public demon_gpu_binding::GpuForceBase {
public:
    GpuDrivingForce(Cloud *C, double dampConst, double amp, double drivingShift)
        : GpuForceBase(C, DEMON_FORCE_DRIVING), amplitude(amp), driveConst(-dampConst), shift(drivingShift) {}
    DEMON_GPU_FORCE_STAGES
    void contribute(demon_params &p) const {
        p.drive_const = -driveConst;
        p.drive_amplitude = amplitude;
        p.drive_shift = shift;
    }
    void writeForce(fitsfile *const f, int *const e) const {
        demon_gpu_binding::write_key(f, e, DrivingForceFlag, "drivingAmplitude", amplitude, "[N] (DrivingForce)");
        demon_gpu_binding::write_key(f, e, DrivingForceFlag, "drivingConst", driveConst, "[m^2] (DrivingForce)");
        demon_gpu_binding::write_key(f, e, DrivingForceFlag, "drivingShift", shift, "[m] (DrivingForce)");
    }
    void readForce(fitsfile *const f, int *const e) {
        demon_gpu_binding::read_key(f, e, "drivingAmplitude", &amplitude);
        demon_gpu_binding::read_key(f, e, "drivingConst", &driveConst);
        demon_gpu_binding::read_key(f, e, "drivingShift", &shift);
    }
    double amplitude, driveConst, shift;
};

class GpuElectricForce : public demon_gpu_binding::GpuForceBase {
public:
    GpuElectricForce(Cloud *C, double electricField, double plasmaRad)
        : GpuForceBase(C, DEMON_FORCE_ELECTRIC), electric(electricField), radius(plasmaRad) {}
    DEMON_GPU_FORCE_STAGES
    void contribute(demon_params &p) const {
        p.electric_field = electric;
        p.plasma_radius = radius;
    }
    void writeForce(fitsfile *const f, int *const e) const {
        demon_gpu_binding::write_key(f, e, ElectricForceFlag, "ElectricFieldStrength", electric, "[V/m^2] (ElectricForce)");
        demon_gpu_binding::write_key(f, e, ElectricForceFlag, "plasmaRadius", radius, "[m] (ElectricForce)");
    }
    void readForce(fitsfile *const f, int *const e) {
        demon_gpu_binding::read_key(f, e, "electricFieldStrength", &electric);
        demon_gpu_binding::read_key(f, e, "plasmaRadius", &radius);
    }
    double electric, radius;
};

class GpuVertElectricForce : public demon_gpu_binding::GpuForceBase {
public:
    GpuVertElectricForce(Cloud *C, double vertElectricField, double vertDecay)
        : GpuForceBase(C, DEMON_FORCE_VERT_ELECTRIC), vertElectric(vertElectricField), vertDec(vertDecay) {}
    DEMON_GPU_FORCE_STAGES
    void contribute(demon_params &p) const {
        p.vert_electric_field = vertElectric;
        p.vert_decay = vertDec;
    }
    void writeForce(fitsfile *const f, int *const e) const {
        demon_gpu_binding::write_key(f, e, VertElectricForceFlag, "VertElectricFieldStrength", vertElectric,
                                     "[V/m^2] (VertElectricForce)");
        demon_gpu_binding::write_key(f, e, VertElectricForceFlag, "verticalDecay", vertDec, "[m] (VertElectricForce)");
    }
    void readForce(fitsfile *const f, int *const e) {
        demon_gpu_binding::read_key(f, e, "vertElectricFieldStrength", &vertElectric);
        demon_gpu_binding::read_key(f, e, "verticalDecay", &vertDec);
    }
    double vertElectric, vertDec;
};

class GpuGravitationalForce : public demon_gpu_binding::GpuForceBase {
public:
    GpuGravitationalForce(Cloud *C, double g) : GpuForceBase(C, DEMON_FORCE_GRAVITATIONAL), gravitational(g) {}
    DEMON_GPU_FORCE_STAGES
    void contribute(demon_params &p) const { p.gravity = gravitational; }
    void writeForce(fitsfile *const f, int *const e) const {
        demon_gpu_binding::write_key(f, e, GravitationalForceFlag, "GravitationalFieldStrength", gravitational,
                                     "[m/s^2] (GravitationalForce)");
    }
    void readForce(fitsfile *const f, int *const e) {
        demon_gpu_binding::read_key(f, e, "gravitationalFieldStrength", &gravitational);
    }
    double gravitational;
};

// ThermalForceLocalized(C, heatVal1, heatVal2, radius): heatVal1 strictly
// inside the radius, heatVal2 elsewhere (ThermalForceLocalized.cpp:124-133).
class GpuThermalForceLocalized : public demon_gpu_binding::GpuForceBase {
public:
    GpuThermalForceLocalized(Cloud *C, double thermRed1, double thermRed2, double specifiedRadius, uint64_t seed = 0)
        : GpuForceBase(C, DEMON_FORCE_THERMAL_LOCALIZED), heatingRadius(specifiedRadius), heatVal1(thermRed1),
          heatVal2(thermRed2), seed(seed) {}
    DEMON_GPU_FORCE_STAGES
    void contribute(demon_params &p) const {
        p.heat_in = heatVal1;
        p.heat_out = heatVal2;
        p.heat_radius = heatingRadius;
        p.seed = seed;
    }
    void writeForce(fitsfile *const f, int *const e) const {
        demon_gpu_binding::write_key(f, e, ThermalForceLocalizedFlag, "heatingValue1", heatVal1, "[N] (ThermalForceLocalized)");
        demon_gpu_binding::write_key(f, e, ThermalForceLocalizedFlag, "heatingValue2", heatVal2, "[N] (ThermalForceLocalized)");
        demon_gpu_binding::write_key(f, e, ThermalForceLocalizedFlag, "heatingRadius", heatingRadius,
                                     "[m] (ThermalForceLocalized)");
    }
    void readForce(fitsfile *const f, int *const e) {
        demon_gpu_binding::read_key(f, e, "heatingValue1", &heatVal1);
        demon_gpu_binding::read_key(f, e, "heatingValue2", &heatVal2);
        demon_gpu_binding::read_key(f, e, "heatingRadius", &heatingRadius);
    }
    double heatingRadius, heatVal1, heatVal2;
    uint64_t seed;
};

// TimeVaryingThermalForce(C, scale, offset): the thermal force with heatVal =
// scale t + offset at the stage time (TimeVaryingThermalForce.cpp:14-42); it
// writes ThermalForce's keyword and bit as well, as the CPU class does.
class GpuTimeVaryingThermalForce : public demon_gpu_binding::GpuForceBase {
public:
    GpuTimeVaryingThermalForce(Cloud *C, double scale, double offset, uint64_t seed = 0)
        : GpuForceBase(C, DEMON_FORCE_TIME_VARYING_THERMAL), heatVal(offset), heatValScale(scale),
          heatValOffset(offset), seed(seed) {}
    DEMON_GPU_FORCE_STAGES
    void contribute(demon_params &p) const {
        p.heat_scale = heatValScale;
        p.heat_offset = heatValOffset;
        p.seed = seed;
    }
    void writeForce(fitsfile *const f, int *const e) const {
        demon_gpu_binding::write_key(f, e, ThermalForceFlag, "heatingValue", heatVal, "[N] (ThermalForce)");
        demon_gpu_binding::write_key(f, e, TimeVaryingThermalForceFlag, "heatingValueScale", heatValScale,
                                     "[N/s] (TimeVaryingThermalForce)");
        demon_gpu_binding::write_key(f, e, TimeVaryingThermalForceFlag, "heatingValueOffset", heatValOffset,
                                     "[N] (TimeVaryingThermalForce)");
    }
    void readForce(fitsfile *const f, int *const e) {
        demon_gpu_binding::read_key(f, e, "heatingValue", &heatVal);
        demon_gpu_binding::read_key(f, e, "heatingValueScale", &heatValScale);
        demon_gpu_binding::read_key(f, e, "heatingValueOffset", &heatValOffset);
    }
    double heatVal, heatValScale, heatValOffset;
    uint64_t seed;
};

#undef DEMON_GPU_FORCE_STAGES

namespace demon_gpu_binding {

// The device parameter block of a ForceArray of Gpu* forces.
inline demon_params gather_params(const ForceArray &forces, int device) {
    demon_params p{};
    p.shielding = 1.0;
    p.dt_check = 1;  // moveParticles always calls modifyTimeStep (Runge_Kutta4.cpp:25, Runge_Kutta2.cpp:25)
    p.device = device;
    for (Force *const F : forces) {
        GpuForceBase *g = dynamic_cast<GpuForceBase *>(F);
        if (!g) {
            std::fprintf(stderr, "Error: the Gpu integrators need Gpu* forces only\n");
            std::exit(1);
        }
        p.forces |= g->flag();
        g->contribute(p);
    }
    return p;
}

}  // namespace demon_gpu_binding

// order 4: GpuRungeKutta4, 2: GpuRungeKutta2.
template <int ORDER>
class GpuRungeKutta : public Integrator {
public:
    GpuRungeKutta(Cloud *const C, const ForceArray &FA, const double timeStep, const double startTime,
                  int device = 0)
        : Integrator(C, FA, timeStep, startTime), device_(device) {}
    ~GpuRungeKutta() { demon_destroy(ctx_); }

    void moveParticles(const double endTime) {
        const demon_params p = demon_gpu_binding::gather_params(forces, device_);
        if (!ctx_) demon_gpu_binding::require(demon_create(&ctx_, (int64_t)cloud->n, &p), nullptr, "demon_create");
        else demon_gpu_binding::require(demon_set_params(ctx_, &p), ctx_, "demon_set_params");
        demon_gpu_binding::require(
            demon_upload(ctx_, cloud->x, cloud->y, cloud->Vx, cloud->Vy, cloud->charge, cloud->mass), ctx_,
            "demon_upload");
        if (ORDER == 4)
            demon_gpu_binding::require(demon_rk4_move(ctx_, &currentTime, endTime, init_dt, nullptr), ctx_,
                                       "demon_rk4_move");
        else
            demon_gpu_binding::require(demon_rk2_move(ctx_, &currentTime, endTime, init_dt, nullptr), ctx_,
                                       "demon_rk2_move");
        demon_gpu_binding::require(demon_download(ctx_, cloud->x, cloud->y, cloud->Vx, cloud->Vy), ctx_,
                                   "demon_download");
    }

private:
    int device_;
    demon_ctx *ctx_ = nullptr;
};

typedef GpuRungeKutta<4> GpuRungeKutta4;
typedef GpuRungeKutta<2> GpuRungeKutta2;

#endif  // DEMON_GPU_FORCES_H
