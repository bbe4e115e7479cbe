// demon_host.h -- host C++ mirror of the reference's plugin surface for the
// hot path, backed by the C ABI of libdemon_gpu.so.
//
// Same class names, constructor arguments and call contracts as the
// reference (AU-PSL/demon_simulation_code), in namespace demon_amd:
//   Cloud                      Cloud.h:16-87 (public SoA arrays, stage caches)
//   Force / ForceArray         Force.h:15-91 (force1..force4 add into
//                              Cloud::forceX/forceY; ForceFlag bit values)
//   ShieldedCoulombForce       ShieldedCoulombForce.h:15-39
//   ConfinementForce           ConfinementForce.h:13-30
//   DragForce                  DragForce.h:13-30
//   ThermalForce               ThermalForce.h:13-30
//   MagneticForce              MagneticForce.h:14-33
//   TimeVaryingDragForce       TimeVaryingDragForce.h:14-33
//   RectConfinementForce, ConfinementForceVoid, RotationalForce,
//   DrivingForce, ElectricForce, VertElectricForce, GravitationalForce,
//   ThermalForceLocalized, TimeVaryingThermalForce
//                              their headers' constructors (e.g.
//                              RotationalForce.h:20-21, DrivingForce.h:17-18)
//   Integrator, Runge_Kutta4,  Integrator.h:17-38, Runge_Kutta4.h:14-31,
//   Runge_Kutta2               Runge_Kutta2.h
//   checkParams                driver.cpp:626-824 (Params.cfg)
//   writeForce / readForce     the FITS keywords of each force (demon_fits.h)
// The forces evaluate on the device; Runge_Kutta4::moveParticles runs the
// whole while (t < end) loop on the device and copies the state back.
// There is no CPU code path: a Force that is not one of the device forces
// makes Runge_Kutta4 throw.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/demon_gpu.h"
#include "demon_fits.h"

namespace demon_amd {

using cloud_index = int64_t;
constexpr double kPi = 3.14159265358979323846;

// ForceFlag values of the reference (Force.h:75-91).
enum ForceFlag : long {
    ConfinementForceFlag = 1,
    DragForceFlag = 2,
    ShieldedCoulombForceFlag = 4,
    GravitationalForceFlag = 8,
    ThermalForceFlag = 16,
    ThermalForceLocalizedFlag = 32,
    DrivingForceFlag = 64,
    RotationalForceFlag = 128,
    TimeVaryingDragForceFlag = 256,
    TimeVaryingThermalForceFlag = 512,
    MagneticForceFlag = 1024,
    ConfinementForceVoidFlag = 2048,
    ElectricForceFlag = 4096,
    RectConfinementForceFlag = 8192,
    VertElectricForceFlag = 16384
};
using force_flags = long;

class Error : public std::runtime_error {
public:
    Error(int code, const std::string &msg) : std::runtime_error(msg), code(code) {}
    int code;
};

// Throws Error when a C-ABI call fails.
void check(int rc, const demon_ctx *ctx);

class Cloud {
public:
    explicit Cloud(cloud_index numPar, int device = 0);
    ~Cloud();
    Cloud(const Cloud &) = delete;
    Cloud &operator=(const Cloud &) = delete;

    const cloud_index n;
    // current positions, velocities, charges and masses (Cloud.h:21-22)
    std::vector<double> x, y, Vx, Vy, charge, mass;
    // force accumulators used by force1..force4 (Cloud.h:27); zeroed here,
    // unlike the reference (Cloud.cpp:31 leaves them uninitialised)
    std::vector<double> forceX, forceY;
    // stage states read by force2..force4 (Cloud.h:28, CacheOperator.cpp)
    std::vector<double> xCache, yCache, VxCache, VyCache;

    static double interParticleSpacing;   // Cloud.h:31
    static const double electronCharge;   // -1.602e-19 C (Cloud.cpp:15)
    static const double epsilon0;         // 8.8541878e-12 F/m (Cloud.cpp:16)
    static double dustParticleMassDensity;
    static double justX, justY, velX, velY;

    // Cloud::initializeGrid (Cloud.cpp:95-135): square grid, Gaussian charge
    // (qMean, qSigma in units of e) and radius (rMean, rSigma).  The draws use
    // std::mt19937_64 with an explicit seed instead of the wall clock.
    static std::unique_ptr<Cloud> initializeGrid(cloud_index numParticles, cloud_index rowX,
                                                 cloud_index rowY, double rMean, double rSigma,
                                                 double qMean, double qSigma, uint64_t seed = 0,
                                                 int device = 0);
    // Hexagonal-lattice patch (demon_hex_lattice), uniform charge and mass.
    static std::unique_ptr<Cloud> initializeHexLattice(cloud_index numParticles, double spacing,
                                                       double jitter, uint64_t seed, double charge,
                                                       double mass, int device = 0);

    // Stage getters of the reference (Cloud.cpp:277-475), element-wise:
    // stage 1 reads x/Vx, stages 2-4 the caches.
    const double *stageX(int stage) const { return stage == 1 ? x.data() : xCache.data(); }
    const double *stageY(int stage) const { return stage == 1 ? y.data() : yCache.data(); }
    const double *stageVx(int stage) const { return stage == 1 ? Vx.data() : VxCache.data(); }
    const double *stageVy(int stage) const { return stage == 1 ? Vy.data() : VyCache.data(); }

    // Device mirror.  The context is created on first use with the forces'
    // parameters; syncToDevice uploads x, y, Vx, Vy, charge, mass.
    demon_ctx *context(const demon_params &p);
    void syncToDevice();
    void syncFromDevice();
    int device() const { return device_; }

private:
    demon_ctx *ctx_ = nullptr;
    demon_params ctxParams_{};
    int device_;
};

class Force {
public:
    Cloud *const cloud;
    explicit Force(Cloud *C) : cloud(C) {}
    virtual ~Force() {}
    virtual void force1(double currentTime) = 0;
    virtual void force2(double currentTime) = 0;
    virtual void force3(double currentTime) = 0;
    virtual void force4(double currentTime) = 0;
    // The device parameters this force contributes and its ForceFlag bit;
    // Runge_Kutta4 builds one parameter block from all forces.
    virtual force_flags flag() const = 0;
    virtual void contribute(demon_params &p) const = 0;
    // Force::writeForce / readForce: the FORCES bits and the parameter
    // keywords of the output file's primary header (e.g.
    // ShieldedCoulombForce.cpp:239-277).
    virtual void writeForce(long &flags, std::vector<FitsKey> &keys) const = 0;
    virtual void readForce(const FitsCloud &file) = 0;

protected:
    // Evaluates this force on the cloud's stage-S state at the stage time t on
    // the device and adds it into cloud->forceX / forceY (the contract of
    // Force::forceS).
    void evaluate(int stage, uint64_t step, double t);
};

typedef std::vector<Force *> ForceArray;

class ShieldedCoulombForce : public Force {
public:
    ShieldedCoulombForce(Cloud *C, double shieldingConstant) : Force(C), shielding(shieldingConstant) {}
    void force1(double t) override { evaluate(1, 0, t); }
    void force2(double t) override { evaluate(2, 0, t); }
    void force3(double t) override { evaluate(3, 0, t); }
    void force4(double t) override { evaluate(4, 0, t); }
    force_flags flag() const override { return ShieldedCoulombForceFlag; }
    void contribute(demon_params &p) const override { p.shielding = shielding; }
    void writeForce(long &flags, std::vector<FitsKey> &keys) const override {
        flags |= ShieldedCoulombForceFlag;
        keys.push_back({"shieldingConstant", shielding, "[m^-1] (ShieldedCoulombForce)"});
    }
    void readForce(const FitsCloud &f) override { shielding = f.key("shieldingConstant", shielding); }
    double shielding;  // inverse shielding length [1/m]
};

class ConfinementForce : public Force {
public:
    ConfinementForce(Cloud *C, double confineConst) : Force(C), confine(confineConst) {}
    void force1(double t) override { evaluate(1, 0, t); }
    void force2(double t) override { evaluate(2, 0, t); }
    void force3(double t) override { evaluate(3, 0, t); }
    void force4(double t) override { evaluate(4, 0, t); }
    force_flags flag() const override { return ConfinementForceFlag; }
    void contribute(demon_params &p) const override { p.confine = confine; }
    void writeForce(long &flags, std::vector<FitsKey> &keys) const override {
        flags |= ConfinementForceFlag;
        keys.push_back({"confineConst", confine, "[V/m^2] (ConfinementForce)"});
    }
    void readForce(const FitsCloud &f) override { confine = f.key("confineConst", confine); }
    double confine;  // [V/m^2]
};

class DragForce : public Force {
public:
    DragForce(Cloud *C, double gamma) : Force(C), dragConst(-gamma) {}
    void force1(double t) override { evaluate(1, 0, t); }
    void force2(double t) override { evaluate(2, 0, t); }
    void force3(double t) override { evaluate(3, 0, t); }
    void force4(double t) override { evaluate(4, 0, t); }
    force_flags flag() const override { return DragForceFlag; }
    void contribute(demon_params &p) const override { p.drag_gamma = -dragConst; }
    void writeForce(long &flags, std::vector<FitsKey> &keys) const override {
        flags |= DragForceFlag;
        keys.push_back({"dragConst", dragConst, "[Hz] (DragForce)"});
    }
    void readForce(const FitsCloud &f) override { dragConst = f.key("dragConst", dragConst); }
    double dragConst;  // -gamma [1/s] (DragForce.h:17)
};

class ThermalForce : public Force {
public:
    ThermalForce(Cloud *C, double redFactor, uint64_t seed = 0) : Force(C), heatVal(redFactor), seed(seed) {}
    void force1(double t) override { evaluate(1, step_++, t); }
    void force2(double t) override { evaluate(2, step_, t); }
    void force3(double t) override { evaluate(3, step_, t); }
    void force4(double t) override { evaluate(4, step_, t); }
    force_flags flag() const override { return ThermalForceFlag; }
    void contribute(demon_params &p) const override {
        p.heat_val = heatVal;
        p.seed = seed;
    }
    void writeForce(long &flags, std::vector<FitsKey> &keys) const override {
        flags |= ThermalForceFlag;
        keys.push_back({"heatingValue", heatVal, "[N] (ThermalForce)"});
    }
    void readForce(const FitsCloud &f) override { heatVal = f.key("heatingValue", heatVal); }
    double heatVal;  // [N]
    uint64_t seed;

private:
    uint64_t step_ = 0;
};

// MagneticForce (MagneticForce.h:14-33): F += q B (vy, -vx), B along z.
class MagneticForce : public Force {
public:
    MagneticForce(Cloud *C, double magneticField) : Force(C), BField(magneticField) {}
    void force1(double t) override { evaluate(1, 0, t); }
    void force2(double t) override { evaluate(2, 0, t); }
    void force3(double t) override { evaluate(3, 0, t); }
    void force4(double t) override { evaluate(4, 0, t); }
    force_flags flag() const override { return MagneticForceFlag; }
    void contribute(demon_params &p) const override { p.magnetic_field = BField; }
    void writeForce(long &flags, std::vector<FitsKey> &keys) const override {
        flags |= MagneticForceFlag;
        keys.push_back({"magneticField", BField, "[T] (MagneticForce)"});
    }
    void readForce(const FitsCloud &f) override { BField = f.key("magneticField", BField); }
    double BField;  // [T]
};

// TimeVaryingDragForce (TimeVaryingDragForce.h:14-33): the drag constant at
// the stage time t is -(scale t + offset).  The reference's driver gives it
// offset = dragGamma and then leaves DragForce out (driver.cpp:322-323,394).
class TimeVaryingDragForce : public Force {
public:
    TimeVaryingDragForce(Cloud *C, double scale, double offset)
        : Force(C), scaleConst(scale), offsetConst(offset) {}
    void force1(double t) override { evaluate(1, 0, t); }
    void force2(double t) override { evaluate(2, 0, t); }
    void force3(double t) override { evaluate(3, 0, t); }
    void force4(double t) override { evaluate(4, 0, t); }
    force_flags flag() const override { return TimeVaryingDragForceFlag; }
    void contribute(demon_params &p) const override {
        p.drag_scale = scaleConst;
        p.drag_gamma = offsetConst;
    }
    // as the reference's: DragForce's keyword (its dragConst is +offset
    // between stages, TimeVaryingDragForce.h:16-17) and bit, then its own
    // (TimeVaryingDragForce.cpp:41-64)
    void writeForce(long &flags, std::vector<FitsKey> &keys) const override {
        flags |= DragForceFlag | TimeVaryingDragForceFlag;
        keys.push_back({"dragConst", offsetConst, "[Hz] (DragForce)"});
        keys.push_back({"TVDragScaleConst", scaleConst, "[Hz/s] (TimeVaryingDragForce)"});
        keys.push_back({"TVDragOffsetConst", offsetConst, "[Hz] (TimeVaryingDragForce)"});
    }
    void readForce(const FitsCloud &f) override {
        scaleConst = f.key("TVDragScaleConst", scaleConst);
        offsetConst = f.key("TVDragOffsetConst", offsetConst);
    }
    double scaleConst;   // [Hz/s]
    double offsetConst;  // [Hz]
};

#define DEMON_HOST_STAGES(STEP_FIRST, STEP)                         \
    void force1(double t) override { evaluate(1, STEP_FIRST, t); }  \
    void force2(double t) override { evaluate(2, STEP, t); }        \
    void force3(double t) override { evaluate(3, STEP, t); }        \
    void force4(double t) override { evaluate(4, STEP, t); }

// RectConfinementForce (RectConfinementForce.h): F += (q cX x, q cY y).
class RectConfinementForce : public Force {
public:
    RectConfinementForce(Cloud *C, double confineConstX, double confineConstY)
        : Force(C), confineX(confineConstX), confineY(confineConstY) {}
    DEMON_HOST_STAGES(0, 0)
    force_flags flag() const override { return RectConfinementForceFlag; }
    void contribute(demon_params &p) const override {
        p.confine_x = confineX;
        p.confine_y = confineY;
    }
    void writeForce(long &flags, std::vector<FitsKey> &keys) const override {
        flags |= RectConfinementForceFlag;
        keys.push_back({"confineConstX", confineX, "[V/m^2] (RectConfinementForce)"});
        keys.push_back({"confineConstY", confineY, "[V/m^2] (RectConfinementForce)"});
    }
    void readForce(const FitsCloud &f) override {
        confineX = f.key("confineConstX", confineX);
        confineY = f.key("confineConstY", confineY);
    }
    double confineX, confineY;  // [V/m^2]
};

// ConfinementForceVoid (ConfinementForceVoid.h:14-20): the confinement force
// plus -q d e^{-d r} r_hat, one plugin; it writes ConfinementForce's keyword
// and bit too (ConfinementForceVoid.cpp:57-80).
class ConfinementForceVoid : public ConfinementForce {
public:
    ConfinementForceVoid(Cloud *C, double confineConst, double voidDecay)
        : ConfinementForce(C, confineConst), decay(voidDecay) {}
    force_flags flag() const override { return ConfinementForceVoidFlag; }
    void contribute(demon_params &p) const override {
        p.confine = confine;
        p.void_decay = decay;
    }
    void writeForce(long &flags, std::vector<FitsKey> &keys) const override {
        ConfinementForce::writeForce(flags, keys);
        flags |= ConfinementForceVoidFlag;
        keys.push_back({"decay", decay, "[m^-1] (ConfinementForceVoid)"});
    }
    void readForce(const FitsCloud &f) override {
        ConfinementForce::readForce(f);
        decay = f.key("decay", decay);
    }
    double decay;  // [1/m]
};

// RotationalForce (RotationalForce.h:20-21): c (-y, x)/r for rmin < r < rmax.
class RotationalForce : public Force {
public:
    RotationalForce(Cloud *C, double rmin, double rmax, double rotConst)
        : Force(C), innerRad(rmin), outerRad(rmax), rotationalConst(rotConst) {}
    DEMON_HOST_STAGES(0, 0)
    force_flags flag() const override { return RotationalForceFlag; }
    void contribute(demon_params &p) const override {
        p.rot_inner = innerRad;
        p.rot_outer = outerRad;
        p.rot_const = rotationalConst;
    }
    void writeForce(long &flags, std::vector<FitsKey> &keys) const override {
        flags |= RotationalForceFlag;
        keys.push_back({"rotationalConst", rotationalConst, "[N] (RotationalForce)"});
        keys.push_back({"innerRadius", innerRad, "[m] (RotationalForce)"});
        keys.push_back({"outerRadius", outerRad, "[m] (RotationalForce)"});
    }
    void readForce(const FitsCloud &f) override {
        rotationalConst = f.key("rotationalConst", rotationalConst);
        innerRad = f.key("innerRadius", innerRad);
        outerRad = f.key("outerRadius", outerRad);
    }
    double innerRad, outerRad;  // [m]
    double rotationalConst;     // [N]
};

// DrivingForce (DrivingForce.h:17-18): the member driveConst is -dampConst,
// and the FITS keyword holds the member.
class DrivingForce : public Force {
public:
    DrivingForce(Cloud *C, double dampConst, double amp, double drivingShift)
        : Force(C), amplitude(amp), driveConst(-dampConst), shift(drivingShift) {}
    DEMON_HOST_STAGES(0, 0)
    force_flags flag() const override { return DrivingForceFlag; }
    void contribute(demon_params &p) const override {
        p.drive_const = -driveConst;
        p.drive_amplitude = amplitude;
        p.drive_shift = shift;
    }
    void writeForce(long &flags, std::vector<FitsKey> &keys) const override {
        flags |= DrivingForceFlag;
        keys.push_back({"drivingAmplitude", amplitude, "[N] (DrivingForce)"});
        keys.push_back({"drivingConst", driveConst, "[m^2] (DrivingForce)"});
        keys.push_back({"drivingShift", shift, "[m] (DrivingForce)"});
    }
    void readForce(const FitsCloud &f) override {
        amplitude = f.key("drivingAmplitude", amplitude);
        driveConst = f.key("drivingConst", driveConst);
        shift = f.key("drivingShift", shift);
    }
    double amplitude;   // [N]
    double driveConst;  // [m^2]
    double shift;       // [m]
};

// ElectricForce (ElectricForce.h:16-17): F += q E e^{-r/R} r.
class ElectricForce : public Force {
public:
    ElectricForce(Cloud *C, double electricField, double plasmaRad)
        : Force(C), electric(electricField), radius(plasmaRad) {}
    DEMON_HOST_STAGES(0, 0)
    force_flags flag() const override { return ElectricForceFlag; }
    void contribute(demon_params &p) const override {
        p.electric_field = electric;
        p.plasma_radius = radius;
    }
    void writeForce(long &flags, std::vector<FitsKey> &keys) const override {
        flags |= ElectricForceFlag;
        keys.push_back({"ElectricFieldStrength", electric, "[V/m^2] (ElectricForce)"});
        keys.push_back({"plasmaRadius", radius, "[m] (ElectricForce)"});
    }
    void readForce(const FitsCloud &f) override {
        electric = f.key("electricFieldStrength", electric);
        radius = f.key("plasmaRadius", radius);
    }
    double electric;  // [V/m^2]
    double radius;    // [m]
};

// VertElectricForce (VertElectricForce.h:16-17): Fy += q E e^{y/d}.
class VertElectricForce : public Force {
public:
    VertElectricForce(Cloud *C, double vertElectricField, double vertDecay)
        : Force(C), vertElectric(vertElectricField), vertDec(vertDecay) {}
    DEMON_HOST_STAGES(0, 0)
    force_flags flag() const override { return VertElectricForceFlag; }
    void contribute(demon_params &p) const override {
        p.vert_electric_field = vertElectric;
        p.vert_decay = vertDec;
    }
    void writeForce(long &flags, std::vector<FitsKey> &keys) const override {
        flags |= VertElectricForceFlag;
        keys.push_back({"VertElectricFieldStrength", vertElectric, "[V/m^2] (VertElectricForce)"});
        keys.push_back({"verticalDecay", vertDec, "[m] (VertElectricForce)"});
    }
    void readForce(const FitsCloud &f) override {
        vertElectric = f.key("vertElectricFieldStrength", vertElectric);
        vertDec = f.key("verticalDecay", vertDec);
    }
    double vertElectric;  // [V/m^2]
    double vertDec;       // [m]
};

// GravitationalForce: Fy += m g.
class GravitationalForce : public Force {
public:
    GravitationalForce(Cloud *C, double g) : Force(C), gravitational(g) {}
    DEMON_HOST_STAGES(0, 0)
    force_flags flag() const override { return GravitationalForceFlag; }
    void contribute(demon_params &p) const override { p.gravity = gravitational; }
    void writeForce(long &flags, std::vector<FitsKey> &keys) const override {
        flags |= GravitationalForceFlag;
        keys.push_back({"GravitationalFieldStrength", gravitational, "[m/s^2] (GravitationalForce)"});
    }
    void readForce(const FitsCloud &f) override {
        gravitational = f.key("gravitationalFieldStrength", gravitational);
    }
    double gravitational;  // [m/s^2]
};

// ThermalForceLocalized (ThermalForceLocalized.h:16-31): the thermal force
// with heatVal1 strictly inside heatingRadius and heatVal2 elsewhere.
class ThermalForceLocalized : public Force {
public:
    ThermalForceLocalized(Cloud *C, double thermRed1, double thermRed2, double specifiedRadius, uint64_t seed = 0)
        : Force(C), heatingRadius(specifiedRadius), heatVal1(thermRed1), heatVal2(thermRed2), seed(seed) {}
    DEMON_HOST_STAGES(step_++, step_)
    force_flags flag() const override { return ThermalForceLocalizedFlag; }
    void contribute(demon_params &p) const override {
        p.heat_in = heatVal1;
        p.heat_out = heatVal2;
        p.heat_radius = heatingRadius;
        p.seed = seed;
    }
    void writeForce(long &flags, std::vector<FitsKey> &keys) const override {
        flags |= ThermalForceLocalizedFlag;
        keys.push_back({"heatingValue1", heatVal1, "[N] (ThermalForceLocalized)"});
        keys.push_back({"heatingValue2", heatVal2, "[N] (ThermalForceLocalized)"});
        keys.push_back({"heatingRadius", heatingRadius, "[m] (ThermalForceLocalized)"});
    }
    void readForce(const FitsCloud &f) override {
        heatVal1 = f.key("heatingValue1", heatVal1);
        heatVal2 = f.key("heatingValue2", heatVal2);
        heatingRadius = f.key("heatingRadius", heatingRadius);
    }
    double heatingRadius;       // [m]
    double heatVal1, heatVal2;  // [N]
    uint64_t seed;

private:
    uint64_t step_ = 0;
};

// TimeVaryingThermalForce (TimeVaryingThermalForce.h:14-30): the thermal
// force with heatVal = scale t + offset at the stage time; it writes
// ThermalForce's keyword and bit too (TimeVaryingThermalForce.cpp:44-69).
class TimeVaryingThermalForce : public ThermalForce {
public:
    TimeVaryingThermalForce(Cloud *C, double scale, double offset, uint64_t seed = 0)
        : ThermalForce(C, offset, seed), heatValScale(scale), heatValOffset(offset) {}
    force_flags flag() const override { return TimeVaryingThermalForceFlag; }
    void contribute(demon_params &p) const override {
        p.heat_scale = heatValScale;
        p.heat_offset = heatValOffset;
        p.seed = seed;
    }
    void writeForce(long &flags, std::vector<FitsKey> &keys) const override {
        ThermalForce::writeForce(flags, keys);
        flags |= TimeVaryingThermalForceFlag;
        keys.push_back({"heatingValueScale", heatValScale, "[N/s] (TimeVaryingThermalForce)"});
        keys.push_back({"heatingValueOffset", heatValOffset, "[N] (TimeVaryingThermalForce)"});
    }
    void readForce(const FitsCloud &f) override {
        ThermalForce::readForce(f);
        heatValScale = f.key("heatingValueScale", heatValScale);
        heatValOffset = f.key("heatingValueOffset", heatValOffset);
    }
    double heatValScale;   // [N/s]
    double heatValOffset;  // [N]
};

#undef DEMON_HOST_STAGES

class Integrator {
public:
    Integrator(Cloud *C, const ForceArray &FA, double timeStep, double startTime)
        : currentTime(startTime), cloud(C), forces(FA), init_dt(timeStep) {}
    virtual ~Integrator() {}
    double currentTime;
    virtual void moveParticles(double endTime) = 0;

protected:
    Cloud *const cloud;
    const ForceArray &forces;
    const double init_dt;
};

class Runge_Kutta4 : public Integrator {
public:
    Runge_Kutta4(Cloud *C, const ForceArray &FA, double timeStep, double startTime)
        : Integrator(C, FA, timeStep, startTime) {}
    // Runge_Kutta4::moveParticles (Runge_Kutta4.cpp:23-137) with the dt check
    // of Integrator::modifyTimeStep, entirely on the device.
    void moveParticles(double endTime) override;
    // Steps taken by the last moveParticles call.
    int64_t lastSteps() const { return lastSteps_; }

private:
    int64_t lastSteps_ = 0;
};

class Runge_Kutta2 : public Integrator {
public:
    Runge_Kutta2(Cloud *C, const ForceArray &FA, double timeStep, double startTime)
        : Integrator(C, FA, timeStep, startTime) {}
    // Runge_Kutta2::moveParticles (Runge_Kutta2.cpp:23-78): the midpoint rule
    // with the same dt check, entirely on the device.
    void moveParticles(double endTime) override;
    int64_t lastSteps() const { return lastSteps_; }

private:
    int64_t lastSteps_ = 0;
};

// The driver's parameter set (driver.cpp:66-127 defaults) and Params.cfg
// parsing (checkParams, driver.cpp:626-824).
struct Params {
    double voidDecay = 0.4, magneticFieldStrength = 1.0;
    cloud_index numParticles = 4;
    double startTime = 0.0, dataTimeStep = 0.01, simTimeStep = 0.01 / 100.0, spacing = 0.003;
    double endTime = 5.0, confinementConst = 100, confinementConstX = 100, confinementConstY = 1000;
    double shieldingConstant = 2E4, dragGamma = 10.0, thermRed = 1E-14, thermRed1 = 1E-14;
    double thermScale = 1E-14, thermOffset = 0.0, heatRadius = 0.001, driveConst = 0.00001;
    double waveAmplitude = 1E-13, waveShift = 0.007, machSpeed = 0.2, massFactor = 100;
    double qMean = 6000.0, qSigma = 100.0, rMean = 1.45E-6, rSigma = 0.0;
    double rmin = 0.003 * 5.0, rmax = 0.003 * 10.0, rotConst = 1E-15, dragScale = -1.0;
    double electricFieldStrength = 0, plasmaRadius = 1, vertElectricFieldStrength = 0;
    double verticalDecay = 1, gravitationalFieldStrength = 0, justifyX = 0, justifyY = 0;
    double massDensity = 2200, velocityX = 0.0, velocityY = 0.0;
    force_flags usedForces = 0;
};

// Reads whitespace-separated `key value` pairs; unknown keys are ignored and
// a later duplicate wins.  `forceFlags` is a comma-terminated list of letters
// (B D L R S T v V w E F G); a letter after the last comma is dropped, as in
// the reference.  Conflicting flags (e.g. T with L) and a flag given twice
// throw Error (the reference prints the message and exits).
void checkParams(const std::string &path, Params &params);
void checkParamsText(const std::string &text, Params &params);

}  // namespace demon_amd
