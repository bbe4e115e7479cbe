// demon_host.cpp -- implementation of the host mirror (see demon_host.h).
#include "demon_host.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>

namespace demon_amd {

const double Cloud::electronCharge = -1.602E-19;
const double Cloud::epsilon0 = 8.8541878E-12;
double Cloud::interParticleSpacing = 0.003;
double Cloud::dustParticleMassDensity = 2200;
double Cloud::justX = 0.0;
double Cloud::justY = 0.0;
double Cloud::velX = 0.0;
double Cloud::velY = 0.0;

void check(int rc, const demon_ctx *ctx) {
    if (rc != DEMON_OK) throw Error(rc, demon_last_error(ctx));
}

Cloud::Cloud(cloud_index numPar, int device)
    : n(numPar), x(n), y(n), Vx(n), Vy(n), charge(n), mass(n, 1.0), forceX(n, 0.0), forceY(n, 0.0),
      xCache(n), yCache(n), VxCache(n), VyCache(n), device_(device) {
    if (n < 2) throw Error(DEMON_ERR_ARG, "a cloud needs at least 2 particles");
}

Cloud::~Cloud() { demon_destroy(ctx_); }

std::unique_ptr<Cloud> Cloud::initializeGrid(cloud_index numParticles, cloud_index rowX,
                                             cloud_index rowY, double rMean, double rSigma,
                                             double qMean, double qSigma, uint64_t seed, int device) {
    std::unique_ptr<Cloud> c(new Cloud(numParticles, device));
    // as square a grid as possible when only numParticles is given (Cloud.cpp:103-112)
    if (rowY == 0) {
        rowX = (cloud_index)std::floor(std::sqrt((double)numParticles));
        while (numParticles % rowX != 0) ++rowX;
        rowY = numParticles / rowX;
    }
    const double hx = ((double)rowX - 1.0) / 2.0 * interParticleSpacing;
    const double hy = ((double)rowY - 1.0) / 2.0 * interParticleSpacing;
    std::mt19937_64 eng(seed);
    std::normal_distribution<double> qd(qMean, qSigma > 0 ? qSigma : 1e-300);
    for (cloud_index i = 0; i < c->n; ++i)
        c->charge[i] = (qSigma > 0 ? qd(eng) : qMean) * electronCharge;        // Cloud.cpp:61-65
    const double pm = (4.0 / 3.0) * kPi * dustParticleMassDensity;             // Cloud.cpp:74-81
    std::normal_distribution<double> rd(rMean, rSigma > 0 ? rSigma : 1e-300);
    for (cloud_index i = 0; i < c->n; ++i) {
        const double r = rSigma > 0 ? rd(eng) : rMean;
        c->mass[i] = pm * r * r * r;
    }
    for (cloud_index i = 0; i < rowX; ++i)
        for (cloud_index j = 0; j < rowY; ++j) {
            c->x[i * rowY + j] = hx - ((double)i * interParticleSpacing) + justX;
            c->y[i * rowY + j] = hy - ((double)j * interParticleSpacing) + justY;
        }
    for (cloud_index l = 0; l < c->n; ++l) {
        c->Vx[l] = velX;
        c->Vy[l] = velY;
    }
    return c;
}

std::unique_ptr<Cloud> Cloud::initializeHexLattice(cloud_index numParticles, double spacing,
                                                   double jitter, uint64_t seed, double charge,
                                                   double mass, int device) {
    std::unique_ptr<Cloud> c(new Cloud(numParticles, device));
    check(demon_hex_lattice(numParticles, spacing, jitter, seed, c->x.data(), c->y.data()), nullptr);
    for (cloud_index i = 0; i < c->n; ++i) {
        c->charge[i] = charge;
        c->mass[i] = mass;
        c->Vx[i] = 0.0;
        c->Vy[i] = 0.0;
    }
    return c;
}

demon_ctx *Cloud::context(const demon_params &p) {
    if (!ctx_) {
        check(demon_create(&ctx_, n, &p), nullptr);
        ctxParams_ = p;
        syncToDevice();
    } else if (std::memcmp(&p, &ctxParams_, sizeof p) != 0) {
        check(demon_set_params(ctx_, &p), ctx_);
        ctxParams_ = p;
    }
    return ctx_;
}

void Cloud::syncToDevice() {
    if (!ctx_) return;
    check(demon_upload(ctx_, x.data(), y.data(), Vx.data(), Vy.data(), charge.data(), mass.data()), ctx_);
}

void Cloud::syncFromDevice() {
    if (!ctx_) return;
    check(demon_download(ctx_, x.data(), y.data(), Vx.data(), Vy.data()), ctx_);
}

static demon_params base_params(int device) {
    demon_params p{};
    p.shielding = 1.0;
    p.dt_check = 1;
    p.device = device;
    return p;
}

void Force::evaluate(int stage, uint64_t step, double t) {
    demon_params p = base_params(cloud->device());
    p.forces = (int32_t)flag();
    contribute(p);
    demon_ctx *ctx = cloud->context(p);
    // charge and mass come from the cloud's host arrays
    check(demon_upload(ctx, cloud->x.data(), cloud->y.data(), cloud->Vx.data(), cloud->Vy.data(),
                       cloud->charge.data(), cloud->mass.data()), ctx);
    check(demon_force_stage_at(ctx, (int32_t)flag(), stage, step, t, cloud->stageX(stage),
                               cloud->stageY(stage), cloud->stageVx(stage), cloud->stageVy(stage),
                               cloud->forceX.data(), cloud->forceY.data()),
          ctx);
}

namespace {

// The device parameter block from the integrator's forces, then the
// moveParticles loop of the given order (4: Runge_Kutta4, 2: Runge_Kutta2).
int64_t move_on_device(Cloud *cloud, const ForceArray &forces, double init_dt, double &currentTime,
                       double endTime, int order) {
    demon_params p = base_params(cloud->device());
    for (const Force *f : forces) {
        if (p.forces & f->flag())
            throw Error(DEMON_ERR_UNSUPPORTED, "moveParticles: two forces of the same kind in one ForceArray");
        p.forces |= (int32_t)f->flag();
        f->contribute(p);
    }
    if ((p.forces & DragForceFlag) && (p.forces & TimeVaryingDragForceFlag))
        throw Error(DEMON_ERR_UNSUPPORTED,
                    "moveParticles: DragForce and TimeVaryingDragForce together (the reference's driver "
                    "uses one or the other)");
    if (!(p.forces & ShieldedCoulombForceFlag))
        throw Error(DEMON_ERR_UNSUPPORTED, "moveParticles: the device integrator needs ShieldedCoulombForce");
    demon_ctx *ctx = cloud->context(p);
    cloud->syncToDevice();
    int64_t steps = 0;
    if (order == 4) check(demon_rk4_move(ctx, &currentTime, endTime, init_dt, &steps), ctx);
    else check(demon_rk2_move(ctx, &currentTime, endTime, init_dt, &steps), ctx);
    cloud->syncFromDevice();
    return steps;
}

}  // namespace

void Runge_Kutta4::moveParticles(double endTime) {
    lastSteps_ = move_on_device(cloud, forces, init_dt, currentTime, endTime, 4);
}

void Runge_Kutta2::moveParticles(double endTime) {
    lastSteps_ = move_on_device(cloud, forces, init_dt, currentTime, endTime, 2);
}

// ---------------------------------------------------------------------------
// Params.cfg
// ---------------------------------------------------------------------------

static void check_force(Params &P, char letter, force_flags flag, std::initializer_list<std::pair<char, force_flags>> conflicts) {
    // checkForce (driver.cpp:483-511)
    if (P.usedForces & flag)
        throw Error(DEMON_ERR_ARG, std::string("option -") + letter + " already set.");
    for (const auto &c : conflicts)
        if (P.usedForces & c.second)
            throw Error(DEMON_ERR_ARG, std::string("option -") + letter + " conflicts with option -" + c.first);
    P.usedForces |= flag;
}

static void apply_flags(Params &P, std::string value) {
    // the reference splits on ',' while more than one character remains
    // (driver.cpp:765-769): the last letter needs a trailing comma.
    std::vector<std::string> flags;
    while (value.size() > 1) {
        const size_t comma = value.find(',');
        flags.push_back(value.substr(0, comma));
        if (comma == std::string::npos) break;
        value = value.substr(comma + 1);
    }
    for (const std::string &f : flags) {
        if (f == "B") check_force(P, 'B', MagneticForceFlag, {});
        if (f == "D") check_force(P, 'D', TimeVaryingDragForceFlag, {});
        if (f == "L") check_force(P, 'L', ThermalForceLocalizedFlag, {{'T', ThermalForceFlag}, {'v', TimeVaryingThermalForceFlag}});
        if (f == "R") check_force(P, 'R', RectConfinementForceFlag, {});
        if (f == "S") check_force(P, 'S', RotationalForceFlag, {});
        if (f == "T") check_force(P, 'T', ThermalForceFlag, {{'L', ThermalForceLocalizedFlag}, {'v', TimeVaryingThermalForceFlag}});
        if (f == "v") check_force(P, 'v', TimeVaryingThermalForceFlag, {{'L', ThermalForceLocalizedFlag}, {'T', ThermalForceFlag}});
        if (f == "V") check_force(P, 'V', ConfinementForceVoidFlag, {});
        if (f == "w") check_force(P, 'w', DrivingForceFlag, {});
        if (f == "E") check_force(P, 'E', ElectricForceFlag, {});
        if (f == "F") check_force(P, 'F', VertElectricForceFlag, {});
        if (f == "G") check_force(P, 'G', GravitationalForceFlag, {});
    }
}

void checkParamsText(const std::string &text, Params &P) {
    std::istringstream in(text);
    std::string k, v;
    while (in >> k >> v) {
        const double d = std::atof(v.c_str());
        if (k == "voidDecay") P.voidDecay = d;
        else if (k == "magneticFieldStrength") P.magneticFieldStrength = d;
        else if (k == "numParticles") P.numParticles = (cloud_index)d;
        else if (k == "startTime") P.startTime = d;
        else if (k == "dataTimeStep") P.dataTimeStep = d;
        else if (k == "simTimeStep") P.simTimeStep = d;
        else if (k == "spacing") { P.spacing = d; Cloud::interParticleSpacing = d; }
        else if (k == "endTime") P.endTime = d;
        else if (k == "confinementConst") P.confinementConst = d;
        else if (k == "confinementConstX") P.confinementConstX = d;
        else if (k == "confinementConstY") P.confinementConstY = d;
        else if (k == "shieldingConstant") P.shieldingConstant = d;
        else if (k == "dragGamma") P.dragGamma = d;
        else if (k == "thermRed") P.thermRed = d;
        else if (k == "thermRed1") P.thermRed1 = d;
        else if (k == "thermScale") P.thermScale = d;
        else if (k == "thermOffset") P.thermOffset = d;
        else if (k == "heatRadius") P.heatRadius = d;
        else if (k == "driveConst") P.driveConst = d;
        else if (k == "waveAmplitude") P.waveAmplitude = d;
        else if (k == "waveShift") P.waveShift = d;
        else if (k == "machSpeed") P.machSpeed = d;
        else if (k == "massFactor") P.massFactor = d;
        else if (k == "qMean") P.qMean = d;
        else if (k == "qSigma") P.qSigma = d;
        else if (k == "rMean") P.rMean = d;
        else if (k == "rSigma") P.rSigma = d;
        else if (k == "rmin") P.rmin = d;
        else if (k == "rmax") P.rmax = d;
        else if (k == "rotConst") P.rotConst = d;
        else if (k == "dragScale") P.dragScale = d;
        else if (k == "electricFieldStrength") P.electricFieldStrength = d;
        else if (k == "plasmaRadius") P.plasmaRadius = d;
        else if (k == "vertElectricFieldStrength") P.vertElectricFieldStrength = d;
        else if (k == "verticalDecay") P.verticalDecay = d;
        else if (k == "gravitationalFieldStrength") P.gravitationalFieldStrength = d;
        else if (k == "justifyX") { P.justifyX = d; Cloud::justX = d; }
        else if (k == "justifyY") { P.justifyY = d; Cloud::justY = d; }
        else if (k == "massDensity") { P.massDensity = d; Cloud::dustParticleMassDensity = d; }
        else if (k == "velocityX") { P.velocityX = d; Cloud::velX = d; }
        else if (k == "velocityY") { P.velocityY = d; Cloud::velY = d; }
        else if (k == "forceFlags") apply_flags(P, v);
    }
}

void checkParams(const std::string &path, Params &P) {
    std::ifstream f(path);
    if (!f) throw Error(DEMON_ERR_ARG, "Incorrect parameter filepath: " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    checkParamsText(ss.str(), P);
}

}  // namespace demon_amd
