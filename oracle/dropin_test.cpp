// dropin_test.cpp -- the reference's own classes with the GPU binding
// (demon_simulation_code_amd/host/refbind/GpuForces.h) swapped in, against
// the reference alone.  TEST INFRASTRUCTURE ONLY; built by `make dropin`
// into _ref/ from the reference sources where they lie.
//
// Prints one line per variant:  <variant> <max |dr|/max|r|> <max |dv|/max|v|> <t>
//   alacarte  reference Runge_Kutta4 + Confinement + Drag, GpuShieldedCoulombForce
//   gpuforces reference Runge_Kutta4 + GpuConfinement + GpuDrag + GpuShieldedCoulomb
//   gpurk4    GpuRungeKutta4 with the three Gpu* forces
//   bd_gpuforces, bd_gpurk4: the same two against the reference with the
//             driver's forceFlags B,D set (Confinement, Magnetic,
//             ShieldedCoulomb, TimeVaryingDrag; driver.cpp:373-394)
//   rk2_gpurk2  GpuRungeKutta2 with the three Gpu* forces against the
//             reference's Runge_Kutta2 (the driver's -I)
//   plug_gpuforces, plug_gpurk4: the deterministic further plugins in the
//             driver's order (ConfinementVoid, Drag, Driving,
//             RectConfinement, Rotational, ShieldedCoulomb, Electric,
//             Gravitational, VertElectric; driver.cpp:372-402) as Gpu* forces
//             under the reference's Runge_Kutta4, and under GpuRungeKutta4
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "Cloud.h"
#include "ConfinementForce.h"
#include "ConfinementForceVoid.h"
#include "DragForce.h"
#include "DrivingForce.h"
#include "ElectricForce.h"
#include "GpuForces.h"
#include "GravitationalForce.h"
#include "MagneticForce.h"
#include "RectConfinementForce.h"
#include "RotationalForce.h"
#include "VertElectricForce.h"
#include "Runge_Kutta2.h"
#include "Runge_Kutta4.h"
#include "ShieldedCoulombForce.h"
#include "TimeVaryingDragForce.h"

double Cloud::interParticleSpacing = 0.003;
double Cloud::dustParticleMassDensity = 2200;
double Cloud::justX = 0.0;
double Cloud::justY = 0.0;
double Cloud::velX = 0.0;
double Cloud::velY = 0.0;

namespace {

const double kQ = 6000 * -1.602e-19, kM = 2.8e-14, kKappa = 2000.0, kGamma = 2.0, kDt = 1e-3;
const double kConf = 5e-11 / (6000 * 1.602e-19);
const double kB = 0.8, kScale = -3.0;
// further plugins, sized for the 1024-particle lattice (radius about 8 mm)
const double kVoid = 300.0, kRectX = kConf, kRectY = 2 * kConf, kRotIn = 2e-3, kRotOut = 6e-3, kRot = 2e-14;
const double kDamp = 1e-5, kAmp = 2e-14, kShift = 7e-3, kE = -0.5 * kConf, kPlasma = 5e-3, kVertE = 20.0;
const double kVertDecay = 8e-3, kG = -0.5;

struct State {
    std::vector<double> x, y, vx, vy;
    double t;
};

Cloud *make(int n, const std::vector<double> &x, const std::vector<double> &y) {
    Cloud *c = new Cloud(n);
    for (int i = 0; i < n; ++i) {
        c->x[i] = x[i];
        c->y[i] = y[i];
        c->Vx[i] = c->Vy[i] = 0.0;
        c->charge[i] = kQ * (1.0 + 0.05 * std::sin(0.7 * i));
        c->mass[i] = kM;
        c->forceX[i] = c->forceY[i] = 0.0;
    }
    return c;
}

State run(int variant, int n, int steps, const std::vector<double> &x0, const std::vector<double> &y0) {
    Cloud *c = make(n, x0, y0);
    ForceArray fa;
    if (variant >= 9) {
        // the further plugins: reference classes (9) or Gpu* forces (10, 11)
        if (variant == 9) {
            fa.push_back(new ConfinementForceVoid(c, kConf, kVoid));
            fa.push_back(new DragForce(c, kGamma));
            fa.push_back(new DrivingForce(c, kDamp, kAmp, kShift));
            fa.push_back(new RectConfinementForce(c, kRectX, kRectY));
            fa.push_back(new RotationalForce(c, kRotIn, kRotOut, kRot));
            fa.push_back(new ShieldedCoulombForce(c, kKappa));
            fa.push_back(new ElectricForce(c, kE, kPlasma));
            fa.push_back(new GravitationalForce(c, kG));
            fa.push_back(new VertElectricForce(c, kVertE, kVertDecay));
        } else {
            fa.push_back(new GpuConfinementForceVoid(c, kConf, kVoid));
            fa.push_back(new GpuDragForce(c, kGamma));
            fa.push_back(new GpuDrivingForce(c, kDamp, kAmp, kShift));
            fa.push_back(new GpuRectConfinementForce(c, kRectX, kRectY));
            fa.push_back(new GpuRotationalForce(c, kRotIn, kRotOut, kRot));
            fa.push_back(new GpuShieldedCoulombForce(c, kKappa));
            fa.push_back(new GpuElectricForce(c, kE, kPlasma));
            fa.push_back(new GpuGravitationalForce(c, kG));
            fa.push_back(new GpuVertElectricForce(c, kVertE, kVertDecay));
        }
    } else if (variant >= 7) {
        // Runge_Kutta2: reference forces (7) or Gpu* forces (8)
        if (variant == 7) {
            fa.push_back(new ConfinementForce(c, kConf));
            fa.push_back(new DragForce(c, kGamma));
            fa.push_back(new ShieldedCoulombForce(c, kKappa));
        } else {
            fa.push_back(new GpuConfinementForce(c, kConf));
            fa.push_back(new GpuDragForce(c, kGamma));
            fa.push_back(new GpuShieldedCoulombForce(c, kKappa));
        }
    } else if (variant >= 4) {
        // forceFlags B,D: reference classes (4), Gpu* forces (5, 6)
        if (variant == 4) {
            fa.push_back(new ConfinementForce(c, kConf));
            fa.push_back(new MagneticForce(c, kB));
            fa.push_back(new ShieldedCoulombForce(c, kKappa));
            fa.push_back(new TimeVaryingDragForce(c, kScale, kGamma));
        } else {
            fa.push_back(new GpuConfinementForce(c, kConf));
            fa.push_back(new GpuMagneticForce(c, kB));
            fa.push_back(new GpuShieldedCoulombForce(c, kKappa));
            fa.push_back(new GpuTimeVaryingDragForce(c, kScale, kGamma));
        }
    } else if (variant <= 1) {
        fa.push_back(new ConfinementForce(c, kConf));
        fa.push_back(new DragForce(c, kGamma));
        if (variant == 0) fa.push_back(new ShieldedCoulombForce(c, kKappa));
        else fa.push_back(new GpuShieldedCoulombForce(c, kKappa));
    } else {
        fa.push_back(new GpuConfinementForce(c, kConf));
        fa.push_back(new GpuDragForce(c, kGamma));
        fa.push_back(new GpuShieldedCoulombForce(c, kKappa));
    }
    Integrator *I;
    if (variant == 3 || variant == 6 || variant == 11) I = new GpuRungeKutta4(c, fa, kDt, 0.0);
    else if (variant == 7) I = new Runge_Kutta2(c, fa, kDt, 0.0);
    else if (variant == 8) I = new GpuRungeKutta2(c, fa, kDt, 0.0);
    else I = new Runge_Kutta4(c, fa, kDt, 0.0);
    for (int s = 0; s < steps; ++s) I->moveParticles(std::nextafter(I->currentTime, INFINITY));
    State st{std::vector<double>(c->x, c->x + n), std::vector<double>(c->y, c->y + n),
             std::vector<double>(c->Vx, c->Vx + n), std::vector<double>(c->Vy, c->Vy + n), I->currentTime};
    delete I;
    for (Force *f : fa) delete f;
    delete c;
    return st;
}

}  // namespace

int main(int argc, char **argv) {
    const int n = argc > 1 ? std::atoi(argv[1]) : 1024;
    const int steps = argc > 2 ? std::atoi(argv[2]) : 20;
    std::vector<double> x(n), y(n);
    if (demon_hex_lattice(n, 500e-6, 50e-6, 21, x.data(), y.data()) != DEMON_OK) return 2;
    const char *names[12] = {"reference", "alacarte", "gpuforces", "gpurk4", "bd_reference", "bd_gpuforces",
                             "bd_gpurk4", "rk2_reference", "rk2_gpurk2", "plug_reference", "plug_gpuforces",
                             "plug_gpurk4"};
    for (int base : {0, 4, 7, 9}) {
        const State ref = run(base, n, steps, x, y);
        double rmax = 0, vmax = 0;
        for (int i = 0; i < n; ++i) {
            rmax = std::fmax(rmax, std::hypot(ref.x[i], ref.y[i]));
            vmax = std::fmax(vmax, std::hypot(ref.vx[i], ref.vy[i]));
        }
        const int last = base == 0 ? 3 : base == 4 ? 6 : base == 7 ? 8 : 11;
        for (int v = base + 1; v <= last; ++v) {
            const State s = run(v, n, steps, x, y);
            double dr = 0, dv = 0;
            for (int i = 0; i < n; ++i) {
                dr = std::fmax(dr, std::hypot(s.x[i] - ref.x[i], s.y[i] - ref.y[i]));
                dv = std::fmax(dv, std::hypot(s.vx[i] - ref.vx[i], s.vy[i] - ref.vy[i]));
            }
            std::printf("%s %.3e %.3e %.17g %.17g\n", names[v], dr / rmax, dv / vmax, s.t, ref.t);
        }
    }
    return 0;
}
