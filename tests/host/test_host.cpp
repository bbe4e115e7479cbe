// test_host.cpp -- exercises the host C++ mirror (libdemon_host.so) for the
// pytest suite (tests/test_host_mirror.py).
//
//   test_host parse FILE            prints every parsed parameter as key=value
//   test_host parse_text 'k v ...'  same for inline text
//   test_host forces N              hex lattice, ShieldedCoulombForce::force1..4 on the
//                                   device; prints x y q fx1 fy1 fx3 fy3 per particle
//   test_host rk4 N STEPS           Runge_Kutta4 with Confinement, Drag and
//                                   ShieldedCoulomb; prints t, steps, then x y vx vy
//   test_host fits_write PATH N ROWS FLAGS
//                                   writes a DEMON output file (no device): the
//                                   forces of FLAGS' keys, the hex cloud, ROWS
//                                   TIME_STEP rows (row k: x + k 1e-6, t = k 0.01)
//   test_host fits_append PATH ROWS appends ROWS rows to an existing file
//                                   (row k of them: x + 1e-3 k, t = 100 + k)
//   test_host fits_read PATH ROW    prints n rows time forces, the keys, then
//                                   mass charge x y vx vy of row ROW (-1: last)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <string>
#include <vector>

#include "demon_host.h"

using namespace demon_amd;

static void dump(const Params &P) {
#define D(k) std::printf(#k "=%.17g\n", (double)P.k)
    D(voidDecay); D(magneticFieldStrength); D(numParticles); D(startTime); D(dataTimeStep);
    D(simTimeStep); D(spacing); D(endTime); D(confinementConst); D(confinementConstX);
    D(confinementConstY); D(shieldingConstant); D(dragGamma); D(thermRed); D(thermRed1);
    D(thermScale); D(thermOffset); D(heatRadius); D(driveConst); D(waveAmplitude); D(waveShift);
    D(machSpeed); D(massFactor); D(qMean); D(qSigma); D(rMean); D(rSigma); D(rmin); D(rmax);
    D(rotConst); D(dragScale); D(electricFieldStrength); D(plasmaRadius);
    D(vertElectricFieldStrength); D(verticalDecay); D(gravitationalFieldStrength); D(justifyX);
    D(justifyY); D(massDensity); D(velocityX); D(velocityY); D(usedForces);
#undef D
}

int main(int argc, char **argv) {
    if (argc < 3) return 2;
    const std::string mode = argv[1];
    try {
        if (mode == "parse" || mode == "parse_text") {
            Params P;
            if (mode == "parse") checkParams(argv[2], P);
            else checkParamsText(argv[2], P);
            dump(P);
            return 0;
        }
        if (mode == "fits_read") {
            const FitsCloud f = readFits(argv[2], std::atol(argv[3]));
            std::printf("%lld %lld %.17g %ld\n", (long long)f.n, (long long)f.rows, f.time, f.forces);
            std::printf("%zu\n", f.keys.size());
            for (const FitsKey &k : f.keys) std::printf("%s %.17g\n", k.name.c_str(), k.value);
            for (int64_t i = 0; i < f.n; ++i)
                std::printf("%.17g %.17g %.17g %.17g %.17g %.17g\n", f.mass[(size_t)i], f.charge[(size_t)i],
                            f.x[(size_t)i], f.y[(size_t)i], f.vx[(size_t)i], f.vy[(size_t)i]);
            return 0;
        }
        if (mode == "fits_append") {
            FitsWriter w(argv[2]);
            const FitsCloud f = readFits(argv[2], -1);
            const long rows = std::atol(argv[3]);
            std::vector<double> x = f.x;
            for (long k = 1; k <= rows; ++k) {
                for (auto &v : x) v += 1e-3;
                w.writeTimeStep(100.0 + (double)k, x.data(), f.y.data(), f.vx.data(), f.vy.data());
            }
            return 0;
        }
        if (mode == "fits_write") {
            const long n = std::atol(argv[3]);
            const long rows = std::atol(argv[4]);
            const long flags = std::atol(argv[5]);
            auto cloud = Cloud::initializeHexLattice(n, 500e-6, 50e-6, 1, 6000 * Cloud::electronCharge, 2.8e-14);
            ConfinementForce conf(cloud.get(), 52018.31);
            DragForce drag(cloud.get(), 2.0);
            MagneticForce mag(cloud.get(), 0.75);
            ShieldedCoulombForce yuk(cloud.get(), 2000.0);
            ThermalForce therm(cloud.get(), 3e-15);
            TimeVaryingDragForce tvd(cloud.get(), -1.5, 4.0);
            Cloud *c = cloud.get();
            ConfinementForceVoid voidf(c, 52018.31, 0.4);
            DrivingForce drive(c, 1e-5, 1e-13, 0.007);
            RectConfinementForce rect(c, 100.5, 1000.25);
            RotationalForce rot(c, 0.015, 0.03, 1e-15);
            ThermalForceLocalized local(c, 2e-15, 1e-15, 0.001);
            TimeVaryingThermalForce tvtherm(c, 1e-14, 5e-16);
            ElectricForce elec(c, 30.0, 1.5);
            GravitationalForce grav(c, -9.81);
            VertElectricForce vert(c, 12.5, 0.75);
            ForceArray fa;
            // the driver's construction order (driver.cpp:372-402)
            for (Force *f : std::initializer_list<Force *>{&conf, &voidf, &drag, &drive, &mag, &rect, &rot, &yuk,
                                                          &therm, &local, &tvd, &tvtherm, &elec, &grav, &vert}) {
                if (!(flags & f->flag())) continue;
                if (f == &drag && (flags & TimeVaryingDragForceFlag)) continue;
                fa.push_back(f);
            }
            long fl = 0;
            std::vector<FitsKey> keys;
            for (const Force *f : fa) f->writeForce(fl, keys);
            for (long i = 0; i < n; ++i) {
                cloud->Vx[i] = 1e-4 * std::sin(0.3 * (double)i);
                cloud->Vy[i] = -2e-4 * std::cos(0.7 * (double)i);
            }
            FitsWriter w(argv[2], fl, keys, n, cloud->mass.data(), cloud->charge.data(), cloud->x.data(),
                         cloud->y.data(), cloud->Vx.data(), cloud->Vy.data());
            for (long k = 1; k < rows; ++k) {
                std::vector<double> x = cloud->x;
                for (auto &v : x) v += 1e-6 * (double)k;
                w.writeTimeStep(0.01 * (double)k, x.data(), cloud->y.data(), cloud->Vx.data(), cloud->Vy.data());
            }
            return 0;
        }
        const long n = std::atol(argv[2]);
        const double q = 6000 * Cloud::electronCharge, m = 2.8e-14;
        auto cloud = Cloud::initializeHexLattice(n, 500e-6, 50e-6, 1, q, m);
        if (mode == "forces") {
            ShieldedCoulombForce f(cloud.get(), 2000.0);
            f.force1(0.0);
            std::vector<double> f1x = cloud->forceX, f1y = cloud->forceY;
            // stage 3 reads the caches: put a shifted copy there
            for (long i = 0; i < n; ++i) {
                cloud->xCache[i] = cloud->x[i] + 1e-5;
                cloud->yCache[i] = cloud->y[i] - 2e-5;
                cloud->forceX[i] = cloud->forceY[i] = 0.0;
            }
            f.force3(0.0);
            for (long i = 0; i < n; ++i)
                std::printf("%.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", cloud->x[i], cloud->y[i],
                            cloud->charge[i], f1x[i], f1y[i], cloud->forceX[i], cloud->forceY[i]);
            return 0;
        }
        if (mode == "rk4") {
            const long steps = std::atol(argv[3]);
            ConfinementForce conf(cloud.get(), 5e-11 / 9.612e-16);
            DragForce drag(cloud.get(), 2.0);
            ShieldedCoulombForce yuk(cloud.get(), 2000.0);
            ForceArray fa{&conf, &drag, &yuk};
            Runge_Kutta4 I(cloud.get(), fa, 1e-3, 0.0);
            long total = 0;
            // two calls: the second resumes from the downloaded state
            I.moveParticles(1e-3 * (double)(steps / 2) - 1e-9);
            total += I.lastSteps();
            I.moveParticles(1e-3 * (double)steps - 1e-9);
            total += I.lastSteps();
            std::printf("%.17g %ld\n", I.currentTime, total);
            for (long i = 0; i < n; ++i)
                std::printf("%.17g %.17g %.17g %.17g\n", cloud->x[i], cloud->y[i], cloud->Vx[i], cloud->Vy[i]);
            return 0;
        }
    } catch (const Error &e) {
        std::printf("error %d %s\n", e.code, e.what());
        return 1;
    }
    return 2;
}
