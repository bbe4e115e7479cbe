// test_host.cpp -- exercises the host C++ mirror (libdemon_host.so) for the
// pytest suite (tests/test_host_mirror.py).
//
//   test_host parse FILE            prints every parsed parameter as key=value
//   test_host parse_text 'k v ...'  same for inline text
//   test_host forces N              hex lattice, ShieldedCoulombForce::force1..4 on the
//                                   device; prints x y q fx1 fy1 fx3 fy3 per particle
//   test_host rk4 N STEPS           Runge_Kutta4 with Confinement, Drag and
//                                   ShieldedCoulomb; prints t, steps, then x y vx vy
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

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
