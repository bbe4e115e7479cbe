// driver.cpp -- the `demon` command: the reference driver's main loop
// (driver.cpp:315-475) over the device integrator.
//
//   demon [-P Params.cfg] [-n rows [cols]] [-t dt] [-e end] [-o dataTimeStep]
//         [-s shielding] [-C confine] [-g gamma] [-T heat] [-B field]
//         [-D scale offset] [-R X Y] [-V decay] [-S c rmin rmax]
//         [-w amplitude shift driveConst] [-E field radius] [-F field decay]
//         [-G g] [-L radius heat_in heat_out] [-v scale offset] [-I] [-M v factor]
//         [-q mean sigma] [-r mean sigma] [-i spacing] [-p x y] [-k vx vy]
//         [-d density] [-O data.fits | snapshots.csv] [-c file.fits]
//         [-f file.fits] [--hex N] [--seed S] [--device D] [-h]
//
// Options are applied left to right, so options after -P override the file
// (driver.cpp:834-1030).  Always-on forces: ShieldedCoulomb; Confinement
// unless RectConfinement (-R) or ConfinementVoid (-V) replaces it; Drag
// unless TimeVaryingDrag (-D) replaces it (driver.cpp:322-326).  -T, -L and
// -v (the thermal forces) exclude each other (driver.cpp:908-973).  -I
// selects Runge_Kutta2 (driver.cpp:437-438).  -M speed factor fires particle
// 0 into the cloud (the Mach cone experiment, driver.cpp:425-434).
// Output (driver.cpp:356-420, 443-450): the reference's data.fits (or the -O
// name) with a TIME_STEP row every dataTimeStep (demon_fits.h); -c continues
// a file (its forces, parameters, last row and time) and appends to it; -f
// starts from a file's last row and writes a new file.  A name ending in .csv
// writes rows (t, index, x, y, vx, vy) as text instead.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <iostream>
#include <string>

#include "demon_host.h"

using namespace demon_amd;

namespace {

void help() {
    std::cout << "demon: Yukawa all-pairs + RK4 on an MI355X\n"
                 "  -P file        read Params.cfg\n"
                 "  -n R [C]       R*C particles on a grid (rounded up to a multiple of 4)\n"
                 "  --hex N        N particles on a jittered hexagonal lattice instead\n"
                 "  -t dt -e end -o dataTimeStep\n"
                 "  -s kappa -C confine -g gamma -T heat -B field -D scale offset -I (RK2)\n"
                 "  -R X Y (rect) -V decay (void) -S c rmin rmax (shear) -w amp shift driveConst\n"
                 "  -E field radius -F field decay (vertical) -G g -L radius in out -v scale offset\n"
                 "  -M speed massFactor  fire particle 0 into the cloud (Mach cone)\n"
                 "  -q qMean qSigma -r rMean rSigma -i spacing -p x y -k vx vy -d density\n"
                 "  -O data.fits   output file (default data.fits; *.csv writes text rows)\n"
                 "  -c file.fits   continue a run; -f file.fits  start from a run's final state\n"
                 "  --seed S; --device D; -h help\n";
}

bool is_number(const char *s) {
    char *end = nullptr;
    std::strtod(s, &end);
    return end && *end == '\0' && end != s;
}

struct Cli {
    Params P;
    bool rk4 = true;  // driver.cpp:64
    bool mach = false;  // driver.cpp:63
    cloud_index rowX = 4, rowY = 0;
    int64_t hex = 0;
    uint64_t seed = 0;
    int device = 0;
    std::string out = "data.fits";  // driver.cpp:361
    std::string cont, finals;       // -c, -f
};

// checkForce (driver.cpp:483-511): a force option once, and not with the
// options it excludes.
void use_force(Params &P, char letter, force_flags flag, force_flags excludes) {
    if (P.usedForces & flag) throw Error(DEMON_ERR_ARG, std::string("option -") + letter + " already set");
    if (P.usedForces & excludes)
        throw Error(DEMON_ERR_ARG, std::string("option -") + letter + " conflicts with another thermal option");
    P.usedForces |= flag;
}

double take(int argc, char **argv, int &i, double fallback, const char *what) {
    if (i + 1 < argc && is_number(argv[i + 1])) return std::atof(argv[++i]);
    std::cout << "Warning: option incomplete. Using default " << what << " (" << fallback << ")." << std::endl;
    return fallback;
}

// setParticleRows (driver.cpp:274-306): at least 4, a multiple of 4.
void set_rows(Cli &c) {
    cloud_index n = c.rowY == 0 ? c.rowX : c.rowX * c.rowY;
    if (n < 4) {
        c.rowX = 2;
        c.rowY = 2;
        n = 4;
    }
    while (n % 4) {
        if (c.rowY == 0) {
            ++c.rowX;
            n = c.rowX;
        } else if (c.rowY < c.rowX) {
            ++c.rowY;
            n = c.rowX * c.rowY;
        } else {
            ++c.rowX;
            n = c.rowX * c.rowY;
        }
    }
    c.P.numParticles = n;
}

Cli parse(int argc, char **argv) {
    Cli c;
    c.P.numParticles = 4;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        Params &P = c.P;
        if (a == "-h") { help(); std::exit(0); }
        else if (a == "-P") {
            if (i + 1 >= argc) throw Error(DEMON_ERR_ARG, "-P needs a file");
            checkParams(argv[++i], P);
            c.rowX = P.numParticles;
            c.rowY = 0;
        } else if (a == "-n") {
            c.rowX = (cloud_index)take(argc, argv, i, (double)c.rowX, "number of row x particles");
            if (i + 1 < argc && is_number(argv[i + 1])) c.rowY = (cloud_index)std::atof(argv[++i]);
            set_rows(c);
        } else if (a == "--hex") c.hex = (int64_t)take(argc, argv, i, 1024, "hex particles");
        else if (a == "--seed") c.seed = (uint64_t)take(argc, argv, i, 0, "seed");
        else if (a == "--device") c.device = (int)take(argc, argv, i, 0, "device");
        else if (a == "-t") P.simTimeStep = take(argc, argv, i, P.simTimeStep, "time step");
        else if (a == "-e") P.endTime = take(argc, argv, i, P.endTime, "end time");
        else if (a == "-o") P.dataTimeStep = take(argc, argv, i, P.dataTimeStep, "data time step");
        else if (a == "-s") P.shieldingConstant = take(argc, argv, i, P.shieldingConstant, "shielding constant");
        else if (a == "-C") P.confinementConst = take(argc, argv, i, P.confinementConst, "confinementConst");
        else if (a == "-g") P.dragGamma = take(argc, argv, i, P.dragGamma, "dragGamma");
        else if (a == "-T") {
            use_force(P, 'T', ThermalForceFlag, ThermalForceLocalizedFlag | TimeVaryingThermalForceFlag);
            P.thermRed = take(argc, argv, i, P.thermRed, "heat factor");
        } else if (a == "-L") {
            use_force(P, 'L', ThermalForceLocalizedFlag, ThermalForceFlag | TimeVaryingThermalForceFlag);
            P.heatRadius = take(argc, argv, i, P.heatRadius, "radius");
            P.thermRed = take(argc, argv, i, P.thermRed, "heat factor1");
            P.thermRed1 = take(argc, argv, i, P.thermRed1, "heat factor2");
        } else if (a == "-v") {
            use_force(P, 'v', TimeVaryingThermalForceFlag, ThermalForceFlag | ThermalForceLocalizedFlag);
            P.thermScale = take(argc, argv, i, P.thermScale, "heat value scale");
            P.thermOffset = take(argc, argv, i, P.thermOffset, "heat value offset");
        } else if (a == "-B") {
            use_force(P, 'B', MagneticForceFlag, 0);
            P.magneticFieldStrength = take(argc, argv, i, P.magneticFieldStrength, "magnetic field");
        } else if (a == "-D") {
            use_force(P, 'D', TimeVaryingDragForceFlag, 0);
            P.dragScale = take(argc, argv, i, P.dragScale, "scale factor");
            P.dragGamma = take(argc, argv, i, P.dragGamma, "offset");
        } else if (a == "-R") {
            use_force(P, 'R', RectConfinementForceFlag, 0);
            P.confinementConstX = take(argc, argv, i, P.confinementConstX, "confine constantX");
            P.confinementConstY = take(argc, argv, i, P.confinementConstY, "confine constantY");
        } else if (a == "-V") {
            use_force(P, 'V', ConfinementForceVoidFlag, 0);
            P.voidDecay = take(argc, argv, i, P.voidDecay, "void decay");
        } else if (a == "-S") {
            use_force(P, 'S', RotationalForceFlag, 0);
            P.rotConst = take(argc, argv, i, P.rotConst, "force constant");
            P.rmin = take(argc, argv, i, P.rmin, "rmin");
            P.rmax = take(argc, argv, i, P.rmax, "rmax");
        } else if (a == "-w") {
            use_force(P, 'w', DrivingForceFlag, 0);
            P.waveAmplitude = take(argc, argv, i, P.waveAmplitude, "amplitude");
            P.waveShift = take(argc, argv, i, P.waveShift, "wave shift");
            P.driveConst = take(argc, argv, i, P.driveConst, "driving constant");
        } else if (a == "-E") {
            use_force(P, 'E', ElectricForceFlag, 0);
            P.electricFieldStrength = take(argc, argv, i, P.electricFieldStrength, "electric field const");
            P.plasmaRadius = take(argc, argv, i, P.plasmaRadius, "plasma radius");
        } else if (a == "-F") {
            use_force(P, 'F', VertElectricForceFlag, 0);
            P.vertElectricFieldStrength = take(argc, argv, i, P.vertElectricFieldStrength, "vert electric field const");
            P.verticalDecay = take(argc, argv, i, P.verticalDecay, "vertical decay const");
        } else if (a == "-G") {
            use_force(P, 'G', GravitationalForceFlag, 0);
            P.gravitationalFieldStrength = take(argc, argv, i, P.gravitationalFieldStrength, "gravitational field const");
        } else if (a == "-M") {
            c.mach = true;
            P.machSpeed = take(argc, argv, i, P.machSpeed, "velocity");
            P.massFactor = take(argc, argv, i, P.massFactor, "mass");
        } else if (a == "-I") {
            c.rk4 = false;
        } else if (a == "-q") {
            P.qMean = take(argc, argv, i, P.qMean, "mean number of charges");
            P.qSigma = take(argc, argv, i, P.qSigma, "number of charges sigma");
        } else if (a == "-r") {
            P.rMean = take(argc, argv, i, P.rMean, "mean dust radius");
            P.rSigma = take(argc, argv, i, P.rSigma, "dust radius sigma");
        } else if (a == "-i") Cloud::interParticleSpacing = take(argc, argv, i, Cloud::interParticleSpacing, "spacing");
        else if (a == "-p") {
            Cloud::justX = take(argc, argv, i, Cloud::justX, "justify x");
            Cloud::justY = take(argc, argv, i, Cloud::justY, "justify y");
        } else if (a == "-k") {
            Cloud::velX = take(argc, argv, i, Cloud::velX, "velocity x");
            Cloud::velY = take(argc, argv, i, Cloud::velY, "velocity y");
        } else if (a == "-d") Cloud::dustParticleMassDensity = take(argc, argv, i, Cloud::dustParticleMassDensity, "massDensity");
        else if (a == "-O") {
            if (i + 1 >= argc) throw Error(DEMON_ERR_ARG, "-O needs a file");
            c.out = argv[++i];
        } else if (a == "-c") {
            if (i + 1 >= argc) throw Error(DEMON_ERR_ARG, "-c needs a file");
            c.cont = argv[++i];
        } else if (a == "-f") {
            if (i + 1 >= argc) throw Error(DEMON_ERR_ARG, "-f needs a file");
            c.finals = argv[++i];
        } else {
            throw Error(DEMON_ERR_UNSUPPORTED, "Unknown or unsupported option " + a);
        }
    }
    return c;
}

}  // namespace

int main(int argc, char **argv) {
    const auto start = std::chrono::steady_clock::now();
    try {
        Cli c = parse(argc, argv);
        Params &P = c.P;
        if (!c.cont.empty() && !c.finals.empty()) throw Error(DEMON_ERR_ARG, "-c and -f exclude each other");
        std::unique_ptr<Cloud> cloud;
        FitsCloud file;
        const std::string &src = !c.cont.empty() ? c.cont : c.finals;
        if (!src.empty()) {
            // Cloud::initializeFromFile (driver.cpp:332-356): the last TIME_STEP row
            file = readFits(src, -1);
            cloud.reset(new Cloud(file.n, c.device));
            cloud->x = file.x;
            cloud->y = file.y;
            cloud->Vx = file.vx;
            cloud->Vy = file.vy;
            cloud->charge = file.charge;
            cloud->mass = file.mass;
            if (!c.cont.empty()) {
                P.usedForces = file.forces;  // driver.cpp:340
                P.startTime = file.time;     // driver.cpp:343
                if (P.usedForces & ~(force_flags)0x7FFF)
                    throw Error(DEMON_ERR_UNSUPPORTED, src + " has FORCES bits that are no ForceFlag");
                // A plugin that writes its base class's bit too (Void: Confinement,
                // TimeVaryingDrag: Drag, TimeVaryingThermal: Thermal) stands for one
                // force: continue with the forces of the run that wrote the file.
                if (P.usedForces & ConfinementForceVoidFlag) P.usedForces &= ~(force_flags)ConfinementForceFlag;
                if (P.usedForces & TimeVaryingThermalForceFlag) P.usedForces &= ~(force_flags)ThermalForceFlag;
            }
        } else if (c.hex > 0)
            cloud = Cloud::initializeHexLattice(c.hex, Cloud::interParticleSpacing, 0.1 * Cloud::interParticleSpacing,
                                                c.seed, P.qMean * Cloud::electronCharge,
                                                (4.0 / 3.0) * kPi * Cloud::dustParticleMassDensity * P.rMean * P.rMean * P.rMean,
                                                c.device);
        else
            cloud = Cloud::initializeGrid(P.numParticles, c.rowX, c.rowY, P.rMean, P.rSigma, P.qMean,
                                          P.qSigma, c.seed, c.device);
        // driver.cpp:319-326: the always-on forces unless a substitute is used
        if (!(P.usedForces & TimeVaryingDragForceFlag)) P.usedForces |= DragForceFlag;
        else P.usedForces &= ~(force_flags)DragForceFlag;
        if (!(P.usedForces & (RectConfinementForceFlag | ConfinementForceVoidFlag))) P.usedForces |= ConfinementForceFlag;
        P.usedForces |= ShieldedCoulombForceFlag;
        // driver.cpp:372-402: construction order is summation order
        Cloud *cl = cloud.get();
        ConfinementForce conf(cl, P.confinementConst);
        ConfinementForceVoid voidf(cl, P.confinementConst, P.voidDecay);
        DragForce drag(cl, P.dragGamma);
        DrivingForce drive(cl, P.driveConst, P.waveAmplitude, P.waveShift);
        MagneticForce mag(cl, P.magneticFieldStrength);
        RectConfinementForce rect(cl, P.confinementConstX, P.confinementConstY);
        RotationalForce rot(cl, P.rmin, P.rmax, P.rotConst);
        ShieldedCoulombForce yuk(cl, P.shieldingConstant);
        ThermalForce therm(cl, P.thermRed, c.seed);
        ThermalForceLocalized local(cl, P.thermRed, P.thermRed1, P.heatRadius, c.seed);
        TimeVaryingDragForce tvdrag(cl, P.dragScale, P.dragGamma);
        TimeVaryingThermalForce tvtherm(cl, P.thermScale, P.thermOffset, c.seed);
        ElectricForce elec(cl, P.electricFieldStrength, P.plasmaRadius);
        GravitationalForce grav(cl, P.gravitationalFieldStrength);
        VertElectricForce vert(cl, P.vertElectricFieldStrength, P.verticalDecay);
        ForceArray forces;
        for (Force *f : std::initializer_list<Force *>{&conf, &voidf, &drag, &drive, &mag, &rect, &rot, &yuk, &therm,
                                                      &local, &tvdrag, &tvtherm, &elec, &grav, &vert})
            if (P.usedForces & f->flag()) forces.push_back(f);
        // -c: the forces' parameters from the file (driver.cpp:405-408)
        if (!c.cont.empty())
            for (Force *f : forces) f->readForce(file);

        const bool csv = c.cont.empty() && c.out.size() > 4 && c.out.compare(c.out.size() - 4, 4, ".csv") == 0;
        FILE *out = nullptr;
        std::unique_ptr<FitsWriter> fits;
        if (csv) {
            out = std::fopen(c.out.c_str(), "w");
            if (!out) throw Error(DEMON_ERR_ARG, "cannot open " + c.out);
        } else if (!c.cont.empty()) {
            fits.reset(new FitsWriter(c.cont));  // driver.cpp:420: append to TIME_STEP
        } else {
            // writeForce in force order, then Cloud::writeCloudSetup (driver.cpp:410-418)
            long flags = 0;
            std::vector<FitsKey> keys;
            for (const Force *f : forces) f->writeForce(flags, keys);
            fits.reset(new FitsWriter(c.out, flags, keys, cloud->n, cloud->mass.data(), cloud->charge.data(),
                                      cloud->x.data(), cloud->y.data(), cloud->Vx.data(), cloud->Vy.data()));
        }
        auto snapshot = [&](double t) {
            if (fits) {
                fits->writeTimeStep(t, cloud->x.data(), cloud->y.data(), cloud->Vx.data(), cloud->Vy.data());
                return;
            }
            for (cloud_index i = 0; i < cloud->n; ++i)
                std::fprintf(out, "%.17g,%lld,%.17g,%.17g,%.17g,%.17g\n", t, (long long)i, cloud->x[i],
                             cloud->y[i], cloud->Vx[i], cloud->Vy[i]);
            std::fflush(out);
        };
        if (csv) snapshot(P.startTime);
        if (c.mach) {
            // driver.cpp:425-434: after the first row is written, particle 0
            // becomes the bullet, left of the cloud; the distance uses the
            // driver's own spacing variable, 0.003 m, which no option changes
            cloud->x[0] = -0.75 * std::sqrt((double)cloud->n) * 0.003;
            cloud->y[0] = 0.0;
            cloud->Vx[0] = P.machSpeed;
            cloud->Vy[0] = 0.0;
            cloud->mass[0] *= P.massFactor;
        }
        Runge_Kutta4 I4(cloud.get(), forces, P.simTimeStep, P.startTime);
        Runge_Kutta2 I2(cloud.get(), forces, P.simTimeStep, P.startTime);
        Integrator &I = c.rk4 ? static_cast<Integrator &>(I4) : static_cast<Integrator &>(I2);
        double startTime = P.startTime;
        int64_t steps = 0;
        while (startTime < P.endTime) {  // driver.cpp:443-450
            std::cout << "\rCurrent Time: " << I.currentTime << "s (" << I.currentTime / P.endTime * 100.0
                      << "% Complete)" << std::flush;
            I.moveParticles(startTime += P.dataTimeStep);
            steps += c.rk4 ? I4.lastSteps() : I2.lastSteps();
            snapshot(I.currentTime);  // Cloud::writeTimeStep (driver.cpp:447)
        }
        if (out) std::fclose(out);
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        std::cout << "\nN = " << cloud->n << ", " << steps << (c.rk4 ? " RK4" : " RK2") << " steps, t = " << I.currentTime
                  << " s, elapsed " << secs << " s" << std::endl;
    } catch (const Error &e) {
        std::cerr << "Error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
