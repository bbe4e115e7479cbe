// fft_analysis.cpp -- the `demon_fft` command: the reference's FFTAnalysis
// (FFTAnalysis.cpp:54-130, 213-283) on a DEMON output file, with the
// transform on the device (demon_trajectory_spectrum).
//
//   demon_fft -a N -i data.fits -o output.dat [--device D] [-h]
//
// -a N averages the transforms of the first N particles' X_POSITION series
// (fewer if the file has fewer); without -a nothing is computed, as in the
// reference.  output.dat gets one line per frequency bin i < rows/2:
// "i / (rows * 0.01)<TAB>amplitude", the reference's labels (which assume
// dataTimeStep = 0.01 s) and its stream formatting.
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "demon_host.h"

using namespace demon_amd;

int main(int argc, char **argv) {
    std::string in = "data.fits", out = "output.dat";
    long nuse = 1;  // FFTAnalysis.cpp:43
    bool average = false;
    int device = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        const bool has = i + 1 < argc;
        if (a == "-h") {
            std::cout << "\n  demon_fft\n\n -a 1          average FFT of the first N particles' trajectories\n"
                         " -i data.fits  input file\n -o output.dat output file\n --device D\n\n";
            return 0;
        } else if (a == "-a") {
            average = true;
            if (has && std::strspn(argv[i + 1], "0123456789") == std::strlen(argv[i + 1])) nuse = std::atol(argv[++i]);
            else std::cout << "Warning: -a option incomplete. Using default average number (" << nuse << ")." << std::endl;
        } else if (a == "-i" && has) {
            in = argv[++i];
        } else if (a == "-o" && has) {
            out = argv[++i];
        } else if (a == "--device" && has) {
            device = std::atoi(argv[++i]);
        }
    }
    if (!average) return 0;
    try {
        const FitsSeries s = readFitsSeries(in, 0, nuse);
        std::vector<double> amp((size_t)(s.rows / 2));
        check(demon_trajectory_spectrum(s.data.data(), s.rows, s.nuse, s.nuse, amp.data(), device), nullptr);
        std::ofstream f(out);
        if (!f) throw Error(DEMON_ERR_ARG, "cannot open " + out);
        for (int64_t i = 0; i < s.rows / 2; ++i)
            f << ((double)i) / ((double)s.rows * 0.01) << "\t" << amp[(size_t)i] << "\n";
    } catch (const Error &e) {
        std::cerr << "Error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
