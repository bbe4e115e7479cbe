// demon_fits.h -- the reference's FITS output (data.fits) without cfitsio:
// the same HDUs, keywords, columns and row order, written and read by this
// file's own code.
//
//   primary HDU   an empty 16-bit image (driver.cpp:364-367) with the FORCES
//                 bitmask and each force's parameter keyword, as its
//                 writeForce writes them (e.g. ShieldedCoulombForce.cpp:239-266:
//                 fits_write_key_dbl with 6 decimals; keywords longer than 8
//                 characters become HIERARCH cards)
//   CLOUD         binary table, N rows of MASS [kg], CHARGE [C] (D columns)
//                 (Cloud::writeCloudSetup, Cloud.cpp:198-248)
//   TIME_STEP     binary table, one row per snapshot: TIME [s] (D) and
//                 X_POSITION, Y_POSITION [m], X_VELOCITY, Y_VELOCITY [m/s]
//                 (ND vectors); the first row is written at time 0.0 and
//                 every row is flushed (Cloud.cpp:228-271)
//
// A file written here is read by the reference's Cloud::initializeFromFile
// and readForce (cfitsio), and this reader reads the reference's files: the
// -c (continue) and -f (final state) options of the driver use it.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace demon_amd {

// A force parameter keyword of the primary header.
struct FitsKey {
    std::string name;
    double value;
    std::string comment;
};

// A DEMON output file open for appending TIME_STEP rows.
class FitsWriter {
public:
    // Creates path (replacing it): the primary HDU with FORCES = flags and the
    // given keys, the CLOUD table and the TIME_STEP table with its first row
    // at time 0.0 from x, y, vx, vy (driver.cpp:364-400, Cloud.cpp:198-248).
    FitsWriter(const std::string &path, long flags, const std::vector<FitsKey> &keys, int64_t n,
               const double *mass, const double *charge, const double *x, const double *y,
               const double *vx, const double *vy);
    // Opens an existing DEMON file to append rows (the driver's -c).
    explicit FitsWriter(const std::string &path);
    ~FitsWriter();
    FitsWriter(const FitsWriter &) = delete;
    FitsWriter &operator=(const FitsWriter &) = delete;

    // Cloud::writeTimeStep (Cloud.cpp:258-271): appends one row and flushes.
    void writeTimeStep(double t, const double *x, const double *y, const double *vx, const double *vy);
    int64_t rows() const { return rows_; }
    int64_t n() const { return n_; }

private:
    void update_rows();
    std::FILE *f_ = nullptr;
    int64_t n_ = 0, rows_ = 0;
    long naxis2_card_ = 0;  // byte offset of the TIME_STEP NAXIS2 card
    long data_start_ = 0;   // byte offset of the TIME_STEP data
};

// Contents of a DEMON output file (Cloud::initializeFromFile, Cloud.cpp:146-190,
// and the forces' readForce).
struct FitsCloud {
    int64_t n = 0, rows = 0;
    std::vector<double> mass, charge;
    double time = 0.0;                     // TIME of the row read
    std::vector<double> x, y, vx, vy;      // that row's state
    long forces = 0;                       // FORCES of the primary header
    std::vector<FitsKey> keys;             // every other numeric keyword of the primary header
    double key(const std::string &name, double fallback) const;
};

// Reads the CLOUD table and row `row` of TIME_STEP (-1: the last row).
FitsCloud readFits(const std::string &path, int64_t row = -1);

// One TIME_STEP column over every row (FFTAnalysis.cpp:245 reads X_POSITION
// row by row): column 0 X_POSITION, 1 Y_POSITION, 2 X_VELOCITY, 3 Y_VELOCITY;
// the first nuse particles of each row (all of them when nuse < 1 or > n),
// row-major [rows][nuse].
struct FitsSeries {
    int64_t n = 0, rows = 0, nuse = 0;
    std::vector<double> data;
};
FitsSeries readFitsSeries(const std::string &path, int column, int64_t nuse);

}  // namespace demon_amd
