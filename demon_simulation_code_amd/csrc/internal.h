// internal.h -- device data layout and kernel launchers shared by the
// translation units of libdemon_gpu.so.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

namespace demon {

// 1/(4*pi*eps0) with eps0 = 8.8541878e-12, as ShieldedCoulombForce.cpp:21.
constexpr double kCoulomb = 1.0 / (4.0 * 3.14159265358979323846 * 8.8541878E-12);
// Relative widening of the cutoff radius^2 in the squared-distance prefilter.
constexpr double kCutoffSlack = 1.0e-9;
// |r|^2 is stored scaled by (1 - 2^-46) so that the expanded-form
// squared distance |ri|^2 + |rj|^2 - 2 ri.rj errs on the near side by more
// than its rounding error (<= 16 u (|ri|^2 + |rj|^2), u = 2^-53).
constexpr double kNormShrink = 1.0 - 0x1.0p-46;
// Particles are padded to a multiple of this (covers every i tile shape).
constexpr int kPad = 1024;
// Maximum number of j slices of the all-pairs kernel.
constexpr int kMaxSlices = 64;

// One RK stage's state, in the device (tile) order: positions, velocities and
// the prefilter norm nc = -(xs^2 + ys^2) * kNormShrink (padding: NaN positions, nc = -inf).
// box (optional): per 64-particle block the bounding box of its particles'
// positions, (xmin, xmax, ymin, ymax); +inf/-inf for a block of padding only.
struct StageBuf {
    double *xs, *ys, *vxs, *vys, *nc;
    double *box = nullptr;
};

// Particles per bounding-box block, and per superblock of the Newton-3
// all-pairs kernel (four box blocks, one workgroup).
constexpr int kBoxBlock = 64;
constexpr int kSuper = 256;
// Largest nsb/2 of the Newton-3 kernel: the stage kernels hold a
// superblock's tile flags as a bitmask of this many bits in LDS.
constexpr int kMaxHalf = 1024;

// Bounding box of the wave's 64 consecutive particles (one box block), over
// the lanes with real == true; written by lane 0.  Every lane of the wave
// must call it.
__device__ __forceinline__ void wave_box(bool real, double x, double y, double *box) {
    const double inf = __builtin_inf();
    double x0 = real ? x : inf, x1 = real ? x : -inf;
    double y0 = real ? y : inf, y1 = real ? y : -inf;
    for (int off = 32; off > 0; off >>= 1) {
        x0 = fmin(x0, __shfl_xor(x0, off));
        x1 = fmax(x1, __shfl_xor(x1, off));
        y0 = fmin(y0, __shfl_xor(y0, off));
        y1 = fmax(y1, __shfl_xor(y1, off));
    }
    if ((threadIdx.x & 63) == 0) {
        box[0] = x0;
        box[1] = x1;
        box[2] = y0;
        box[3] = y1;
    }
}

// Where the all-pairs kernels leave a stage's Yukawa forces (device order).
// Gather kernel: part[s][2][npad] (s < slices), partial sums per j slice.
// Newton-3 kernel: the i-side sums of superblock X in the work items
// first[X] .. first[X + 1] - 1, part[item][2][256] (item order), and the
// reactions react[d - 1][2][npad] (d = 1 .. nsb/2) from superblock row
// X - d (mod nsb), valid where rflag[X nsb/2 + d - 1] is nonzero.
struct PairSums {
    const double *part;
    int32_t slices;
    const double *react;       // null for the gather kernel
    const uint32_t *rflag;
    int32_t nsb;
    const int32_t *first;      // Newton-3 item ranges per superblock (null: gather)
};

// Newton-3 work plan of one launch, written by the plan kernel: the live
// tiles of every superblock row, cut into items of `span` consecutive live
// tiles of one row, each item one workgroup's share.
// Work queues of the Newton-3 pair kernel: queue 8 k + s holds the items of
// rows (8 k + s) nsb/64 .. (8 k + s + 1) nsb/64 - 1, for XCD k.
constexpr int kN3Queues = 64;
// A row's live tile list holds the offsets d; kN3Swap set: the tile's i side
// is superblock T's blocks (the pair kernel honours it for one-tile items).
constexpr int32_t kN3Swap = 1 << 30;
constexpr int32_t kN3Offset = kN3Swap - 1;

struct N3Plan {
    int32_t span;              // live tiles per item (the last item of a row may hold fewer)
    int32_t nitems;            // items of the launch
    uint32_t next[64];         // item counters of the pair kernel, one per queue (kN3Queues)
};

// Entries of the Newton-3 pair kernels' exp table, e^{-k/64} for k < 642
// (u < 10.03).
constexpr int kExpTabN = 642;

// The table on the host: each entry the double nearest to a long-double
// evaluation.
inline void exp_table_host(double *tab) {
    for (int k = 0; k < kExpTabN; ++k) tab[k] = (double)std::exp(-(long double)k / 64.0L);
}

struct N3List {
    int32_t *list;             // [nsb][nsb/2 + 1] live offsets d of each row, increasing (0: the row itself)
    int32_t *cnt;              // [nsb] live tiles per row
    int32_t *first;            // [nsb + 1] first item of each row (first[nsb] = nitems)
    N3Plan *plan;
    int32_t target;            // items wanted per launch: span = ceil(live tiles / target)
    int32_t max_items;         // >= target + nsb
    int32_t cus;               // compute units (the persistent pair kernel fills each)
    const double *exptab;      // e^{-k/64}, k < 642: the pair kernel's exp table (copied to LDS)
};

// Device scalars of the stepping loop.
constexpr int kCountSlots = 64;

struct Scalars {
    double t;                 // integrator currentTime
    double t_end;             // moveParticles(endTime)
    double init_dt;           // Integrator::init_dt
    double dt_step;           // dt of the current step (after the dt check)
    unsigned long long step;        // RNG step counter
    unsigned long long steps_done;  // steps taken since upload
    unsigned long long remaining;   // steps still allowed in this call
    unsigned long long dt_reductions;
    unsigned long long rmax2_bits;  // max |r|^2 of the stage-1 positions (double bits)
    unsigned int dtmin_bits;        // min FP32 squared distance of tested pairs (float bits)
    int active;                     // 1 while the current step runs
    int err;                        // sticky device-side error code
    int pad_;
    // ordered pairs inside the cutoff [0] and given a distance test [1],
    // summed over launches; a workgroup adds to slot (row + slice) mod
    // kCountSlots, so that the atomics of many workgroups do not serialise
    // on one address
    unsigned long long pair_counts[kCountSlots][2];
    // off-diagonal superblock tiles the Newton-3 kernel evaluated (each
    // writes 256 reactions) and its work items (each writes 256 i-side
    // sums), summed over launches
    unsigned long long n3_tiles, n3_items;
    // the adjoint pair kernel's counts, as pair_counts (a fit call starts
    // them at zero)
    unsigned long long adj_counts[kCountSlots][2];
};

struct PairLaunch {
    const double *xs, *ys, *nc, *q;
    const float *qf;         // q rounded to FP32 (Newton-3 kernel, FP32 pair arithmetic)
    const double *box;       // per-block bounding boxes of xs/ys, or null
    double *react;           // Newton-3 kernel: reactions [nsb/2][2][npad]
    uint32_t *rflag;         // Newton-3 kernel: tile flags [nsb][nsb/2]
    N3List lst;              // Newton-3 kernel: tile lists and work plan
    const int32_t *orig;
    double *part;            // [slices][2][npad] partial force sums
    Scalars *sc;
    int64_t n, npad;
    int32_t jslice;          // j extent per slice (multiple of the chunk)
    int32_t slices;
    int32_t i_per_lane;      // 1, 2 or 4
    int32_t dtcheck;         // fold the FP32 dt check into the near path
    int32_t gated;           // skip when sc->active == 0
    int32_t fp32pair;        // FP32 pair arithmetic
    double kappa, rc2p;      // shielding and (10/kappa)^2 (1 + kCutoffSlack)
};

hipError_t launch_pairs(const PairLaunch &a, hipStream_t s);
// The Newton-3 pass: the tile-list and plan kernels (launch_n3_tiles), then
// the pair kernel (in this order on one stream, for every evaluation).
hipError_t launch_n3_tiles(const PairLaunch &a, hipStream_t s);
hipError_t launch_pairs_n3(const PairLaunch &a, hipStream_t s);

// Constants of the per-particle force plugins (everything but the pairs),
// one field per constructor argument of the reference's classes.
struct ForceParams {
    double confine, gamma, heat;           // Confinement C, Drag gamma, Thermal heatVal
    double bfield, drag_scale;             // Magnetic B; TimeVaryingDrag scale (offset gamma)
    double confine_x, confine_y;           // RectConfinement
    double void_decay;                     // ConfinementVoid (with confine)
    double rot_inner, rot_outer, rot_const;// Rotational
    double drive_const, drive_amplitude, drive_shift;  // Driving (drive_const = dampConst)
    double electric_field, plasma_radius;  // Electric
    double vert_electric_field, vert_decay;// VertElectric
    double gravity;                        // Gravitational
    double heat_in, heat_out, heat_radius; // ThermalLocalized
    double heat_scale, heat_offset;        // TimeVaryingThermal
    uint64_t seed;                         // thermal RNG key
};

struct UpdateLaunch {
    StageBuf cur, nxt;
    double *x, *y, *vx, *vy;        // base state
    const double *q, *m;
    const int32_t *orig;
    PairSums pairs;
    double *acc1, *acc23;           // [4][npad]: (k, l, m, n) of stage 1; stage 2 (+3)
    Scalars *sc;
    int64_t n, npad;
    int32_t forces;
    int32_t dtcheck;
    ForceParams fp;
    double *store;                  // optional copy of the current stage record: {xs ys vxs vys nc}[npad]
                                    // then the box table [npad/64][4] (store_doubles(npad) in all)
};

// stage 1..4: Runge_Kutta4's stages; kRk2Final: Runge_Kutta2's second stage
// and combine (its first stage is stage 1).
constexpr int kRk2Final = 5;
hipError_t launch_update(int stage, const UpdateLaunch &a, hipStream_t s);

// Doubles in one stored stage record (UpdateLaunch::store) and its view.
constexpr size_t store_doubles(int64_t npad) { return 5 * (size_t)npad + (size_t)npad / 16; }
__host__ __device__ inline StageBuf store_view(double *slot, int64_t npad) {
    const size_t P = (size_t)npad;
    StageBuf b{slot, slot + P, slot + 2 * P, slot + 3 * P, slot + 4 * P};
    b.box = slot + 5 * P;
    return b;
}
hipError_t launch_boundary(Scalars *sc, int closing, hipStream_t s);

// Force::forceS evaluation on a stage state: out = forces(xs, ys, vxs, vys).
struct ForceOutLaunch {
    StageBuf st;
    const double *q, *m;
    const int32_t *orig;
    PairSums pairs;
    double *fx, *fy;
    int64_t n, npad;
    int32_t forces;
    ForceParams fp;
    double t;                       // stage time (the time-dependent forces)
    uint64_t step;
    int32_t stage;
};
hipError_t launch_force_out(const ForceOutLaunch &a, hipStream_t s);

// nc = -(xs^2+ys^2)*kNormShrink for i < n; padding gets xs=ys=NaN, v=0, nc=-inf.
hipError_t launch_prepare_stage(StageBuf st, int64_t n, int64_t npad, hipStream_t s);
// `count` consecutive stored stage records from base, every particle slot
// emptied (padding form) so that only what a step writes is live.
hipError_t launch_prepare_records(double *base, int64_t count, int64_t npad, hipStream_t s);
// rmax2 = max |r|^2 over the stage positions (into sc->rmax2_bits, after reset).
hipError_t launch_rmax(StageBuf st, int64_t n, Scalars *sc, hipStream_t s);
hipError_t launch_round_fp32(const double *src, float *dst, int64_t npad, hipStream_t s);

// Discrete adjoint (adjoint.hip).  st: a stored stage record (with boxes).
hipError_t launch_adj_pairs(const StageBuf &st, const double *q, const double *ux, const double *uy,
                            double *part, int64_t n, int64_t npad, int32_t jslice, int32_t slices,
                            double kappa, double rc2p, Scalars *sc, hipStream_t s);
hipError_t launch_adj_begin(double *bar, double *ibar, double *u, double *g, const double *seed,
                            const double *m, int64_t n, int64_t npad, hipStream_t s);
hipError_t launch_adj_stage(double *bar, double *ibar, double *u, double *g, int stage,
                            const StageBuf &st, const double *q, const double *m, const double *part,
                            int32_t slices, int32_t forces, double dt, double confine, double gamma,
                            int64_t n, int64_t npad, hipStream_t s);
hipError_t launch_observe(const double *x, const double *y, const double *tgt, double *seed,
                          double *lossp, double norm, int64_t n, int64_t npad, hipStream_t s);

// Analysis (analysis.hip).
// totals[0..3] = {sum m, sum m vx, sum m vy, sum m |v|^2}, totals[4] =
// sum m |v - v_cm|^2; scratch holds npad/256 * 4 doubles.
hipError_t launch_kinetic(const double *vx, const double *vy, const double *m, int64_t n, int64_t npad,
                          double *scratch, double *totals, hipStream_t s);
// Unordered pairs i < j of the n positions binned by squared distance:
// bin b holds edges2[b] <= d^2 < edges2[b + 1], d^2 = dx*dx + dy*dy.
constexpr int32_t kMaxHistBins = 2048;
struct PairHistLaunch {
    const double *x, *y;
    int64_t n;
    int32_t nb;                   // 256-particle blocks
    int32_t nbins;                // <= kMaxHistBins
    const double *edges2;         // nbins + 1 squared bin edges
    double r2max;                 // edges2[nbins]
    float inv_width;              // nbins / r_max: the first bin guess
    unsigned long long *counts;   // nbins, accumulated into
};
hipError_t launch_pair_hist(const PairHistLaunch &a, hipStream_t s);
// amp[k] = |DFT_k(sum_{p<nuse} series[t][p])| / nuse for k < rows/2;
// s_buf: rows doubles, wre/wim: rows doubles each.
hipError_t launch_spectrum(const double *series, int64_t rows, int64_t nuse, double *s_buf, double *wre,
                           double *wim, double *amp, hipStream_t s);

}  // namespace demon
