// pair_core.h -- the per-wavefront all-pairs tile shared by the single-cloud
// kernel (pairs.hip) and the batched ensemble kernel (batch.hip).  See
// pairs.hip for the algorithm.
#pragma once
#include "internal.h"

namespace demon {
namespace pair {

constexpr int kChunk = 8;  // j particles per fast-path chunk

enum : int32_t { kFlagDtCheck = 1, kFlagGated = 4 };

// Relative half-width of the band around the cutoff radius^2 inside which
// the contribution test falls back to the reference's exact form
// RN(sqrt(r2)) * kappa < 10 (ShieldedCoulombForce.cpp:175-178).
constexpr double kBand = 1.0e-12;

// Pairs tested by the reference's modifyTimeStep, in caller indices: all six
// pairs inside a group of four, and cross-group pairs except j = i+1 (mod 4)
// for i in the earlier group (its four shuffles cover offsets 0, -1, -2 only,
// one of them twice: Integrator.cpp:112-129).
__device__ __forceinline__ bool dtcheck_covers(int32_t a, int32_t b) {
    const int32_t lo = a < b ? a : b;
    const int32_t hi = a < b ? b : a;
    if ((lo >> 2) == (hi >> 2)) return lo != hi;
    return (hi & 3) != ((lo + 1) & 3);
}

__device__ __forceinline__ void dtcheck_pair(double xi, double yi, double xj, double yj,
                                             int32_t oi, int32_t oj, unsigned int *dtmin) {
    if (!dtcheck_covers(oi, oj)) return;
    // (float) conversions and unfused FP32 arithmetic as loadFloatVector and
    // isWithInDistance (Integrator.cpp:162-182).
    const float sx = (float)xi - (float)xj;
    const float sy = (float)yi - (float)yj;
    const float s = sx * sx + sy * sy;
    // Only distances at or below the first threshold (1e-4f) can reduce dt.
    if (s <= 1.0001e-8f) atomicMin(dtmin, __float_as_uint(s));
}

// Cutoff radius^2 (10/kappa)^2 of a launch and the band around it.
struct Cutoff {
    double lo2, hi2, kappa;
};

__device__ __forceinline__ Cutoff make_cutoff(double kappa, double rc2p) {
    const double rc2 = rc2p / (1.0 + kCutoffSlack);
    return Cutoff{rc2 * (1.0 - kBand), rc2 * (1.0 + kBand), kappa};
}

// Does the pair contribute, exactly as the reference decides?  Outside the
// band r2 alone decides; inside it the reference's own rounding is replayed.
__device__ __forceinline__ bool contributes(double r2, const Cutoff &c) {
    if (r2 < c.lo2) return true;
    if (r2 >= c.hi2) return false;
    return sqrt(r2) * c.kappa < 10.0;
}

// 1/sqrt(r2) to about 1 ulp: the hardware estimate refined by two Newton
// steps y <- y + y (1 - r2 y^2) / 2 (each doubles the correct bits).
__device__ __forceinline__ double rsqrt_fp64(double r2) {
    double y = __builtin_amdgcn_rsq(r2);
    double e = fma(-r2 * y, y, 1.0);
    y = fma(0.5 * y, e, y);
    e = fma(-r2 * y, y, 1.0);
    return fma(0.5 * y, e, y);
}

// FP64 pair term (i <- j) for a candidate pair:
//   F_i += C q_i q_j (1 + u) e^-u / r^3 * (r_i - r_j),  u = kappa r,
// the reference's cross-pair form (ShieldedCoulombForce.cpp:183-184).  The
// distance is the reference's (dx = xi - xj, unfused dx*dx + dy*dy); r and
// 1/r^3 come from one refined reciprocal square root instead of a square
// root and a division.
__device__ __forceinline__ void near_fp64(double xi, double yi, double qi, int32_t oi, double xj,
                                          double yj, double qj, int32_t oj, const Cutoff &cut,
                                          int32_t flags, unsigned int *dtmin, double &ax,
                                          double &ay) {
    const double dx = xi - xj;
    const double dy = yi - yj;
    const double r2 = dx * dx + dy * dy;  // unfused (-ffp-contract=off)
    if (contributes(r2, cut)) {
        const double y = rsqrt_fp64(r2);
        const double u = (r2 * y) * cut.kappa;
        const double fc = kCoulomb * (qi * qj) * (1.0 + u) * exp(-u) * (y * y * y);
        ax = ax + fc * dx;
        ay = ay + fc * dy;
    }
    if (flags & kFlagDtCheck) dtcheck_pair(xi, yi, xj, yj, oi, oj, dtmin);
}

// FP32 pair term: distance, exp and force in FP32; the contribution test and
// the accumulation stay FP64.
__device__ __forceinline__ void near_fp32(double xi, double yi, double qi, int32_t oi, double xj,
                                          double yj, double qj, int32_t oj, const Cutoff &cut,
                                          int32_t flags, unsigned int *dtmin, double &ax,
                                          double &ay) {
    const double dx = xi - xj;
    const double dy = yi - yj;
    const double r2 = dx * dx + dy * dy;
    if (contributes(r2, cut)) {
        const float fxd = (float)dx, fyd = (float)dy;
        const float r2f = fxd * fxd + fyd * fyd;
        const float rinv = rsqrtf(r2f);
        const float uf = r2f * rinv * (float)cut.kappa;
        const float cqf = (float)(kCoulomb * (qi * qj));
        const float fc = cqf * (1.0f + uf) * expf(-uf) * (rinv * rinv * rinv);
        ax = ax + (double)(fc * fxd);
        ay = ay + (double)(fc * fyd);
    }
    if (flags & kFlagDtCheck) dtcheck_pair(xi, yi, xj, yj, oi, oj, dtmin);
}

// Candidate radius^2 of a launch: the cutoff, widened in the stage-1 launch
// so that it also covers every pair the FP32 dt check could flag (r <= 1e-4 m
// plus the FP32 rounding of the coordinates, <= 2^-23 |r| each).
__device__ __forceinline__ double candidate_r2(double rc2p, int32_t flags, unsigned long long rmax2_bits) {
    double rf2 = rc2p;
    if (flags & kFlagDtCheck) {
        const double rmax = sqrt(__longlong_as_double((long long)rmax2_bits));
        const double dr = 1.001e-4 + 0x1.0p-21 * rmax;
        rf2 = fmax(rf2, dr * dr);
    }
    return rf2;
}

// One wavefront's tile: lanes own i = ibase + 64 r (r < R), j runs over
// [j_begin, j_end) (multiples of kChunk, wave-uniform).  Accumulates into
// ax/ay.  Pointers are a cloud's arrays (device order).
template <int R, bool F32>
__device__ __forceinline__ void wave_tile(const double *__restrict__ xs, const double *__restrict__ ys,
                                          const double *__restrict__ nc, const double *__restrict__ q,
                                          const int32_t *__restrict__ orig, int32_t ibase,
                                          int32_t j_begin, int32_t j_end, int32_t flags, double kappa,
                                          double rc2p, double rf2, unsigned int *dtmin, double (&ax)[R],
                                          double (&ay)[R]) {
    const Cutoff cut = make_cutoff(kappa, rc2p);
    double xi[R], yi[R], xm[R], ym[R], nT[R], qi[R];
    int32_t oi[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int32_t i = ibase + 64 * r;
        xi[r] = xs[i];
        yi[r] = ys[i];
        xm[r] = -2.0 * xi[r];
        ym[r] = -2.0 * yi[r];
        nT[r] = -(rf2 + nc[i]);
        qi[r] = q[i];
        oi[r] = orig[i];
    }
    // the wave's own i range [wlo, whi): only chunks overlapping it can hold j == i
    const int32_t wlo = ibase - (int32_t)(threadIdx.x & 63);
    const int32_t whi = wlo + 64 * R;
    for (int32_t j0 = j_begin; j0 < j_end; j0 += kChunk) {
        double cx[kChunk], cy[kChunk], cn[kChunk];
#pragma unroll
        for (int u = 0; u < kChunk; ++u) {
            cx[u] = xs[j0 + u];
            cy[u] = ys[j0 + u];
            cn[u] = nc[j0 + u];
        }
        // candidate masks; their OR is formed on the scalar unit
        bool cand[kChunk][R];
        uint64_t hit = 0;
#pragma unroll
        for (int u = 0; u < kChunk; ++u) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const double w = fma(xm[r], cx[u], fma(ym[r], cy[u], nT[r]));
                cand[u][r] = w < cn[u];
                hit |= __ballot(cand[u][r]);
            }
        }
        if (hit) {
            const bool own = j0 < whi && j0 + kChunk > wlo;
#pragma unroll
            for (int u = 0; u < kChunk; ++u) {
                const int32_t j = j0 + u;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    if (cand[u][r] && (!own || j != ibase + 64 * r)) {
                        if (F32)
                            near_fp32(xi[r], yi[r], qi[r], oi[r], cx[u], cy[u], q[j], orig[j], cut,
                                      flags, dtmin, ax[r], ay[r]);
                        else
                            near_fp64(xi[r], yi[r], qi[r], oi[r], cx[u], cy[u], q[j], orig[j], cut,
                                      flags, dtmin, ax[r], ay[r]);
                    }
                }
            }
        }
    }
}

}  // namespace pair
}  // namespace demon
