// pair_core.h -- the per-wavefront all-pairs tile shared by the single-cloud
// kernel (pairs.hip) and the batched ensemble kernel (batch.hip).  See
// pairs.hip for the algorithm.
#pragma once
#include "internal.h"

namespace demon {
namespace pair {

constexpr int kChunk = 8;  // j particles per fast-path chunk

enum : int32_t { kFlagDtCheck = 1, kFlagPartner = 2, kFlagGated = 4 };

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

// Exact FP64 pair term (i <- j); the candidate test has already passed.
__device__ __forceinline__ void near_fp64(double xi, double yi, double qi, int32_t oi, double xj,
                                          double yj, double qj, int32_t oj, double kappa,
                                          int32_t flags, unsigned int *dtmin, double &ax,
                                          double &ay) {
    const double dx = xi - xj;
    const double dy = yi - yj;
    const double r = sqrt(dx * dx + dy * dy);  // unfused (-ffp-contract=off), IEEE sqrt
    const double u = r * kappa;
    if (u < 10.0) {
        const double cq = kCoulomb * (qi * qj) * (1.0 + u);
        double fc;
        if ((flags & kFlagPartner) && oj == (oi ^ 1))
            fc = cq / (r * r * r * exp(u));      // intra-pair form, ShieldedCoulombForce.cpp:149-150
        else
            fc = cq * exp(-u) / (r * r * r);     // cross form, :183-184
        ax = ax + fc * dx;
        ay = ay + fc * dy;
    }
    if (flags & kFlagDtCheck) dtcheck_pair(xi, yi, xj, yj, oi, oj, dtmin);
}

// FP32 pair term: distance, exp and force in FP32; the classification and the
// accumulation stay FP64.
__device__ __forceinline__ void near_fp32(double xi, double yi, double qi, int32_t oi, double xj,
                                          double yj, double qj, int32_t oj, double kappa,
                                          double rc2, int32_t flags, unsigned int *dtmin,
                                          double &ax, double &ay) {
    const double dx = xi - xj;
    const double dy = yi - yj;
    const double r2 = dx * dx + dy * dy;
    if (r2 < rc2) {
        const float fxd = (float)dx, fyd = (float)dy;
        const float r2f = fxd * fxd + fyd * fyd;
        const float rinv = rsqrtf(r2f);
        const float uf = r2f * rinv * (float)kappa;
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
    const double rc2 = rc2p / (1.0 + kCutoffSlack);  // FP32 path: r^2 < (10/kappa)^2
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
    for (int32_t j0 = j_begin; j0 < j_end; j0 += kChunk) {
        double cx[kChunk], cy[kChunk], cn[kChunk];
#pragma unroll
        for (int u = 0; u < kChunk; ++u) {
            cx[u] = xs[j0 + u];
            cy[u] = ys[j0 + u];
            cn[u] = nc[j0 + u];
        }
        // wave-wide candidate mask of the chunk (each compare's lane mask is
        // OR-ed on the scalar unit)
        uint64_t hit = 0;
#pragma unroll
        for (int u = 0; u < kChunk; ++u) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const double w = fma(xm[r], cx[u], fma(ym[r], cy[u], nT[r]));
                hit |= __ballot(w < cn[u]);
            }
        }
        if (hit) {
#pragma unroll 1
            for (int u = 0; u < kChunk; ++u) {
                const int32_t j = j0 + u;
                const double xj = xs[j], yj = ys[j], nj = nc[j];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const double w = fma(xm[r], xj, fma(ym[r], yj, nT[r]));
                    if (w < nj && j != ibase + 64 * r) {
                        if (F32)
                            near_fp32(xi[r], yi[r], qi[r], oi[r], xj, yj, q[j], orig[j], kappa,
                                      rc2, flags, dtmin, ax[r], ay[r]);
                        else
                            near_fp64(xi[r], yi[r], qi[r], oi[r], xj, yj, q[j], orig[j], kappa,
                                      flags, dtmin, ax[r], ay[r]);
                    }
                }
            }
        }
    }
}

}  // namespace pair
}  // namespace demon
