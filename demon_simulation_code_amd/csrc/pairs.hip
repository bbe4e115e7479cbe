// pairs.hip -- all-pairs shielded-Coulomb (Yukawa) force for gfx950.
//
// Replaces the O(N^2) loop of ShieldedCoulombForce::force1..4
// (ShieldedCoulombForce.cpp:37-237).  Physics per ordered pair (i <- j), as in
// the reference (:145-150, :175-184):
//     d = r_i - r_j, r = sqrt(dx^2 + dy^2), u = r * kappa,
//     if u < 10:  F_i += C q_i q_j (1 + u) e^{-u} / r^3 * d      (C = 1/(4 pi eps0)).
//
// Layout: the device keeps particles in a spatially sorted order (Morton
// order of the uploaded positions) so that a wavefront's 64 i-particles form
// a compact patch.  Each lane owns R i-particles and accumulates their force
// in registers (gather form: no atomics, deterministic).  j-particles are
// wave-uniform, so their records (x, y, nc) are read with scalar loads and
// enter the VALU as SGPR operands; no LDS is needed.
//
// Per pair the fast path is two FMAs and a compare: the squared distance is
// tested in the expanded form
//     w = (-2 x_i) x_j + (-2 y_i) y_j - (Rf^2 + nc_i)  <  nc_j ,
//     nc = -|r|^2 (1 - 2^-46),
// which says r^2 < Rf^2 + 2^-46 (|r_i|^2 + |r_j|^2): a superset of the
// reference's cutoff set (the margin exceeds the expansion's rounding error).
// Only when some lane of the wave has a candidate in the current 8-j chunk
// does the wave take the exact path, which recomputes the distance exactly as
// the reference (unfused dx*dx + dy*dy, correctly rounded sqrt, u = r*kappa,
// u < 10), so the set of contributing pairs is the reference's.
//
// The FP32 dt check of Integrator::modifyTimeStep (Integrator.cpp:68-134) is
// folded into the stage-1 launch: every candidate pair the reference tests
// (its lane-shuffle coverage, see dtcheck_covers) contributes its FP32
// squared distance to a device-wide minimum.
#include "internal.h"

namespace demon {

namespace {

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

template <int R, bool F32>
__global__ __launch_bounds__(256) void k_pairs(const double *__restrict__ xs,
                                               const double *__restrict__ ys,
                                               const double *__restrict__ nc,
                                               const double *__restrict__ q,
                                               const int32_t *__restrict__ orig,
                                               double *__restrict__ part,
                                               Scalars *__restrict__ sc, int64_t npad,
                                               int32_t jslice, int32_t flags, double kappa,
                                               double rc2p) {
    if ((flags & kFlagGated) && !sc->active) return;

    // Candidate radius: the cutoff, widened in the stage-1 launch so that it
    // also covers every pair the FP32 dt check could flag (r <= 1e-4 m plus
    // the FP32 rounding of the coordinates, <= 2^-23 |r| each).
    double rf2 = rc2p;
    if (flags & kFlagDtCheck) {
        const double rmax = sqrt(__longlong_as_double((long long)sc->rmax2_bits));
        const double dr = 1.001e-4 + 0x1.0p-21 * rmax;
        rf2 = fmax(rf2, dr * dr);
    }
    const double rc2 = rc2p / (1.0 + kCutoffSlack);  // FP32 path: r^2 < (10/kappa)^2

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t ibase = (int64_t)blockIdx.x * (256 * R) + wave * (64 * R) + lane;

    double xi[R], yi[R], xm[R], ym[R], nT[R], qi[R], ax[R], ay[R];
    int32_t oi[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int64_t i = ibase + 64 * r;
        xi[r] = xs[i];
        yi[r] = ys[i];
        xm[r] = -2.0 * xi[r];
        ym[r] = -2.0 * yi[r];
        nT[r] = -(rf2 + nc[i]);
        qi[r] = q[i];
        oi[r] = orig[i];
        ax[r] = 0.0;
        ay[r] = 0.0;
    }

    const int32_t j_begin = (int32_t)blockIdx.y * jslice;
    const int32_t j_end = j_begin + jslice < (int32_t)npad ? j_begin + jslice : (int32_t)npad;
    unsigned int *dtmin = &sc->dtmin_bits;

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
                    if (w < nj && (int64_t)j != ibase + 64 * r) {
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

    double *px = part + (size_t)(2 * blockIdx.y) * npad;
    double *py = px + npad;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        px[ibase + 64 * r] = ax[r];
        py[ibase + 64 * r] = ay[r];
    }
}

template <int R, bool F32>
hipError_t launch_shape(const PairLaunch &a, hipStream_t s) {
    const dim3 grid((unsigned)(a.npad / (256 * R)), (unsigned)a.slices);
    int32_t flags = 0;
    if (a.dtcheck) flags |= kFlagDtCheck;
    if (a.partner) flags |= kFlagPartner;
    if (a.gated) flags |= kFlagGated;
    hipLaunchKernelGGL((k_pairs<R, F32>), grid, dim3(256), 0, s, a.xs, a.ys, a.nc, a.q, a.orig,
                       a.part, a.sc, a.npad, a.jslice, flags, a.kappa, a.rc2p);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_pairs(const PairLaunch &a, hipStream_t s) {
    if (a.npad % kPad || a.jslice % kChunk || a.slices < 1) return hipErrorInvalidValue;
    if ((int64_t)a.slices * a.jslice < a.npad) return hipErrorInvalidValue;
    switch (a.i_per_lane * 2 + (a.fp32pair ? 1 : 0)) {
        case 2: return launch_shape<1, false>(a, s);
        case 3: return launch_shape<1, true>(a, s);
        case 4: return launch_shape<2, false>(a, s);
        case 5: return launch_shape<2, true>(a, s);
        case 8: return launch_shape<4, false>(a, s);
        case 9: return launch_shape<4, true>(a, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace demon
