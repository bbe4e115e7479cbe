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
// RN(sqrt(r2)) * kappa < 10 (ShieldedCoulombForce.cpp:175-178).  Outside the
// band any r2 within a few ulps of the reference's decides the same way.
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

// The reference's own contribution test for a pair inside the band:
// unfused dx*dx + dy*dy, correctly rounded sqrt, u = r * kappa, u < 10.
__device__ __forceinline__ bool exact_contributes(double dx, double dy, double kappa) {
    const double r2 = dx * dx + dy * dy;  // unfused (-ffp-contract=off)
    return sqrt(r2) * kappa < 10.0;
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

// Taylor coefficients 1/k! of e^f for k = 12 .. 3, passed to the kernels by
// value.  As kernel arguments they arrive in SGPRs, so every Horner step is
// one three-operand v_fma_f64 with an SGPR addend; as literals they would
// sit in VGPRs, where the compiler's two-operand v_fmac_f64 (which
// overwrites its addend) costs a register copy per coefficient per pair.
struct ExpCoef {
    double c[10];
};

inline ExpCoef exp_taylor_coef() {
    ExpCoef e{};
    double f = 1.0;  // k!
    for (int k = 1; k <= 12; ++k) {
        f *= k;
        if (k >= 3) e.c[12 - k] = 1.0 / f;
    }
    return e;
}

// e^-u for 0 <= u <= 16 (callers pass u < 10), to about 1 ulp.
// Cody-Waite reduction e^-u = 2^n e^f with n = rint(-u / ln 2) and
// f = -u - n ln 2 in [-ln2/2, ln2/2], ln 2 split as in fdlibm's exp
// (ln2_hi has 11 trailing zero bits, so n * ln2_hi is exact); e^f from its
// Taylor polynomial of degree 12 (remainder |f|^13/13! < 2.3e-16 relative,
// coefficients in ec);
// the scaling by 2^n is exact (no overflow or underflow in this range).
__device__ __forceinline__ double exp_neg(double u, const ExpCoef &ec) {
    constexpr double kInvLn2 = 1.4426950408889634074;
    constexpr double kLn2Hi = 6.93147180369123816490e-01;
    constexpr double kLn2Lo = 1.90821492927058770002e-10;
    const double n = __builtin_rint(u * -kInvLn2);
    double f = fma(n, -kLn2Hi, -u);
    f = fma(n, -kLn2Lo, f);
    double p = ec.c[0];
#pragma unroll
    for (int k = 1; k < 10; ++k) p = fma(p, f, ec.c[k]);
    p = fma(p, f, 0.5);
    p = fma(p, f, 1.0);
    p = fma(p, f, 1.0);
    return __builtin_ldexp(p, (int)n);
}

// Pair term (i <- j) for a contributing pair, added into (ax, ay):
//   F_i += C q_i q_j (1 + u) e^-u / r^3 * (r_i - r_j),  u = kappa r,
// the reference's cross-pair formula (ShieldedCoulombForce.cpp:183-184);
// cqi = C q_i.  r and 1/r^3 come from one refined reciprocal square root.
__device__ __forceinline__ void force_fp64(double dx, double dy, double r2, double cqi, double qj,
                                           double kappa, const ExpCoef &ec, double &ax, double &ay) {
    const double y = rsqrt_fp64(r2);
    const double u = (r2 * y) * kappa;
    const double s = (cqi * qj) * ((1.0 + u) * exp_neg(u, ec)) * ((y * y) * y);
    ax = fma(s, dx, ax);
    ay = fma(s, dy, ay);
}

// FP32 pair term: distance, exp and force in FP32 (hardware reciprocal
// square root and 2^x); the FP64 difference is rounded once, the result is
// accumulated in FP64.
__device__ __forceinline__ void force_fp32(double dx, double dy, float cqi, float qj, float kappa,
                                           double &ax, double &ay) {
    const float fx = (float)dx, fy = (float)dy;
    const float r2 = fmaf(fx, fx, fy * fy);
    const float y = __builtin_amdgcn_rsqf(r2);
    const float u = (r2 * y) * kappa;
    const float e = __builtin_amdgcn_exp2f(u * -1.44269504f);
    const float s = (cqi * qj) * ((1.0f + u) * e) * ((y * y) * y);
    ax = ax + (double)(s * fx);
    ay = ay + (double)(s * fy);
}

// Radii^2 of a launch: the candidate radius (the cutoff, widened in the
// stage-1 launch so that it also covers every pair the FP32 dt check could
// flag: r <= 1e-4 m plus the FP32 rounding of the coordinates, <= 2^-23 |r|
// each) and the dt-check radius itself.
struct Radii {
    double rf2, dt2;
};

__device__ __forceinline__ Radii launch_radii(double rc2p, int32_t flags, unsigned long long rmax2_bits) {
    Radii r{rc2p, -1.0};
    if (flags & kFlagDtCheck) {
        const double rmax = sqrt(__longlong_as_double((long long)rmax2_bits));
        const double dr = 1.001e-4 + 0x1.0p-21 * rmax;
        r.dt2 = dr * dr;
        r.rf2 = fmax(r.rf2, r.dt2);
    }
    return r;
}

// Per-wave i-particle registers of a tile.
template <int R>
struct ITile {
    double xi[R], yi[R], xm[R], ym[R], nT[R], cqi[R];
    float cqf[R];
    int32_t ibase;     // this lane's first i; i = ibase + 64 r
    int32_t wlo, whi;  // the wave's i range
};

// Exact per-pair path over all (u, r) pairs of one chunk: distance,
// contribution test (CHECK; without it every pair is known to contribute),
// force, dt check.  SELF: the chunk overlaps the wave's own i range, so
// j == i is skipped.  Padding slots (j >= n, and the lanes' own i >= n) hold
// NaN coordinates, which fail every test.
template <int R, bool F32, bool CHECK, bool SELF>
__device__ __forceinline__ void chunk_exact(const ITile<R> &t, const double (&cx)[kChunk],
                                            const double (&cy)[kChunk], const double (&cq)[kChunk],
                                            int32_t j0, int32_t n, const Cutoff &cut, float kappa_f,
                                            const ExpCoef &ec, bool dtc, double dt2,
                                            const int32_t *__restrict__ orig,
                                            unsigned int *dtmin, uint64_t &n_near, double (&ax)[R],
                                            double (&ay)[R]) {
#pragma unroll
    for (int u = 0; u < kChunk; ++u) {
        const int32_t j = j0 + u;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const double dx = t.xi[r] - cx[u];
            const double dy = t.yi[r] - cy[u];
            const double r2 = fma(dx, dx, dy * dy);
            bool in = true;
            if (CHECK) {
                // outside the band r2 decides; inside it, the reference's test
                in = r2 < cut.lo2;
                const bool band = !in && r2 < cut.hi2;
                if (__builtin_expect(__ballot(band) != 0, 0)) {
                    if (band) in = exact_contributes(dx, dy, cut.kappa);
                }
            }
            if (SELF) in = in && j != t.ibase + 64 * r;
            n_near += (uint64_t)__builtin_popcountll(__ballot(in));
            if (in) {
                if (F32)
                    force_fp32(dx, dy, t.cqf[r], (float)cq[u], kappa_f, ax[r], ay[r]);
                else
                    force_fp64(dx, dy, r2, t.cqi[r], cq[u], cut.kappa, ec, ax[r], ay[r]);
            }
            if (dtc) {
                const bool close = r2 <= dt2;
                if (__builtin_expect(__ballot(close) != 0, 0)) {
                    if (close) dtcheck_pair(t.xi[r], t.yi[r], cx[u], cy[u], orig[t.ibase + 64 * r], orig[j], dtmin);
                }
            }
        }
    }
}

// One wavefront's tile: lanes own i = ibase + 64 r (r < R), j runs over
// [j_begin, j_end) (multiples of 64, wave-uniform).  Accumulates into ax/ay
// and adds the wave's pair counts to n_near (ordered pairs inside the
// cutoff) and n_tested (ordered pairs not culled by the boxes).  Pointers are
// a cloud's arrays (device order); n its particle count.
//
// With box tables each 64-j block is first classified against the wave's
// i-box: blocks whose nearest points are beyond the candidate radius are
// skipped; blocks whose farthest points lie inside the cutoff (less the band)
// contribute every pair, so their chunks go straight to the force; the rest
// run the per-chunk prefilter.  Without box tables every block takes the
// prefilter path.
template <int R, bool F32>
__device__ __forceinline__ void wave_tile(const double *__restrict__ xs, const double *__restrict__ ys,
                                          const double *__restrict__ nc, const double *__restrict__ q,
                                          const int32_t *__restrict__ orig, const double *__restrict__ box,
                                          int32_t n, int32_t ibase, int32_t j_begin, int32_t j_end,
                                          int32_t flags, double kappa, double rc2p, Radii rad,
                                          const ExpCoef &ec, unsigned int *dtmin, double (&ax)[R],
                                          double (&ay)[R], uint64_t &n_near, uint64_t &n_tested) {
    const Cutoff cut = make_cutoff(kappa, rc2p);
    const bool dtc = (flags & kFlagDtCheck) != 0;
    const float kappa_f = (float)kappa;
    ITile<R> t;
    t.ibase = ibase;
    t.wlo = ibase - (int32_t)(threadIdx.x & 63);
    t.whi = t.wlo + 64 * R;
    const double nan = __builtin_nan("");
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int32_t i = ibase + 64 * r;
        t.xi[r] = i < n ? xs[i] : nan;
        t.yi[r] = i < n ? ys[i] : nan;
        t.xm[r] = -2.0 * t.xi[r];
        t.ym[r] = -2.0 * t.yi[r];
        t.nT[r] = -(rad.rf2 + nc[i]);
        t.cqi[r] = kCoulomb * q[i];
        t.cqf[r] = (float)t.cqi[r];
    }
    const double inf = __builtin_inf();
    double bx0 = inf, bx1 = -inf, by0 = inf, by1 = -inf;  // the wave's i-box
    if (box) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const double *b = box + 4 * (t.wlo / kBoxBlock + r);
            bx0 = fmin(bx0, b[0]);
            bx1 = fmax(bx1, b[1]);
            by0 = fmin(by0, b[2]);
            by1 = fmax(by1, b[3]);
        }
    }
    const bool ireal = t.whi <= n;  // no padding among the wave's i
    for (int32_t jb = j_begin; jb < j_end; jb += kBoxBlock) {
        const int32_t je = jb + kBoxBlock < j_end ? jb + kBoxBlock : j_end;
        bool dense = false;
        if (box) {
            const double *b = box + 4 * (jb / kBoxBlock);
            const double gx = fmax(0.0, fmax(b[0] - bx1, bx0 - b[1]));
            const double gy = fmax(0.0, fmax(b[2] - by1, by0 - b[3]));
            if (gx * gx + gy * gy > rad.rf2) continue;  // no pair within the candidate radius
            const double hx = fmax(b[1] - bx0, bx1 - b[0]);
            const double hy = fmax(b[3] - by0, by1 - b[2]);
            // every pair inside the cutoff, clear of the band (1e-9 covers rounding)
            dense = ireal && jb + kBoxBlock <= n && (hx * hx + hy * hy) * (1.0 + 1e-9) < cut.lo2;
        }
        n_tested += (uint64_t)(je - jb) * (64 * R);
        for (int32_t j0 = jb; j0 < je; j0 += kChunk) {
            double cx[kChunk], cy[kChunk], cq[kChunk];
#pragma unroll
            for (int u = 0; u < kChunk; ++u) {
                cx[u] = xs[j0 + u];
                cy[u] = ys[j0 + u];
            }
            if (!dense) {
                // prefilter: any lane with a candidate in the chunk?
                uint64_t hit = 0;
#pragma unroll
                for (int u = 0; u < kChunk; ++u) {
                    const double cn = nc[j0 + u];
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const double w = fma(t.xm[r], cx[u], fma(t.ym[r], cy[u], t.nT[r]));
                        hit |= __ballot(w < cn);
                    }
                }
                if (!hit) continue;
            }
#pragma unroll
            for (int u = 0; u < kChunk; ++u) cq[u] = q[j0 + u];
            const bool own = j0 < t.whi && j0 + kChunk > t.wlo;
            if (own)
                chunk_exact<R, F32, true, true>(t, cx, cy, cq, j0, n, cut, kappa_f, ec, dtc, rad.dt2, orig,
                                                    dtmin, n_near, ax, ay);
            else if (dense)
                chunk_exact<R, F32, false, false>(t, cx, cy, cq, j0, n, cut, kappa_f, ec, dtc, rad.dt2, orig,
                                                    dtmin, n_near, ax, ay);
            else
                chunk_exact<R, F32, true, false>(t, cx, cy, cq, j0, n, cut, kappa_f, ec, dtc, rad.dt2, orig,
                                                    dtmin, n_near, ax, ay);
        }
    }
}

}  // namespace pair
}  // namespace demon
