// pair_core.h -- the per-wavefront all-pairs tile shared by the single-cloud
// kernel (pairs.hip) and the batched ensemble kernel (batch.hip).  See
// pairs.hip for the algorithm.
#pragma once
#include "internal.h"

namespace demon {
namespace pair {

constexpr int kChunk = 8;  // j particles per prefilter chunk (gather kernels)

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

// The lane's running minimum of the reference's FP32 squared distance over
// the tested pairs it has seen.
__device__ __forceinline__ void dtcheck_pair(double xi, double yi, double xj, double yj,
                                             int32_t oi, int32_t oj, float &smin) {
    if (!dtcheck_covers(oi, oj)) return;
    // (float) conversions and unfused FP32 arithmetic as loadFloatVector and
    // isWithInDistance (Integrator.cpp:162-182).
    const float sx = (float)xi - (float)xj;
    const float sy = (float)yi - (float)yj;
    const float s = sx * sx + sy * sy;
    smin = fminf(smin, s);
}

// End of a wave's tile: the wave's minimum into the device-wide minimum, one
// atomic per wave.  Only distances at or below the first threshold (1e-4f)
// can reduce dt.  Every lane of the wave must call it.
__device__ __forceinline__ void commit_dtmin(float smin, unsigned int *dtmin) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) smin = fminf(smin, __shfl_xor(smin, m));
    if ((threadIdx.x & 63) == 0 && smin <= 1.0001e-8f) atomicMin(dtmin, __float_as_uint(smin));
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

// 1/sqrt(r2) and sqrt(r2), each to about 1 ulp: the hardware estimate y0
// (relative error far below 2^-20) refined by one Halley step.  With
// e = 1 - r2 y0^2, 1/sqrt(r2) = y0 (1 - e)^-1/2 = y0 (1 + e/2 + 3e^2/8 + ...),
// so y = y0 (1 + c) with c = e (1/2 + 3e/8) leaves a relative error of about
// 5e^3/16, and sqrt(r2) = r2 y = t (1 + c) with t = r2 y0 comes from the
// same correction.
struct InvSqrt {
    double y, r;  // 1/sqrt(r2), sqrt(r2)
};

__device__ __forceinline__ InvSqrt inv_sqrt_fp64(double r2) {
    const double y0 = __builtin_amdgcn_rsq(r2);
    const double t = r2 * y0;
    const double e = fma(-t, y0, 1.0);
    const double c = e * fma(e, 0.375, 0.5);
    return InvSqrt{fma(y0, c, y0), fma(t, c, t)};
}

__device__ __forceinline__ double rsqrt_fp64(double r2) { return inv_sqrt_fp64(r2).y; }

// The same with one Newton step, c = e/2: a relative error of about 3e^2/8
// (3.5e-15 for the measured |e| <= 2^-23.2 of v_rsq_f64), one operation
// fewer.  For the Newton-3 kernel's pair force, whose bar is 1e-12 relative.
__device__ __forceinline__ InvSqrt inv_sqrt_fp64_newton(double r2) {
    const double y0 = __builtin_amdgcn_rsq(r2);
    const double t = r2 * y0;
    const double c = 0.5 * fma(-t, y0, 1.0);
    return InvSqrt{fma(y0, c, y0), fma(t, c, t)};
}

// Coefficients q_0 .. q_9 (highest power first) of the polynomial q with
// e^f ~ 1 + f + f^2 q(f) on [-ln2/2, ln2/2]: the Chebyshev interpolant of
// degree 9 of (e^f - 1 - f) / f^2 (near-minimax; 1.6e-17 relative from e^f with the
// coefficients rounded to double), produced by gen_exp_coef.py.  Passed to the
// kernels by value: as kernel arguments they arrive in SGPRs, so every Horner
// step is one three-operand v_fma_f64 with an SGPR addend; as literals they
// would sit in VGPRs, where the compiler's two-operand v_fmac_f64 (which
// overwrites its addend) costs a register copy per coefficient per pair.
struct ExpCoef {
    double c[10];
};

inline ExpCoef exp_coef() {
    return ExpCoef{{2.5100375832561234e-08, 2.7620075879983367e-07, 2.7557268480310024e-06,
                    2.4801521322368692e-05, 0.00019841269863040545, 0.0013888888917196719,
                    0.008333333333330065, 0.041666666666624164, 0.16666666666666669,
                    0.5000000000000001}};
}

// e^-u for 0 <= u <= 16 (callers pass u < 10), to about 1.5 ulp.
// Cody-Waite reduction e^-u = 2^n e^f with n = rint(-u / ln 2) and
// f = -u - n ln 2 in [-ln2/2, ln2/2], ln 2 split as in fdlibm's exp
// (ln2_hi has 11 trailing zero bits, so n * ln2_hi is exact).  n is rounded
// by adding 1.5 * 2^52 inside the fma that forms -u / ln 2: the sum's low
// word is then n itself (two's complement), so neither a rounding nor a
// conversion instruction is needed.  e^f = 1 + f (1 + f q(f)) by Horner; the
// scaling by 2^n is exact (no overflow or underflow in this range).
__device__ __forceinline__ double exp_neg(double u, const ExpCoef &ec) {
    constexpr double kLog2e = 1.4426950408889634074;
    constexpr double kRound = 0x1.8p52;
    constexpr double kLn2Hi = 6.93147180369123816490e-01;
    constexpr double kLn2Lo = 1.90821492927058770002e-10;
    const double m = fma(u, -kLog2e, kRound);
    const double n = m - kRound;
    const int ni = __double2loint(m);
    double f = fma(n, -kLn2Hi, -u);
    f = fma(n, -kLn2Lo, f);
    double p = ec.c[0];
#pragma unroll
    for (int k = 1; k < 10; ++k) p = fma(p, f, ec.c[k]);
    p = fma(p, f, 1.0);
    p = fma(p, f, 1.0);
    return __builtin_ldexp(p, ni);
}

// e^{-kappa r} for 0 <= kappa r < 10.03 (callers pass kappa r < 10) from
// a table in LDS, tab[k] = e^{-k/64} (k < kExpTabN, each the double nearest
// to a long-double evaluation on the host): n = rint(64 kappa r) by the same
// magic-number rounding (m = 1.5 2^52 + n, with 64 kappa exact), -n/64 from
// m exactly (fma(m, -1/64, 1.5 2^46): both products and the sum need at
// most 53 bits), d = kappa r - n/64 with one rounding, and e^-d by the
// degree-4 Chebyshev interpolant of e^-d on |d| <= 1/128 (gen_exp_coef.py:
// 1.5e-14 relative with the coefficients rounded to double, the constant
// term 1), well inside the pair force's 1e-12 bar.  Its coefficients
// tc[0..3] (highest power first) and the reduction's constants tc[5..8] =
// 1.5 2^46, 64, 1.5 2^52, -1/64 arrive as kernel arguments (SGPRs), as
// exp_neg's do: as literals the rounding constant would be copied into a
// VGPR pair for every evaluation.  The argument is r, not u = kappa r: the
// caller forms (1 + u)/r^3 without u (pair_scale_noq, pair_scale_env).  Four
// FMA-class operations and an add fewer than exp_neg, and no ldexp: the
// table replaces the reduction's ln 2 split and the high-degree tail.
constexpr double kExpTabScale = 64.0;

inline ExpCoef exp_tab_coef() {
    return ExpCoef{{0.041666772630595095, -0.1666673024504682, 0.4999999999983831, -0.9999999999902988, 0.0,
                    0x1.8p52 / kExpTabScale, kExpTabScale, 0x1.8p52, -1.0 / kExpTabScale, 0.0}};
}

__device__ __forceinline__ double exp_neg_tab(double r, double kappa, const double *__restrict__ tab,
                                              const ExpCoef &tc) {
    const double m = fma(r, kappa * tc.c[6], tc.c[7]);
    const int ni = __double2loint(m);
    const double d = fma(r, kappa, fma(m, tc.c[8], tc.c[5]));
    double p = tc.c[0];
#pragma unroll
    for (int k = 1; k < 4; ++k) p = fma(p, d, tc.c[k]);
    p = fma(p, d, 1.0);
    return tab[ni] * p;
}

// Pair scale for a contributing pair (i <- j):
//   F_i += s (r_i - r_j),  s = C q_i q_j (1 + u) e^-u / r^3,  u = kappa r,
// the reference's cross-pair formula (ShieldedCoulombForce.cpp:183-184);
// cqi = C q_i.  r and 1/r^3 come from one refined reciprocal square root;
// (1 + u) / r^3 is one fma.
__device__ __forceinline__ double pair_scale_fp64(double r2, double cqi, double qj, double kappa,
                                                  const ExpCoef &ec) {
    const InvSqrt s = inv_sqrt_fp64(r2);
    const double u = s.r * kappa;
    const double y3 = (s.y * s.y) * s.y;
    return ((cqi * qj) * fma(u, y3, y3)) * exp_neg(u, ec);
}

// FP32 pair force: distance, exp and force in FP32 (hardware reciprocal
// square root and 2^x) from the FP64 difference rounded once.  kl2 =
// -kappa log2(e), so a = kl2 r gives e^-u = 2^a; kf = kappa, and
// (1 + u) / r^3 = y^3 + kappa r y^3 = y^2 (y + kappa) with y = 1/r, which
// takes one operation fewer than forming 1 + u (the FP32 rounding of r y = 1
// is far below the FP32 path's error).
__device__ __forceinline__ void pair_force_fp32f(double dx, double dy, float cqi, float qj, float kl2, float kf,
                                                 float &fx, float &fy) {
    const float gx = (float)dx, gy = (float)dy;
    const float r2 = fmaf(gx, gx, gy * gy);
    const float y = __builtin_amdgcn_rsqf(r2);
    const float a = (r2 * y) * kl2;
    const float e = __builtin_amdgcn_exp2f(a);
    const float g = (y * y) * (y + kf);
    const float s = ((cqi * qj) * e) * g;
    fx = s * gx;
    fy = s * gy;
}

// -kappa log2(e) in FP32, the FP32 path's exp scale.
__host__ __device__ __forceinline__ float fp32_kl2(double kappa) {
    return (float)(-kappa * 1.4426950408889634074);
}

// The same without the factor q_j (the Newton-3 kernel applies it to the
// i-side term and to the j-side sums).
__device__ __forceinline__ void pair_force_fp32f_noq(double dx, double dy, float cqi, float kl2, float kf, float &fx,
                                                     float &fy) {
    const float gx = (float)dx, gy = (float)dy;
    const float r2 = fmaf(gx, gx, gy * gy);
    const float y = __builtin_amdgcn_rsqf(r2);
    const float a = (r2 * y) * kl2;
    const float e = __builtin_amdgcn_exp2f(a);
    const float g = (y * y) * (y + kf);
    const float s = (cqi * e) * g;
    fx = s * gx;
    fy = s * gy;
}

// The same, with the components returned in FP64 for FP64 accumulation.
__device__ __forceinline__ void pair_force_fp32(double dx, double dy, float cqi, float qj, float kl2, float kf,
                                                double &fx, double &fy) {
    float gx, gy;
    pair_force_fp32f(dx, dy, cqi, qj, kl2, kf, gx, gy);
    fx = (double)gx;
    fy = (double)gy;
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

// A load at a wave-uniform address through the scalar data cache (the
// constant address space), for the j particles of a chunk: every lane needs
// the same values, and no kernel writes them while the pair kernels read
// them.  Through a generic pointer the compiler proves this only when the
// kernel's arguments are restrict-qualified pointers, and otherwise issues
// a vector load per lane group.
template <typename T>
__device__ __forceinline__ T ld_uniform(const T *p) {
    return *(const __attribute__((address_space(4))) T *)p;
}

// Per-wave i-particle registers of a tile.
template <int R>
struct ITile {
    double xi[R], yi[R], xm[R], ym[R], nT[R], cqi[R];
    float cqf[R];
    int32_t ibase;     // this lane's first i; i = ibase + 64 r
    int32_t wlo, whi;  // the wave's i range
};

// Padding slots (i >= n) get NaN coordinates, which fail every test.
template <int R>
__device__ __forceinline__ ITile<R> make_itile(const double *__restrict__ xs, const double *__restrict__ ys,
                                               const double *__restrict__ nc, const double *__restrict__ q,
                                               int32_t n, int32_t ibase, double rf2) {
    ITile<R> t;
    t.ibase = ibase;
    // wave-uniform; readfirstlane tells the compiler so (SGPR, scalar branches)
    t.wlo = __builtin_amdgcn_readfirstlane(ibase - (int32_t)(threadIdx.x & 63));
    t.whi = t.wlo + 64 * R;
    const double nan = __builtin_nan("");
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int32_t i = ibase + 64 * r;
        t.xi[r] = i < n ? xs[i] : nan;
        t.yi[r] = i < n ? ys[i] : nan;
        t.xm[r] = -2.0 * t.xi[r];
        t.ym[r] = -2.0 * t.yi[r];
        t.nT[r] = -(rf2 + nc[i]);
        t.cqi[r] = kCoulomb * q[i];
        t.cqf[r] = (float)t.cqi[r];
    }
    return t;
}

// Axis-aligned bounding box (xmin, xmax, ymin, ymax); empty: +inf, -inf.
struct Box {
    double x0, x1, y0, y1;
};

__device__ __forceinline__ Box empty_box() {
    const double inf = __builtin_inf();
    return Box{inf, -inf, inf, -inf};
}

__device__ __forceinline__ Box box_union(Box a, const double *__restrict__ b) {
    return Box{fmin(a.x0, b[0]), fmax(a.x1, b[1]), fmin(a.y0, b[2]), fmax(a.y1, b[3])};
}

// Squared gap between two boxes (0 if they overlap; +inf if either is empty).
__device__ __forceinline__ double box_gap2(const Box &a, const Box &b) {
    const double gx = fmax(0.0, fmax(b.x0 - a.x1, a.x0 - b.x1));
    const double gy = fmax(0.0, fmax(b.y0 - a.y1, a.y0 - b.y1));
    return gx * gx + gy * gy;
}

// Squared largest distance between points of two boxes.
__device__ __forceinline__ double box_span2(const Box &a, const Box &b) {
    const double hx = fmax(b.x1 - a.x0, a.x1 - b.x0);
    const double hy = fmax(b.y1 - a.y0, a.y1 - b.y0);
    return hx * hx + hy * hy;
}

enum BlockClass { kSkip, kMixed, kDense };

// A 64-j block against a wave's i-box: no pair within the candidate radius
// (skip), every pair inside the cutoff and clear of the band (dense; 1e-9
// covers the rounding of the box arithmetic; needs no padding on either
// side), or per-chunk prefiltering (mixed).
__device__ __forceinline__ BlockClass classify(const Box &ib, const double *__restrict__ jbox, bool ireal,
                                               bool jreal, const Radii &rad, const Cutoff &cut) {
    const Box jb{jbox[0], jbox[1], jbox[2], jbox[3]};
    if (box_gap2(ib, jb) > rad.rf2) return kSkip;
    if (ireal && jreal && box_span2(ib, jb) * (1.0 + 1e-9) < cut.lo2) return kDense;
    return kMixed;
}

// Any lane of the wave with a candidate pair in the chunk?  Per pair two
// FMAs and a compare on the expanded squared distance
//   w = (-2 x_i) x_j + (-2 y_i) y_j - (Rf^2 + nc_i) < nc_j,
//   nc = -|r|^2 (1 - 2^-46): a superset of the pairs within Rf.
template <int R, int C>
__device__ __forceinline__ bool chunk_hit(const ITile<R> &t, const double (&cx)[C],
                                          const double (&cy)[C], const double *__restrict__ nc, int32_t j0) {
    uint64_t hit = 0;
#pragma unroll
    for (int u = 0; u < C; ++u) {
        const double cn = ld_uniform(nc + j0 + u);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const double w = fma(t.xm[r], cx[u], fma(t.ym[r], cy[u], t.nT[r]));
            hit |= __ballot(w < cn);
        }
    }
    return hit != 0;
}

// Launch-wide constants of the exact pair path.
struct PairEnv {
    Cutoff cut;
    Radii rad;
    float kl2_f;  // -kappa log2(e), FP32 pair arithmetic
    bool dtc;
    const int32_t *__restrict__ orig;
    ExpCoef ec;   // exp_neg's coefficients, or exp_neg_tab's with etab
    const double *etab = nullptr;  // e^{-k/64} in LDS (kernels built with TAB)
};

// 1/sqrt(r2) after one Newton step and e^{-kappa r} from the LDS table
// (exp_neg_tab), with the table read issued before the refinement: its index
// n = rint(64 kappa t) comes from the first estimate t = r2 y0 (relative
// error 2^-24.2, so n may differ from rint(64 kappa r) within 5e-7 of a
// half-step), and d = kappa r - n/64 uses the refined r, so |d| stays within
// 1/128 + 5.2e-7 (kappa r < 10), where the degree-4 interpolant keeps its
// 1.5e-14 (beyond its interval the error grows like T_5(|d| / h), h = 1/128:
// by 25 (|d| - h) / h = 0.17% here).
struct InvSqrtExp {
    double y, e;  // 1/sqrt(r2), e^{-kappa sqrt(r2)}
};
__device__ __forceinline__ InvSqrtExp inv_sqrt_exp_tab(double r2, double kappa, const double *__restrict__ tab,
                                                       const ExpCoef &tc) {
    const double y0 = __builtin_amdgcn_rsq(r2);
    const double t = r2 * y0;
    const double m = fma(t, kappa * tc.c[6], tc.c[7]);
    const double te = tab[__double2loint(m)];
    const double c = 0.5 * fma(-t, y0, 1.0);
    const double y = fma(y0, c, y0), r = fma(t, c, t);
    const double d = fma(r, kappa, fma(m, tc.c[8], tc.c[5]));
    double p = tc.c[0];
#pragma unroll
    for (int k = 1; k < 4; ++k) p = fma(p, d, tc.c[k]);
    p = fma(p, d, 1.0);
    return InvSqrtExp{y, te * p};
}

// The pair scale without the factor q_j, C q_i (1 + u) e^-u / r^3, with
// the launch's exp: exp_neg_tab on the LDS table (TAB; with the Newton-step
// reciprocal square root), where (1 + u)/r^3 = y^3 + kappa r y^3 is formed
// as y^2 (y + kappa) (r y = 1 to about 1e-29 after the refinement) and u
// itself is never needed, or exp_neg.
template <bool TAB>
__device__ __forceinline__ double pair_scale_noq(double r2, double cqi, const PairEnv &e) {
    if constexpr (TAB) {
        const InvSqrtExp s = inv_sqrt_exp_tab(r2, e.cut.kappa, e.etab, e.ec);
        const double g = (s.y * s.y) * (s.y + e.cut.kappa);
        return (cqi * g) * s.e;
    } else {
        const InvSqrt s = inv_sqrt_fp64(r2);
        const double u = s.r * e.cut.kappa;
        const double y3 = (s.y * s.y) * s.y;
        return (cqi * fma(u, y3, y3)) * exp_neg(u, e.ec);
    }
}

// The pair scale with the launch's exp: exp_neg_tab on the LDS table (TAB)
// or exp_neg.
template <bool TAB>
__device__ __forceinline__ double pair_scale_env(double r2, double cqi, double qj, const PairEnv &e) {
    if constexpr (TAB) {
        const InvSqrtExp s = inv_sqrt_exp_tab(r2, e.cut.kappa, e.etab, e.ec);
        const double g = (s.y * s.y) * (s.y + e.cut.kappa);
        return ((cqi * qj) * g) * s.e;
    } else {
        return pair_scale_fp64(r2, cqi, qj, e.cut.kappa, e.ec);
    }
}

// Distance and contribution test of one pair: outside the band r2 decides;
// inside it, the reference's own test (CHECK; without it the pair is known
// to contribute).  Then the dt check of the stage-1 launch.
template <bool CHECK>
__device__ __forceinline__ bool pair_in(double dx, double dy, double r2, const PairEnv &e) {
    bool in = true;
    if (CHECK) {
        in = r2 < e.cut.lo2;
        const bool band = !in && r2 < e.cut.hi2;
        if (__builtin_expect(__ballot(band) != 0, 0)) {
            if (band) in = exact_contributes(dx, dy, e.cut.kappa);
        }
    }
    return in;
}

__device__ __forceinline__ void pair_dtcheck(double xi, double yi, double xj, double yj, double r2,
                                             int32_t i, int32_t j, const PairEnv &e, float &smin) {
    if (e.dtc) {
        const bool close = r2 <= e.rad.dt2;
        if (__builtin_expect(__ballot(close) != 0, 0)) {
            if (close) dtcheck_pair(xi, yi, xj, yj, e.orig[i], e.orig[j], smin);
        }
    }
}

// Gather form over all (u, r) pairs of one chunk: accumulates F_i into
// ax/ay.  SELF: the chunk overlaps the wave's own i range, so j == i is
// skipped.
template <int R, int C, bool F32, bool CHECK, bool SELF, bool TAB = false>
__device__ __forceinline__ void chunk_gather(const ITile<R> &t, const double (&cx)[C],
                                             const double (&cy)[C], const double (&cq)[C],
                                             int32_t j0, const PairEnv &e, uint64_t &n_near,
                                             float &smin, double (&ax)[R], double (&ay)[R]) {
#pragma unroll
    for (int u = 0; u < C; ++u) {
        const int32_t j = j0 + u;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const double dx = t.xi[r] - cx[u];
            const double dy = t.yi[r] - cy[u];
            const double r2 = fma(dx, dx, dy * dy);
            bool in = pair_in<CHECK>(dx, dy, r2, e);
            if (SELF) in = in && j != t.ibase + 64 * r;
            n_near += (uint64_t)__builtin_popcountll(__ballot(in));
            if (in) {
                if (F32) {
                    double fx, fy;
                    pair_force_fp32(dx, dy, t.cqf[r], (float)cq[u], e.kl2_f, (float)e.cut.kappa, fx, fy);
                    ax[r] = ax[r] + fx;
                    ay[r] = ay[r] + fy;
                } else {
                    const double s = pair_scale_env<TAB>(r2, t.cqi[r], cq[u], e);
                    ax[r] = fma(s, dx, ax[r]);
                    ay[r] = fma(s, dy, ay[r]);
                }
            }
            pair_dtcheck(t.xi[r], t.yi[r], cx[u], cy[u], r2, t.ibase + 64 * r, j, e, smin);
        }
    }
}

// Gather form over the 64-j blocks [j_begin, j_end) (multiples of 64,
// wave-uniform) in chunks of C, with box classification when box tables are
// given.
// Adds the wave's counts to n_near (ordered pairs inside the cutoff) and
// n_tested (ordered pairs not culled by the boxes).
template <int R, int C, bool F32, bool TAB = false>
__device__ __forceinline__ void gather_blocks(const ITile<R> &t, const double *__restrict__ xs,
                                              const double *__restrict__ ys, const double *__restrict__ nc,
                                              const double *__restrict__ q, const double *__restrict__ box,
                                              const Box &ib, int32_t n, int32_t j_begin, int32_t j_end,
                                              const PairEnv &e, double (&ax)[R], double (&ay)[R],
                                              uint64_t &n_near, uint64_t &n_tested, float &smin) {
    const bool ireal = t.whi <= n;
    for (int32_t jb = j_begin; jb < j_end; jb += kBoxBlock) {
        const int32_t je = jb + kBoxBlock < j_end ? jb + kBoxBlock : j_end;
        BlockClass cls = kMixed;
        if (box) {
            cls = classify(ib, box + 4 * (jb / kBoxBlock), ireal, jb + kBoxBlock <= n, e.rad, e.cut);
            if (cls == kSkip) continue;
        }
        n_tested += (uint64_t)(je - jb) * (64 * R);
        for (int32_t j0 = jb; j0 < je; j0 += C) {
            double cx[C], cy[C], cq[C];
#pragma unroll
            for (int u = 0; u < C; ++u) {
                cx[u] = ld_uniform(xs + j0 + u);
                cy[u] = ld_uniform(ys + j0 + u);
            }
            if constexpr (C == 4) {
                // four j: positions, norms and charges in one batch of
                // scalar loads (one round trip per chunk; the empty
                // statement keeps the charge loads from being sunk behind
                // the prefilter's branch)
                double cn[C];
#pragma unroll
                for (int u = 0; u < C; ++u) {
                    cn[u] = ld_uniform(nc + j0 + u);
                    cq[u] = ld_uniform(q + j0 + u);
                }
                asm volatile("" ::"s"(cx[0]), "s"(cx[1]), "s"(cx[2]), "s"(cx[3]), "s"(cy[0]), "s"(cy[1]), "s"(cy[2]),
                             "s"(cy[3]), "s"(cn[0]), "s"(cn[1]), "s"(cn[2]), "s"(cn[3]), "s"(cq[0]), "s"(cq[1]),
                             "s"(cq[2]), "s"(cq[3]));
                if (cls != kDense) {
                    uint64_t hit = 0;
#pragma unroll
                    for (int u = 0; u < C; ++u)
#pragma unroll
                        for (int r = 0; r < R; ++r)
                            hit |= __ballot(fma(t.xm[r], cx[u], fma(t.ym[r], cy[u], t.nT[r])) < cn[u]);
                    if (!hit) continue;
                }
            } else {
                if (cls != kDense && !chunk_hit<R, C>(t, cx, cy, nc, j0)) continue;
#pragma unroll
                for (int u = 0; u < C; ++u) cq[u] = ld_uniform(q + j0 + u);
            }
            const bool own = j0 < t.whi && j0 + C > t.wlo;
            if (own)
                chunk_gather<R, C, F32, true, true, TAB>(t, cx, cy, cq, j0, e, n_near, smin, ax, ay);
            else if (cls == kDense)
                chunk_gather<R, C, F32, false, false, TAB>(t, cx, cy, cq, j0, e, n_near, smin, ax, ay);
            else
                chunk_gather<R, C, F32, true, false, TAB>(t, cx, cy, cq, j0, e, n_near, smin, ax, ay);
        }
    }
}

// One wavefront's gather tile: lanes own i = ibase + 64 r (r < R), j runs
// over [j_begin, j_end).  Pointers are a cloud's arrays (device order), n
// its particle count; box may be null (no culling).
template <int R, bool F32>
__device__ __forceinline__ void wave_tile(const double *__restrict__ xs, const double *__restrict__ ys,
                                          const double *__restrict__ nc, const double *__restrict__ q,
                                          const int32_t *__restrict__ orig, const double *__restrict__ box,
                                          int32_t n, int32_t ibase, int32_t j_begin, int32_t j_end,
                                          int32_t flags, double kappa, double rc2p, Radii rad,
                                          const ExpCoef &ec, unsigned int *dtmin, double (&ax)[R],
                                          double (&ay)[R], uint64_t &n_near, uint64_t &n_tested) {
    const PairEnv e{make_cutoff(kappa, rc2p), rad, fp32_kl2(kappa), (flags & kFlagDtCheck) != 0, orig, ec};
    const ITile<R> t = make_itile<R>(xs, ys, nc, q, n, ibase, rad.rf2);
    Box ib = empty_box();
    if (box) {
#pragma unroll
        for (int r = 0; r < R; ++r) ib = box_union(ib, box + 4 * (t.wlo / kBoxBlock + r));
    }
    float smin = __builtin_inff();
    gather_blocks<R, kChunk, F32>(t, xs, ys, nc, q, box, ib, n, j_begin, j_end, e, ax, ay, n_near, n_tested,
                                  smin);
    if (e.dtc) commit_dtmin(smin, dtmin);
}

}  // namespace pair
}  // namespace demon
