// pairs_n3.hip -- all-pairs shielded-Coulomb force with Newton's third law:
// every unordered pair is evaluated once and applied to both particles, as
// the reference does (ShieldedCoulombForce.cpp:51-56 add the force to one
// particle and subtract it from the other).  Same physics, contribution test
// and dt check as the gather kernel (pairs.hip); half the pair evaluations.
//
// Work decomposition.  Particles (device order) form superblocks of 256
// (four 64-particle box blocks).  With nsb superblocks, row S pairs with
// superblocks T = S + d (mod nsb) for d = 1 .. nsb/2 (d = nsb/2 only for
// S < nsb/2, so that each unordered superblock pair is taken once), plus
// itself (d = 0, gather form).  Every row does the same amount of work.
// A workgroup of four waves owns row S and a contiguous range of d (a
// slice); wave w holds the 64 i-particles of block 4S + w in registers and
// accumulates their forces across the slice (written once, to part[slice]).
// For each pair tile (S, T) the reaction on every j of T is summed over the
// wave's 64 lanes, then over the four waves in LDS in a fixed order, and
// written to react[d - 1][.][j]; rflag[T nsb/2 + d - 1] says whether the
// tile was evaluated (1) or culled / not owned (0).  Nothing is atomic, so
// the sums are deterministic: the stage kernel adds part[0..slices) and then
// react[d - 1] for d = 1 .. nsb/2 in that order.
#include "pair_core.h"

namespace demon {

namespace {

using namespace pair;

// j particles per chunk: four keeps the chunk's j records (x, y, q) in 24
// SGPRs next to the exp coefficients without spilling.
constexpr int kChunkN3 = 4;

// Wave sums of the 8 per-lane values v[4 c + u] (component c of the pair
// force on j0 + u), transposed through LDS: every lane writes its 8 values
// into scr[v][lane] (rows padded to 66 doubles, which keeps both the writes
// and the 16-byte reads below at the minimum bank load), then lane
// (v = l / 8, p = l % 8) adds scr[v][8p .. 8p + 7] and three butterflies over
// the 8 lanes of value v finish the sum, which lane p = 0 stores into
// red[2 (jj0 + u) + c].  10 VALU per chunk; the order of the additions is
// fixed, so the result is deterministic.
constexpr int kScrRow = 66;

__device__ __forceinline__ void store_wave_sums(const double (&f)[2 * kChunkN3], double *scr, double *red,
                                                int32_t jj0) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int v = 0; v < 2 * kChunkN3; ++v) scr[v * kScrRow + lane] = f[v];
    __builtin_amdgcn_wave_barrier();
    const int v = lane >> 3, p = lane & 7;
    const double *row = scr + v * kScrRow + 8 * p;
    double t = row[0];
#pragma unroll
    for (int k = 1; k < 8; ++k) t += row[k];
    __builtin_amdgcn_wave_barrier();
    t += __shfl_xor(t, 1);
    t += __shfl_xor(t, 2);
    t += __shfl_xor(t, 4);
    if (p == 0) red[2 * (jj0 + (v & 3)) + (v >> 2)] = t;
}

// Newton-3 form over one chunk of 4 j: adds F_ij into ax/ay and stores the
// wave's sum of F_ij for each j into red[2 jj], red[2 jj + 1] (the
// reaction, negated later).
template <bool F32, bool CHECK>
__device__ __forceinline__ void chunk_n3(const ITile<1> &t, const double (&cx)[kChunkN3],
                                         const double (&cy)[kChunkN3], const double (&cq)[kChunkN3],
                                         int32_t j0, const PairEnv &e, uint64_t &n_near, float &smin,
                                         double &ax, double &ay, double *scr, double *red, int32_t jj0) {
    double f[2 * kChunkN3];
#pragma unroll
    for (int u = 0; u < kChunkN3; ++u) {
        const double dx = t.xi[0] - cx[u];
        const double dy = t.yi[0] - cy[u];
        const double r2 = fma(dx, dx, dy * dy);
        const bool in = pair_in<CHECK>(dx, dy, r2, e);
        n_near += 2 * (uint64_t)__builtin_popcountll(__ballot(in));
        double fx = 0.0, fy = 0.0;
        if (in) {
            if (F32) {
                pair_force_fp32(dx, dy, t.cqf[0], (float)cq[u], e.kappa_f, fx, fy);
            } else {
                const double s = pair_scale_fp64(r2, t.cqi[0], cq[u], e.cut.kappa, e.ec);
                fx = s * dx;
                fy = s * dy;
            }
        }
        ax = ax + fx;
        ay = ay + fy;
        f[u] = fx;
        f[kChunkN3 + u] = fy;
        pair_dtcheck(t.xi[0], t.yi[0], cx[u], cy[u], r2, t.ibase, j0 + u, e, smin);
    }
    store_wave_sums(f, scr, red, jj0);
}

template <bool F32>
__global__ __launch_bounds__(256) void k_pairs_n3(const double *__restrict__ xs,
                                                  const double *__restrict__ ys,
                                                  const double *__restrict__ nc,
                                                  const double *__restrict__ q,
                                                  const int32_t *__restrict__ orig,
                                                  const double *__restrict__ box,
                                                  double *__restrict__ part,
                                                  double *__restrict__ react,
                                                  uint32_t *__restrict__ rflag,
                                                  Scalars *__restrict__ sc, int32_t n, int32_t npad,
                                                  int32_t span, int32_t flags, double kappa,
                                                  double rc2p, ExpCoef ec) {
    __shared__ double red[4][2 * kSuper];                  // [wave][2 jj + component]
    __shared__ double scr[4][2 * kChunkN3 * kScrRow];      // per-wave transpose scratch
    if ((flags & kFlagGated) && !sc->active) return;
    const Radii rad = launch_radii(rc2p, flags, sc->rmax2_bits);
    const PairEnv e{make_cutoff(kappa, rc2p), rad, (float)kappa, (flags & kFlagDtCheck) != 0, orig, ec};
    float smin = __builtin_inff();
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int32_t nsb = npad / kSuper;
    const int32_t half = nsb / 2;
    const int32_t S = (int32_t)blockIdx.x;
    const int32_t d0 = (int32_t)blockIdx.y * span;
    const int32_t d1 = d0 + span < half + 1 ? d0 + span : half + 1;
    const int32_t ibase = S * kSuper + wave * 64 + lane;
    const ITile<1> t = make_itile<1>(xs, ys, nc, q, n, ibase, rad.rf2);
    const Box ib = box_union(empty_box(), box + 4 * (t.wlo / kBoxBlock));
    Box sbS = empty_box();
#pragma unroll
    for (int k = 0; k < 4; ++k) sbS = box_union(sbS, box + 4 * (S * 4 + k));
    const bool ireal = t.whi <= n;
    double ax[1] = {0.0}, ay[1] = {0.0};
    uint64_t n_near = 0, n_tested = 0;
    for (int32_t d = d0; d < d1; ++d) {
        if (d == 0) {
            // the superblock with itself, every ordered pair, j == i skipped
            gather_blocks<1, kChunkN3, F32>(t, xs, ys, nc, q, box, ib, n, S * kSuper, S * kSuper + kSuper,
                                            e, ax, ay, n_near, n_tested, smin);
            continue;
        }
        const int32_t T = S + d < nsb ? S + d : S + d - nsb;
        uint32_t *fl = rflag + (size_t)T * half + (d - 1);
        Box sbT = empty_box();
#pragma unroll
        for (int k = 0; k < 4; ++k) sbT = box_union(sbT, box + 4 * (T * 4 + k));
        // the pair (S, S + nsb/2) belongs to the lower row only
        if ((d == half && S >= half) || box_gap2(sbS, sbT) > rad.rf2) {
            if (threadIdx.x == 0) *fl = 0u;
            continue;
        }
        // the previous tile's sums have been read by all threads
        __syncthreads();
        double *r = red[wave];
        for (int k = lane; k < 2 * kSuper; k += 64) r[k] = 0.0;
        for (int jb4 = 0; jb4 < 4; ++jb4) {
            const int32_t jb = T * kSuper + jb4 * kBoxBlock;
            const BlockClass cls = classify(ib, box + 4 * (jb / kBoxBlock), ireal, jb + kBoxBlock <= n, rad, e.cut);
            if (cls == kSkip) continue;
            n_tested += 2 * 64 * kBoxBlock;
            for (int32_t j0 = jb; j0 < jb + kBoxBlock; j0 += kChunkN3) {
                double cx[kChunkN3], cy[kChunkN3], cq[kChunkN3];
#pragma unroll
                for (int u = 0; u < kChunkN3; ++u) {
                    cx[u] = xs[j0 + u];
                    cy[u] = ys[j0 + u];
                }
                if (cls != kDense && !chunk_hit<1, kChunkN3>(t, cx, cy, nc, j0)) continue;
#pragma unroll
                for (int u = 0; u < kChunkN3; ++u) cq[u] = q[j0 + u];
                if (cls == kDense)
                    chunk_n3<F32, false>(t, cx, cy, cq, j0, e, n_near, smin, ax[0], ay[0], scr[wave], r,
                                         j0 - T * kSuper);
                else
                    chunk_n3<F32, true>(t, cx, cy, cq, j0, e, n_near, smin, ax[0], ay[0], scr[wave], r,
                                        j0 - T * kSuper);
            }
        }
        __syncthreads();
        // the four waves' sums in a fixed order; the reaction is their negation
        const int jj = threadIdx.x;
        const double rx = ((red[0][2 * jj] + red[1][2 * jj]) + red[2][2 * jj]) + red[3][2 * jj];
        const double ry = ((red[0][2 * jj + 1] + red[1][2 * jj + 1]) + red[2][2 * jj + 1]) + red[3][2 * jj + 1];
        double *rr = react + (size_t)(d - 1) * 2 * npad;
        rr[T * kSuper + jj] = -rx;
        rr[npad + T * kSuper + jj] = -ry;
        if (threadIdx.x == 0) *fl = 1u;
    }
    if (e.dtc) commit_dtmin(smin, &sc->dtmin_bits);
    double *px = part + (size_t)(2 * blockIdx.y) * npad;
    px[ibase] = ax[0];
    px[npad + ibase] = ay[0];
    if (lane == 0) {
        atomicAdd(&sc->pairs_near, (unsigned long long)n_near);
        atomicAdd(&sc->pairs_tested, (unsigned long long)n_tested);
    }
}

}  // namespace

hipError_t launch_pairs_n3(const PairLaunch &a, hipStream_t s) {
    const int64_t nsb = a.npad / kSuper;
    if (a.npad % kPad || !a.box || !a.react || !a.rflag || a.span < 1 || a.n > a.npad)
        return hipErrorInvalidValue;
    if ((int64_t)a.slices * a.span < nsb / 2 + 1) return hipErrorInvalidValue;
    int32_t flags = 0;
    if (a.dtcheck) flags |= kFlagDtCheck;
    if (a.gated) flags |= kFlagGated;
    const dim3 grid((unsigned)nsb, (unsigned)a.slices);
    const ExpCoef ec = exp_taylor_coef();
    if (a.fp32pair)
        hipLaunchKernelGGL(k_pairs_n3<true>, grid, dim3(256), 0, s, a.xs, a.ys, a.nc, a.q, a.orig, a.box,
                           a.part, a.react, a.rflag, a.sc, (int32_t)a.n, (int32_t)a.npad, a.span, flags,
                           a.kappa, a.rc2p, ec);
    else
        hipLaunchKernelGGL(k_pairs_n3<false>, grid, dim3(256), 0, s, a.xs, a.ys, a.nc, a.q, a.orig, a.box,
                           a.part, a.react, a.rflag, a.sc, (int32_t)a.n, (int32_t)a.npad, a.span, flags,
                           a.kappa, a.rc2p, ec);
    return hipGetLastError();
}

}  // namespace demon
