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
#include "pair_core.h"

namespace demon {

namespace {

using namespace pair;

template <int R, bool F32>
__global__ __launch_bounds__(256) void k_pairs(const double *__restrict__ xs,
                                               const double *__restrict__ ys,
                                               const double *__restrict__ nc,
                                               const double *__restrict__ q,
                                               const int32_t *__restrict__ orig,
                                               const double *__restrict__ box,
                                               double *__restrict__ part,
                                               Scalars *__restrict__ sc, int64_t n, int64_t npad,
                                               int32_t jslice, int32_t flags, double kappa,
                                               double rc2p, ExpCoef ec) {
    if ((flags & kFlagGated) && !sc->active) return;
    const Radii rad = launch_radii(rc2p, flags, sc->rmax2_bits);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int32_t ibase = (int32_t)blockIdx.x * (256 * R) + wave * (64 * R) + lane;
    const int32_t j_begin = (int32_t)blockIdx.y * jslice;
    const int32_t j_end = j_begin + jslice < (int32_t)npad ? j_begin + jslice : (int32_t)npad;
    double ax[R], ay[R];
#pragma unroll
    for (int r = 0; r < R; ++r) ax[r] = ay[r] = 0.0;
    uint64_t n_near = 0, n_tested = 0;
    wave_tile<R, F32>(xs, ys, nc, q, orig, box, (int32_t)n, ibase, j_begin, j_end, flags, kappa,
                      rc2p, rad, ec, &sc->dtmin_bits, ax, ay, n_near, n_tested);
    // one atomic per counter per workgroup, to one of kCountSlots addresses
    // (device-wide atomics on one address serialise); every thread reaches
    // this point
    __shared__ unsigned long long cnt[4][2];
    if (lane == 0) {
        cnt[wave][0] = n_near;
        cnt[wave][1] = n_tested;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long *slot = sc->pair_counts[(blockIdx.x + blockIdx.y) & (kCountSlots - 1)];
        atomicAdd(&slot[0], ((cnt[0][0] + cnt[1][0]) + cnt[2][0]) + cnt[3][0]);
        atomicAdd(&slot[1], ((cnt[0][1] + cnt[1][1]) + cnt[2][1]) + cnt[3][1]);
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
    if (a.gated) flags |= kFlagGated;
    hipLaunchKernelGGL((k_pairs<R, F32>), grid, dim3(256), 0, s, a.xs, a.ys, a.nc, a.q, a.orig,
                       a.box, a.part, a.sc, a.n, a.npad, a.jslice, flags, a.kappa, a.rc2p,
                       exp_coef());
    return hipGetLastError();
}

}  // namespace

hipError_t launch_pairs(const PairLaunch &a, hipStream_t s) {
    // a slice is whole box blocks or a quarter or half of one (never two
    // blocks' parts: the tile tests one block's box for each 64 j)
    const bool js_ok = a.jslice > 0 && (a.jslice % kBoxBlock == 0 || a.jslice == kBoxBlock / 4 ||
                                        a.jslice == kBoxBlock / 2);
    if (a.npad % kPad || !js_ok || a.slices < 1 || a.n > a.npad) return hipErrorInvalidValue;
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
