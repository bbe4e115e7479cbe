// pairs_n3.h -- tile code of the Newton-3 all-pairs pass, shared by the
// single-cloud kernel (pairs_n3.hip) and the batched ensemble (batch.hip).
//
// Every unordered pair is evaluated once and applied to both particles, as
// the reference does (ShieldedCoulombForce.cpp:51-56 add the force to one
// particle and subtract it from the other).  Same physics, contribution test
// and dt check as the gather form (pair_core.h); half the pair evaluations.
//
// Work decomposition.  Particles (device order) form superblocks of 256
// (four 64-particle box blocks).  With nsb superblocks, row S pairs with
// superblocks T = S + d (mod nsb) for d = 1 .. floor(nsb/2) (for even nsb,
// d = nsb/2 only for S < nsb/2, so that each unordered superblock pair is
// taken once), plus itself (d = 0, each unordered pair of the superblock
// once: n3_diag).  Every row does the same amount of work.
// A workgroup of four waves owns row S and a set of offsets d: a run of the
// row's live tiles from the tile list (the single-cloud kernel,
// pairs_n3.hip) or a contiguous range of d (a slice; the batched ensemble);
// wave w holds the 64 i-particles of block 4S + w in registers and
// accumulates their forces across its offsets (written once, to the item's
// or the slice's partial sums).
// For each pair tile (S, T) the reaction on every j of T is summed over the
// wave's 64 lanes, then over the four waves in LDS in a fixed order, and
// written to react[d - 1][.][j]; rflag[T nsb/2 + d - 1] says whether the
// tile was evaluated (1) or culled / not owned (0).  Nothing is atomic, so
// the sums are deterministic: the stage kernel adds part[0..slices) and then
// react[d - 1] for d = 1 .. nsb/2 in that order.
#pragma once

#include <type_traits>

#include "pair_core.h"

namespace demon {
namespace n3 {

using namespace pair;

// j particles per chunk: four keeps the chunk's j records (x, y, q) in 24
// SGPRs next to the exp coefficients without spilling.
constexpr int kChunkN3 = 4;

// v from the lane given by the DPP control CTRL (a permutation within each
// group of 8 lanes here, so every source lane exists), as two 32-bit DPP
// moves.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

constexpr int kDppSwap1 = 0xB1;        // quad_perm [1, 0, 3, 2]: lane ^ 1
constexpr int kDppSwap2 = 0x4E;        // quad_perm [2, 3, 0, 1]: lane ^ 2
constexpr int kDppHalfMirror = 0x141;  // row_half_mirror: lane 7 - lane within 8

// Wave sums of the 8 per-lane values v[4 c + u] (component c of the pair
// force on j0 + u), transposed through LDS: every lane writes its 8 values
// into scr[v][lane], then lane (v = l / 8, p = l % 8) adds scr[v][8p .. 8p+7]
// and three DPP butterflies over the 8 lanes of value v (swap 1, swap 2,
// then the mirror, which pairs the two quad sums) finish the sum, which
// lane p = 0 stores into red[2 (jj0 + u) + c] (FP64).  Rows are padded to
// 66 doubles or 68 floats, which keeps both the writes and the 16-byte reads
// at the minimum bank load.  The order of the additions is fixed, so the
// result is deterministic.
template <typename V>
struct Scr;
template <>
struct Scr<double> {
    static constexpr int kRow = 66;
};
template <>
struct Scr<float> {
    static constexpr int kRow = 68;
};

__device__ __forceinline__ double dpp_swap1(double v) { return dpp_f64<kDppSwap1>(v); }
__device__ __forceinline__ double dpp_swap2(double v) { return dpp_f64<kDppSwap2>(v); }
__device__ __forceinline__ double dpp_mirror(double v) { return dpp_f64<kDppHalfMirror>(v); }
__device__ __forceinline__ float dpp_swap1(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), kDppSwap1, 0xF, 0xF, true));
}
__device__ __forceinline__ float dpp_swap2(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), kDppSwap2, 0xF, 0xF, true));
}
__device__ __forceinline__ float dpp_mirror(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), kDppHalfMirror, 0xF, 0xF, true));
}

// The j of a chunk of four: each one's particle index (the dt check), and
// for this lane the slot in the tile's LDS reaction sums of the j whose sum
// the lane stores (u = (lane / 8) mod 4 in store_wave_sums): j minus the
// tile's first particle.
struct ChunkJ {
    int32_t j[kChunkN3];
    int32_t lslot;
};

__device__ __forceinline__ int32_t lane_chunk_u() { return (int32_t)((threadIdx.x >> 3) & 3); }

template <typename V>
__device__ __forceinline__ void store_wave_sums(const V (&f)[2 * kChunkN3], V *scr, double *red, int32_t lslot) {
    constexpr int kRow = Scr<V>::kRow;
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int v = 0; v < 2 * kChunkN3; ++v) scr[v * kRow + lane] = f[v];
    __builtin_amdgcn_wave_barrier();
    const int v = lane >> 3, p = lane & 7;
    const V *row = scr + v * kRow + 8 * p;
    V t = row[0];
#pragma unroll
    for (int k = 1; k < 8; ++k) t += row[k];
    __builtin_amdgcn_wave_barrier();
    t += dpp_swap1(t);
    t += dpp_swap2(t);
    t += dpp_mirror(t);
    if (p == 0) red[2 * lslot + (v >> 2)] = (double)t;
}

// LDS scratch per wave for the transposes: 8 rows of the row length of the
// pair arithmetic's type (FP32 rows take half the space, which lets more
// workgroups share a CU).
constexpr int kScrDoubles = 2 * kChunkN3 * Scr<double>::kRow;
template <bool F32>
struct ScrSel {
    using type = double[4][kScrDoubles];
};
template <>
struct ScrSel<true> {
    using type = float[4][2 * kChunkN3 * Scr<float>::kRow];
};
template <bool F32>
using ScrLds = typename ScrSel<F32>::type;
// one wave's rows, handed down as double * (the chunk code reads them as
// its own type)
template <bool F32>
__device__ __forceinline__ double *scr_rows(ScrLds<F32> &scr, int wave) {
    return reinterpret_cast<double *>(&scr[wave][0]);
}

// Newton-3 form over one chunk of 4 j: adds F_ij into ax/ay and stores the
// wave's sum of F_ij for each j into red[2 jj], red[2 jj + 1] (the
// reaction, negated later).  CHECK: the chunk may hold pairs outside the
// cutoff (counted into n_near here; a dense block's count is added by the
// caller).  DTC: the stage-1 launch, with the dt check.  FP32 pair
// arithmetic keeps the chunk's values in FP32 (the transposed sums of 64
// terms, and the 4-term i-side sum added to ax/ay in FP64).
template <bool F32>
using PairVal = typename std::conditional<F32, float, double>::type;

template <bool F32, bool CHECK, bool DTC, bool TAB = false>
__device__ __forceinline__ void chunk_n3(const ITile<1> &t, const double (&cx)[kChunkN3],
                                         const double (&cy)[kChunkN3], const PairVal<F32> (&cq)[kChunkN3],
                                         const ChunkJ &cj, const PairEnv &e, uint64_t &n_near, float &smin,
                                         double &ax, double &ay, double *scr, double *red, bool dtc_blk) {
    using V = PairVal<F32>;
    V f[2 * kChunkN3];
    uint64_t hits = 0;  // lanes with some pair of the chunk inside the cutoff (CHECK)
#pragma unroll
    for (int u = 0; u < kChunkN3; ++u) {
        const double dx = t.xi[0] - cx[u];
        const double dy = t.yi[0] - cy[u];
        const double r2 = fma(dx, dx, dy * dy);
        const bool in = pair_in<CHECK>(dx, dy, r2, e);
        // a dense block has 64 real i and every pair inside the cutoff
        if (CHECK) {
            const uint64_t b = __ballot(in);
            n_near += 2 * (uint64_t)__builtin_popcountll(b);
            hits |= b;
        }
        V fx = 0, fy = 0;
        if (in) {
            if constexpr (F32) {
                // without q_j, as the FP64 form below
                pair_force_fp32f_noq(dx, dy, t.cqf[0], e.kl2_f, (float)e.cut.kappa, fx, fy);
            } else {
                // the force without q_j: q_j multiplies the i-side term here
                // and the j-side sums once per j (n3_tile, n3_diag)
                const double s = pair_scale_noq<TAB>(r2, t.cqi[0], e);
                fx = s * dx;
                fy = s * dy;
                ax = fma(cq[u], fx, ax);
                ay = fma(cq[u], fy, ay);
            }
        }
        f[u] = fx;
        f[kChunkN3 + u] = fy;
        if (DTC && dtc_blk) pair_dtcheck(t.xi[0], t.yi[0], cx[u], cy[u], r2, t.ibase, cj.j[u], e, smin);
    }
    // No pair of the chunk inside the cutoff: every f is zero, and the
    // wave's sums of this chunk's j keep the zeros the tile stored in red
    // (each chunk's slots are written once per tile).  hits is uniform.
    if (CHECK && hits == 0) return;
    if constexpr (F32) {
        float sx = cq[0] * f[0], sy = cq[0] * f[kChunkN3];
#pragma unroll
        for (int u = 1; u < kChunkN3; ++u) {
            sx = fmaf(cq[u], f[u], sx);
            sy = fmaf(cq[u], f[kChunkN3 + u], sy);
        }
        ax = ax + (double)sx;
        ay = ay + (double)sy;
    }
    store_wave_sums<V>(f, reinterpret_cast<V *>(scr), red, cj.lslot);
}

// One cloud's arrays and launch constants for the row function.
struct N3Row {
    const double *xs, *ys, *nc, *q;
    const float *qf;  // q in FP32 (FP32 pair arithmetic)
    const int32_t *orig;
    const double *box;
    double *part;     // [slices][2][npad]
    double *react;    // [nsb/2][2][npad]
    uint32_t *rflag;  // [nsb][nsb/2]
    Scalars *sc;
    int32_t n, npad, flags;
    double kappa, rc2p;
};

using RedLds = double[4][2 * kSuper];  // [wave][2 jj + component]
// Bounding boxes of the superblocks T a row visits, staged in LDS a chunk at
// a time: one round trip to memory per chunk instead of one per offset,
// which is what a launch of mostly culled tiles waits on.
constexpr int kBoxChunk = 32;
using TBoxLds = Box[kBoxChunk];

// The wave's i block against one 64-j block (first particle jb) in the
// Newton-3 form: box classification, then the straight-line path (dense) or
// the prefiltered one (mixed); i-side forces into ax/ay, the wave's sums of
// the forces on each j into red[2 (j - t0) + c].
template <bool F32, bool DTC, bool TAB>
__device__ __forceinline__ void n3_jblock(const N3Row &a, const PairEnv &e, const ITile<1> &t, const Box &ib,
                                          bool ireal, int32_t jb, int32_t t0, double &ax, double &ay,
                                          uint64_t &n_near, uint64_t &n_tested, float &smin, double *scr,
                                          double *red) {
    const double *__restrict__ xs = a.xs;
    const double *__restrict__ ys = a.ys;
    const double *__restrict__ nc = a.nc;
    const double *__restrict__ q = a.q;
    const float *__restrict__ qf = a.qf;
    const double *jbx = a.box + 4 * (jb / kBoxBlock);
    const BlockClass cls = classify(ib, jbx, ireal, jb + kBoxBlock <= a.n, e.rad, e.cut);
    if (cls == kSkip) return;
    // the stage-1 dt check only where the boxes come within its radius (no
    // pair of blocks farther apart can be that close)
    const bool dtc_blk = DTC && box_gap2(ib, Box{jbx[0], jbx[1], jbx[2], jbx[3]}) <= e.rad.dt2;
    if (cls == kDense) {
        n_tested += 2 * 64 * kBoxBlock;
        n_near += 2 * 64 * kBoxBlock;
        for (int32_t j0 = jb; j0 < jb + kBoxBlock; j0 += kChunkN3) {
            double cx[kChunkN3], cy[kChunkN3];
            PairVal<F32> cq[kChunkN3];
#pragma unroll
            for (int u = 0; u < kChunkN3; ++u) {
                cx[u] = ld_uniform(xs + j0 + u);
                cy[u] = ld_uniform(ys + j0 + u);
                if constexpr (F32) cq[u] = ld_uniform(qf + j0 + u);
                else cq[u] = ld_uniform(q + j0 + u);
            }
            const ChunkJ cj{{j0, j0 + 1, j0 + 2, j0 + 3}, j0 - t0 + lane_chunk_u()};
            chunk_n3<F32, false, DTC, TAB>(t, cx, cy, cq, cj, e, n_near, smin, ax, ay, scr, red, dtc_blk);
        }
        return;
    }
    n_tested += 2 * 64 * kBoxBlock;
    for (int32_t j0 = jb; j0 < jb + kBoxBlock; j0 += kChunkN3) {
        double cx[kChunkN3], cy[kChunkN3];
        PairVal<F32> cq[kChunkN3];
#pragma unroll
        for (int u = 0; u < kChunkN3; ++u) {
            cx[u] = ld_uniform(xs + j0 + u);
            cy[u] = ld_uniform(ys + j0 + u);
        }
        // the norms and charges in the same batch of scalar loads: one
        // round trip to memory per chunk, which is what a sparse launch's
        // chunks wait on
        double cn[kChunkN3];
#pragma unroll
        for (int u = 0; u < kChunkN3; ++u) {
            cn[u] = ld_uniform(nc + j0 + u);
            if constexpr (F32) cq[u] = ld_uniform(qf + j0 + u);
            else cq[u] = ld_uniform(q + j0 + u);
        }
        // an empty statement that takes all sixteen values: the compiler
        // must have every load back here, so it issues them together rather
        // than sinking the charges behind the prefilter's branch
        asm volatile("" ::"s"(cx[0]), "s"(cx[1]), "s"(cx[2]), "s"(cx[3]), "s"(cy[0]), "s"(cy[1]), "s"(cy[2]),
                     "s"(cy[3]), "s"(cn[0]), "s"(cn[1]), "s"(cn[2]), "s"(cn[3]), "s"(cq[0]), "s"(cq[1]), "s"(cq[2]),
                     "s"(cq[3]));
        uint64_t hit = 0;
#pragma unroll
        for (int u = 0; u < kChunkN3; ++u) {
            const double w = fma(t.xm[0], cx[u], fma(t.ym[0], cy[u], t.nT[0]));
            hit |= __ballot(w < cn[u]);
        }
        if (!hit) continue;
        const ChunkJ cj{{j0, j0 + 1, j0 + 2, j0 + 3}, j0 - t0 + lane_chunk_u()};
        chunk_n3<F32, true, DTC, TAB>(t, cx, cy, cq, cj, e, n_near, smin, ax, ay, scr, red, dtc_blk);
    }
}

// Work a caller may put where a tile's waves wait for each other (just
// before the barrier that ends the tile's pair work): called by every wave
// as it arrives there.  The default does nothing.
struct NoHook {
    __device__ __forceinline__ void operator()() const {}
};

// Tile (S, T): the four 64-j blocks of superblock T against the wave's i
// block; adds the i-side forces into ax/ay and writes the reactions on T's
// 256 particles to react[d - 1] (negated sums over the four waves in a fixed
// order), or, given jdst, to jdst[jj] and jdst[jstr + jj].  Called by all
// 256 threads; ib: the wave's i-block box.
template <bool F32, bool DTC, bool TAB = false, class HOOK = NoHook>
__device__ __forceinline__ void n3_tile(const N3Row &a, const PairEnv &e, const ITile<1> &t, const Box &ib,
                                        bool ireal, int32_t T, int32_t d, double &ax, double &ay,
                                        uint64_t &n_near, uint64_t &n_tested, float &smin, RedLds &red,
                                        ScrLds<F32> &scr, const HOOK &hook = HOOK(), double *jdst = nullptr,
                                        int64_t jstr = 0) {
    const int32_t npad = a.npad;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // the previous tile's sums have been read by all threads
    __syncthreads();
    double *r = red[wave];
    for (int k = lane; k < 2 * kSuper; k += 64) r[k] = 0.0;
    for (int jb4 = 0; jb4 < 4; ++jb4)
        n3_jblock<F32, DTC, TAB>(a, e, t, ib, ireal, T * kSuper + jb4 * kBoxBlock, T * kSuper, ax, ay, n_near,
                                 n_tested, smin, scr_rows<F32>(scr, wave), r);
    hook();
    __syncthreads();
    // the four waves' sums in a fixed order; the reaction is their negation
    // times q_j (the pair sums leave it out; FP32 pairs: the rounded charge)
    const int jj = threadIdx.x;
    double rx = ((red[0][2 * jj] + red[1][2 * jj]) + red[2][2 * jj]) + red[3][2 * jj];
    double ry = ((red[0][2 * jj + 1] + red[1][2 * jj + 1]) + red[2][2 * jj + 1]) + red[3][2 * jj + 1];
    {
        const double qj = F32 ? (double)a.qf[T * kSuper + jj] : a.q[T * kSuper + jj];
        rx = rx * qj;
        ry = ry * qj;
    }
    if (jdst) {
        jdst[jj] = -rx;
        jdst[jstr + jj] = -ry;
        return;
    }
    double *rr = a.react + (size_t)(d - 1) * 2 * npad;
    rr[T * kSuper + jj] = -rx;
    rr[npad + T * kSuper + jj] = -ry;
}

// The superblock S with itself in the Newton-3 form.  Each wave takes its
// own block in the gather form (64 x 63 ordered pairs) and then the block
// pairs (w, w + 1 mod 4) and, for waves 0 and 1, (w, w + 2): each of the six
// block pairs once.  The reactions meet in LDS and go to the thread that
// holds the particle as its i (thread jj holds particle 256 S + jj), in a
// fixed order.  Called by all 256 threads.
template <bool F32, bool DTC, bool TAB, class HOOK = NoHook>
__device__ __forceinline__ void n3_diag(const N3Row &a, const PairEnv &e, const ITile<1> &t, const Box &ib,
                                        bool ireal, int32_t S, double &ax, double &ay, uint64_t &n_near,
                                        uint64_t &n_tested, float &smin, RedLds &red, ScrLds<F32> &scr,
                                        const HOOK &hook = HOOK()) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // the previous tile's sums have been read by all threads
    __syncthreads();
    double *r = red[wave];
    for (int k = lane; k < 2 * kSuper; k += 64) r[k] = 0.0;
    const int32_t own = S * kSuper + wave * kBoxBlock;
    double gx[1] = {0.0}, gy[1] = {0.0};
    gather_blocks<1, kChunkN3, F32, TAB>(t, a.xs, a.ys, a.nc, a.q, a.box, ib, a.n, own, own + kBoxBlock, e, gx, gy,
                                         n_near, n_tested, smin);
    ax = ax + gx[0];
    ay = ay + gy[0];
    n3_jblock<F32, DTC, TAB>(a, e, t, ib, ireal, S * kSuper + ((wave + 1) & 3) * kBoxBlock, S * kSuper, ax, ay,
                             n_near, n_tested, smin, scr_rows<F32>(scr, wave), r);
    if (wave < 2)
        n3_jblock<F32, DTC, TAB>(a, e, t, ib, ireal, S * kSuper + (wave + 2) * kBoxBlock, S * kSuper, ax, ay,
                                 n_near, n_tested, smin, scr_rows<F32>(scr, wave), r);
    hook();
    __syncthreads();
    const int jj = threadIdx.x;
    double rx = ((red[0][2 * jj] + red[1][2 * jj]) + red[2][2 * jj]) + red[3][2 * jj];
    double ry = ((red[0][2 * jj + 1] + red[1][2 * jj + 1]) + red[2][2 * jj + 1]) + red[3][2 * jj + 1];
    {
        const double qj = F32 ? (double)a.qf[S * kSuper + jj] : a.q[S * kSuper + jj];
        rx = rx * qj;
        ry = ry * qj;
    }
    ax = ax - rx;
    ay = ay - ry;
}

// The pair counts of a workgroup's waves, added with one atomic each to one
// of kCountSlots counter addresses (device-wide atomics on one address
// serialise).  Uses red as scratch; called by all 256 threads.
__device__ __forceinline__ void n3_commit_counts(Scalars *sc, int32_t slot_index, uint64_t n_near,
                                                 uint64_t n_tested, RedLds &red) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    __syncthreads();
    unsigned long long *cnt = reinterpret_cast<unsigned long long *>(&red[0][0]);
    if (lane == 0) {
        cnt[2 * wave] = n_near;
        cnt[2 * wave + 1] = n_tested;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long *slot = sc->pair_counts[slot_index & (kCountSlots - 1)];
        atomicAdd(&slot[0], ((cnt[0] + cnt[2]) + cnt[4]) + cnt[6]);
        atomicAdd(&slot[1], ((cnt[1] + cnt[3]) + cnt[5]) + cnt[7]);
    }
    // red is reused by the caller
    __syncthreads();
}

// One workgroup's share in the batched ensemble: row S against superblocks
// S + d for d in [d0, d1), its i-side sums written to part[slice].  Called by
// all 256 threads.
// etab: the exp table e^{-k/64} in LDS (FP64 pair arithmetic), with ec its
// coefficients (exp_tab_coef).
template <bool F32, bool DTC>
__device__ __forceinline__ void n3_row(const N3Row &a, const ExpCoef &ec, const double *etab, int32_t S, int32_t d0,
                                       int32_t d1, int32_t slice, RedLds &red, ScrLds<F32> &scr, TBoxLds &tbox) {
    const double *__restrict__ box = a.box;
    const int32_t n = a.n, npad = a.npad;
    const Radii rad = launch_radii(a.rc2p, a.flags, a.sc->rmax2_bits);
    const PairEnv e{make_cutoff(a.kappa, a.rc2p), rad, fp32_kl2(a.kappa), DTC, a.orig, ec, etab};
    float smin = __builtin_inff();
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int32_t nsb = npad / kSuper;
    const int32_t half = nsb / 2;
    const int32_t ibase = S * kSuper + wave * 64 + lane;
    const ITile<1> t = make_itile<1>(a.xs, a.ys, a.nc, a.q, n, ibase, rad.rf2);
    const Box ib = box_union(empty_box(), box + 4 * (t.wlo / kBoxBlock));
    Box sbS = empty_box();
#pragma unroll
    for (int k = 0; k < 4; ++k) sbS = box_union(sbS, box + 4 * (S * 4 + k));
    const bool ireal = t.whi <= n;
    double ax[1] = {0.0}, ay[1] = {0.0};
    uint64_t n_near = 0, n_tested = 0;
    int32_t tb0 = 0;  // offset d of tbox[0]
    for (int32_t d = d0; d < d1; ++d) {
        if (d == 0) {
            // the superblock with itself, each unordered pair once
            n3_diag<F32, DTC, true>(a, e, t, ib, ireal, S, ax[0], ay[0], n_near, n_tested, smin, red, scr);
            continue;
        }
        if ((d - d0) % kBoxChunk == 0 || d == 1) {
            // stage the boxes of T = S + d .. S + min(d1, d + kBoxChunk) - 1;
            // the barriers are uniform (d is), the first one lets the
            // previous chunk's readers finish
            const int32_t dc = d, dn = d1 < d + kBoxChunk ? d1 : d + kBoxChunk;
            __syncthreads();
            for (int32_t k = (int32_t)threadIdx.x; k < dn - dc; k += 256) {
                const int32_t Tk = S + dc + k < nsb ? S + dc + k : S + dc + k - nsb;
                Box b = empty_box();
#pragma unroll
                for (int m = 0; m < 4; ++m) b = box_union(b, box + 4 * (Tk * 4 + m));
                tbox[k] = b;
            }
            __syncthreads();
            tb0 = dc;
        }
        const int32_t T = S + d < nsb ? S + d : S + d - nsb;
        uint32_t *fl = a.rflag + (size_t)T * half + (d - 1);
        const Box sbT = tbox[d - tb0];
        // for even nsb the pair (S, S + nsb/2) is reached from both rows and
        // belongs to the lower one; for odd nsb every offset is distinct
        if ((d == half && !(nsb & 1) && S >= half) || box_gap2(sbS, sbT) > rad.rf2) {
            if (threadIdx.x == 0) *fl = 0u;
            continue;
        }
        n3_tile<F32, DTC, true>(a, e, t, ib, ireal, T, d, ax[0], ay[0], n_near, n_tested, smin, red, scr);
        if (threadIdx.x == 0) *fl = 1u;
    }
    if (DTC) commit_dtmin(smin, &a.sc->dtmin_bits);
    double *px = a.part + (size_t)(2 * slice) * npad;
    px[ibase] = ax[0];
    px[npad + ibase] = ay[0];
    n3_commit_counts(a.sc, S + slice, n_near, n_tested, red);
}

}  // namespace n3
}  // namespace demon
