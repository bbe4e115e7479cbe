// pairs_n3.hip -- the single-cloud Newton-3 all-pairs pass (the tile code is
// in pairs_n3.h).  Three launches per force evaluation:
//
//  k_n3_tiles   one workgroup per superblock row S: tests the bounding box of
//               S against the box of every superblock T = S + d it owns
//               (d = 1 .. nsb/2), and then S's four 64-particle block boxes
//               against T's (a tile is live when some block pair may hold a
//               pair within reach), writes the tile flags rflag[T][d - 1] and
//               the row's live offsets, in increasing d (d = 0, the row with
//               itself, first), into list[S].
//  k_n3_plan    one workgroup: cuts the rows into work items of `span`
//               consecutive live tiles, span = ceil(live tiles / target), and
//               writes the first item of each row.
//  k_pairs_n3   a persistent grid: every workgroup takes items from atomic
//               counters until none is left; one item = the row's i
//               particles in registers against its span of live tiles, the
//               i-side sums written once, to part[item].
//
// Only live tiles cost anything: on the benchmark lattice about 1% of the
// superblock pairs are within the cutoff, and a launch takes as long as
// those take.  The items of rows k nsb/8 .. (k+1) nsb/8 - 1 go to XCD k,
// whose workgroups are those with blockIdx.x = k mod 8, which keeps the j
// superblocks an XCD visits to 5/8 of the cloud, within its own L2.
#include "pairs_n3.h"

namespace demon {

namespace {

using namespace pair;
using namespace n3;

__device__ __forceinline__ Box superblock_box(const double *__restrict__ box, int32_t S) {
    Box b = empty_box();
#pragma unroll
    for (int k = 0; k < 4; ++k) b = box_union(b, box + 4 * (S * 4 + k));
    return b;
}

// Block-wide exclusive prefix sum of one value per thread (256 threads);
// *total receives the sum.  A scan within each wave (six shuffle steps),
// then the four wave totals through LDS: two barriers.  Ends with a barrier
// (lds may be reused at once).
__device__ __forceinline__ int32_t block_exclusive_scan(int32_t v, int32_t *lds, int32_t *total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int32_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    int32_t base = 0;
    for (int w = 0; w < wave; ++w) base += lds[w];
    *total = lds[0] + lds[1] + lds[2] + lds[3];
    __syncthreads();
    return base + x - v;
}

__global__ __launch_bounds__(256) void k_n3_tiles(N3Row a, N3List l) {
    __shared__ int32_t wcnt[4];
    __shared__ int32_t base_s;
    if ((a.flags & kFlagGated) && !a.sc->active) return;
    const int32_t nsb = a.npad / kSuper, half = nsb / 2;
    const int32_t S = (int32_t)blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const Radii rad = launch_radii(a.rc2p, a.flags, a.sc->rmax2_bits);
    const Box sbS = superblock_box(a.box, S);
    const bool sreal = sbS.x0 <= sbS.x1;  // the row holds real particles
    Box ibS[4];  // the row's four 64-particle blocks
#pragma unroll
    for (int k = 0; k < 4; ++k) ibS[k] = box_union(empty_box(), a.box + 4 * (S * 4 + k));
    int32_t *lst = l.list + (size_t)S * (half + 1);
    if (threadIdx.x == 0) {
        base_s = sreal ? 1 : 0;
        if (sreal) lst[0] = 0;
    }
    __syncthreads();
    for (int32_t d0 = 1; d0 <= half; d0 += 256) {
        const int32_t d = d0 + (int32_t)threadIdx.x;
        bool live = false, swap = false;
        if (d <= half) {
            const int32_t T = S + d < nsb ? S + d : S + d - nsb;
            // for even nsb the pair (S, S + nsb/2) is reached from both rows
            // and belongs to the lower one
            const bool owned = !(d == half && !(nsb & 1) && S >= half);
            if (owned && sreal && box_gap2(sbS, superblock_box(a.box, T)) <= rad.rf2) {
                // live when some block pair is within reach: otherwise the
                // pair kernel would skip all sixteen (the same box test)
                int32_t rowc[4] = {0, 0, 0, 0}, colmax = 0;
#pragma unroll
                for (int kj = 0; kj < 4; ++kj) {
                    const Box jb = box_union(empty_box(), a.box + 4 * (T * 4 + kj));
                    int32_t colc = 0;
#pragma unroll
                    for (int ki = 0; ki < 4; ++ki) {
                        const int32_t m = box_gap2(ibS[ki], jb) <= rad.rf2 ? 1 : 0;
                        rowc[ki] += m;
                        colc += m;
                    }
                    colmax = max(colmax, colc);
                }
                const int32_t rowmax = max(max(rowc[0], rowc[1]), max(rowc[2], rowc[3]));
                live = rowmax > 0;
                // a tile takes as long as its busiest wave, and wave w takes
                // block w of the i-side superblock: T's blocks are the i side
                // when the busiest of them has fewer live block pairs than
                // the busiest of S's
                swap = colmax < rowmax;
            }
            a.rflag[(size_t)T * half + (d - 1)] = live ? 1u : 0u;
        }
        const unsigned long long m = __ballot(live);
        if (lane == 0) wcnt[wave] = __builtin_popcountll(m);
        __syncthreads();
        int32_t off = base_s;
        for (int w = 0; w < wave; ++w) off += wcnt[w];
        if (live) {
            const int32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            lst[off + below] = d | (swap ? kN3Swap : 0);
        }
        __syncthreads();
        if (threadIdx.x == 0) base_s += wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
        __syncthreads();
    }
    if (threadIdx.x == 0) l.cnt[S] = base_s;
}

// One workgroup: cuts the rows' live tiles into items (span = ceil(live /
// target) consecutive tiles of one row), writes the first item of each row
// (first[nsb] = items of the launch) and resets the pair kernel's queue
// counters.
__global__ __launch_bounds__(256) void k_n3_plan(N3Row a, N3List l) {
    constexpr int kRowsPer = 2 * kMaxHalf / 256;  // rows per thread at the largest nsb
    __shared__ int32_t scan[4];
    if ((a.flags & kFlagGated) && !a.sc->active) return;
    const int32_t nsb = a.npad / kSuper;
    if (threadIdx.x < kN3Queues) l.plan->next[threadIdx.x] = 0u;
    const int32_t per = (nsb + 255) / 256;
    const int32_t r0 = min((int32_t)threadIdx.x * per, nsb);
    int32_t c[kRowsPer];
    int32_t tiles = 0, rows = 0;
#pragma unroll
    for (int k = 0; k < kRowsPer; ++k) {
        c[k] = k < per && r0 + k < nsb ? l.cnt[r0 + k] : 0;
        tiles += c[k];
        rows += c[k] > 0;
    }
    int32_t total = 0, real_rows = 0;
    (void)block_exclusive_scan(tiles, scan, &total);
    (void)block_exclusive_scan(rows, scan, &real_rows);
    const int32_t span = max(1, (total + l.target - 1) / l.target);
    int32_t items = 0;
#pragma unroll
    for (int k = 0; k < kRowsPer; ++k) {
        c[k] = (c[k] + span - 1) / span;
        items += c[k];
    }
    int32_t nitems = 0;
    int32_t f = block_exclusive_scan(items, scan, &nitems);
#pragma unroll
    for (int k = 0; k < kRowsPer; ++k) {
        if (k < per && r0 + k < nsb) l.first[r0 + k] = f;
        f += c[k];
    }
    if (threadIdx.x == 0) {
        l.first[nsb] = nitems;
        l.plan->span = span;
        l.plan->nitems = nitems;
        // the diagonal tile of every row with real particles is not counted
        a.sc->n3_tiles += (unsigned long long)(total - real_rows);
        a.sc->n3_items += (unsigned long long)nitems;
    }
}

// The next item of queue qk (rows qr0 .. qr0 + nr - 1, items i0 .. i1 - 1):
// one atomic claims a slot.  Slots go to items in row order except that the
// first half of the last row's items come first: a row's items start with
// its costliest tiles (the superblock with itself, then the nearest ones),
// and the last row's were among the last items a launch waited on (0.6%
// faster on the imploded cloud, no change on the lattice, where the last
// items are also those taken from other queues).  The item's row is the last
// of the queue's rows (one per lane, t = nr: the queue's end) whose first
// item is at most it (first[] does not decrease).  Run by one whole wave; lane 0 leaves
// (it, S, k0, k1, e) in nx (it >= i1: the queue is exhausted), e the list
// entry of an item of one swapped tile, else 0.
struct QueueAt {
    int32_t qk, qr0, nr, i0, i1, span;
    int32_t il;  // the first item of the queue's last row
};

struct QueueArrays {
    const int32_t *first, *cnt, *list;
    uint32_t *next;
    int32_t stride;  // of list: nsb/2 + 1
};

__device__ __forceinline__ void claim_item(const QueueArrays &qa, const QueueAt &q, int32_t *nx) {
    const int lane = threadIdx.x & 63;
    uint32_t s = 0u;
    if (lane == 0 && q.i0 < q.i1) s = atomicAdd(&qa.next[q.qk], 1u);
    const int32_t sl = (int32_t)__builtin_amdgcn_readfirstlane(s);
    const int32_t h = (q.i1 - q.il + 1) >> 1, mid = q.il - q.i0;
    const int32_t it = sl < h ? q.il + sl : sl - h < mid ? q.i0 + (sl - h) : q.il + (sl - mid);
    int32_t S = 0, k0 = 0, k1 = 0, ent = 0;
    if (it < q.i1) {
        const int32_t fl = lane <= q.nr ? qa.first[q.qr0 + lane] : INT32_MAX;
        const int32_t cl = lane < q.nr ? qa.cnt[q.qr0 + lane] : 0;
        const int32_t ri = __builtin_popcountll(__ballot(fl <= it)) - 1;
        S = q.qr0 + ri;
        k0 = (it - __builtin_amdgcn_readlane(fl, ri)) * q.span;
        const int32_t nl = __builtin_amdgcn_readlane(cl, ri);
        k1 = k0 + q.span < nl ? k0 + q.span : nl;
        if (k1 == k0 + 1) {
            ent = __builtin_amdgcn_readfirstlane(qa.list[(size_t)S * qa.stride + k0]);
            if (!(ent & kN3Swap)) ent = 0;
        }
    }
    if (lane == 0) {
        nx[0] = it;
        nx[1] = S;
        nx[2] = k0;
        nx[3] = k1;
        nx[4] = ent;
    }
}

// On an item's last tile, the first wave to finish its share claims the next
// item while the others are still at work (the tile's shares differ: the
// wait at the tile's barrier is where this costs nothing).  Used by the
// FP32-pair kernel, where it made a lattice launch 5% faster; in the FP64
// kernel, whose stage-1 variant is at its register limit, it cost 1%, and
// the FP64 kernel claims at the item boundary instead (ClaimAtBoundary).
struct ClaimAtBarrier {
    QueueArrays qa;
    QueueAt q;
    int32_t *flag, *nx;
    bool last;
    __device__ __forceinline__ void operator()() const {
        if (!last) return;
        int32_t first = 0;
        if ((threadIdx.x & 63) == 0) first = atomicAdd(flag, 1) == 0 ? 1 : 0;
        if (__builtin_amdgcn_readfirstlane(first)) claim_item(qa, q, nx);
    }
};

struct ClaimAtBoundary {
    bool last;
    __device__ __forceinline__ void operator()() const {}
};

// One item: row S's i particles (wave w: block 4 S + w) against the live
// tiles lst[k0 .. k1); the i-side sums go to part[it].  claim runs at the
// barrier of each tile's waves, with claim.last set on the last tile.
// swp: the item's one tile's list entry when it is swapped (else 0): then
// T's particles are the i side, their sums go to the tile's reaction row
// react[d - 1] and S's to part[it], where the stage kernel reads each
// superblock's share in either orientation.
template <bool F32, bool DTC, class CLAIM>
__device__ __forceinline__ void n3_item(const N3Row &a, const PairEnv &e, int32_t S, const int32_t *lst,
                                        int32_t k0, int32_t k1, int32_t it, int32_t swp, uint64_t &n_near,
                                        uint64_t &n_tested, float &smin, RedLds &red, ScrLds<F32> &scr,
                                        CLAIM claim) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int32_t n = a.n, nsb = a.npad / kSuper;
    const int32_t dsw = swp & kN3Offset;
    const int32_t isb = swp ? (S + dsw < nsb ? S + dsw : S + dsw - nsb) : S;
    const int32_t ibase = isb * kSuper + wave * 64 + lane;
    const ITile<1> t = make_itile<1>(a.xs, a.ys, a.nc, a.q, n, ibase, e.rad.rf2);
    const Box ib = box_union(empty_box(), a.box + 4 * (t.wlo / kBoxBlock));
    const bool ireal = t.whi <= n;
    double ax[1] = {0.0}, ay[1] = {0.0};
    if (swp) {
        claim.last = true;
        n3_tile<F32, DTC, true>(a, e, t, ib, ireal, S, dsw, ax[0], ay[0], n_near, n_tested, smin, red, scr, claim,
                                a.part + (size_t)it * (2 * kSuper), kSuper);
        double *rr = a.react + (size_t)(dsw - 1) * 2 * a.npad + (size_t)isb * kSuper;
        rr[wave * 64 + lane] = ax[0];
        rr[a.npad + wave * 64 + lane] = ay[0];
        return;
    }
    for (int32_t k = k0; k < k1; ++k) {
        const int32_t d = __builtin_amdgcn_readfirstlane(lst[k]) & kN3Offset;
        claim.last = k + 1 == k1;
        if (d == 0) {
            // the superblock with itself, each unordered pair once
            n3_diag<F32, DTC, true>(a, e, t, ib, ireal, S, ax[0], ay[0], n_near, n_tested, smin, red, scr, claim);
            continue;
        }
        const int32_t T = S + d < nsb ? S + d : S + d - nsb;
        n3_tile<F32, DTC, true>(a, e, t, ib, ireal, T, d, ax[0], ay[0], n_near, n_tested, smin, red, scr, claim);
    }
    double *px = a.part + (size_t)it * (2 * kSuper);
    px[wave * 64 + lane] = ax[0];
    px[kSuper + wave * 64 + lane] = ay[0];
}

template <bool F32, bool DTC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(F32 ? 6 : 4))) void k_pairs_n3(N3Row a, N3List l,
                                                                                   ExpCoef ec) {
    __shared__ RedLds red;
    __shared__ __align__(16) ScrLds<F32> scr;
    __shared__ double etab[kExpTabN];
    if ((a.flags & kFlagGated) && !a.sc->active) return;
    // The FP64 exp is exp_neg_tab on this table.  The first tile's barrier
    // orders these stores before any read.
    if (!F32)
        for (int k = (int)threadIdx.x; k < kExpTabN; k += 256) etab[k] = l.exptab[k];
    const int32_t nsb = a.npad / kSuper, half = nsb / 2;
    const Radii rad = launch_radii(a.rc2p, a.flags, a.sc->rmax2_bits);
    const PairEnv e{make_cutoff(a.kappa, a.rc2p), rad, fp32_kl2(a.kappa), DTC, a.orig, ec, etab};
    const int32_t span = __builtin_amdgcn_readfirstlane(l.plan->span);
    float smin = __builtin_inff();
    uint64_t n_near = 0, n_tested = 0;
    // Work queues: workgroup g runs on XCD k = g mod 8 and starts with queue
    // 8 k + s, s = (g / 8) mod 8.  One item per atomic; sixty-four counters,
    // since the requests to one counter serialise.  FP64: wave 0 claims
    // each item at the item boundary.  FP32 pairs: wave 0 claims a queue's first
    // item, the first wave to finish its share of the current item's last
    // tile every later one (the next item, its row and tile range wait in
    // nx).  When its queue is exhausted, the
    // workgroup reads all sixty-four counters at once (one per lane of wave
    // 0) and moves to the first queue with items left, in the order: the
    // other queues of its XCD, then those of the other XCDs.
    __shared__ int32_t slot;
    __shared__ int32_t nx[5], claim_flag;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int32_t home = (int32_t)(blockIdx.x & 7), sub = (int32_t)((blockIdx.x >> 3) & 7);
    int32_t qk = 8 * home + sub;
    const QueueArrays qarr{l.first, l.cnt, l.list, l.plan->next, half + 1};
    for (;;) {
        QueueAt qa;
        qa.qk = qk;
        qa.qr0 = qk * nsb / kN3Queues;
        qa.nr = (qk + 1) * nsb / kN3Queues - qa.qr0;
        qa.i0 = __builtin_amdgcn_readfirstlane(l.first[qa.qr0]);
        qa.i1 = __builtin_amdgcn_readfirstlane(l.first[qa.qr0 + qa.nr]);
        qa.il = __builtin_amdgcn_readfirstlane(l.first[qa.qr0 + (qa.nr > 0 ? qa.nr - 1 : 0)]);
        qa.span = span;
        if constexpr (F32) {
            __syncthreads();  // the previous item and the queue scan are done with nx and slot
            if (wave == 0) claim_item(qarr, qa, nx);
            __syncthreads();
            for (;;) {
                const int32_t it = __builtin_amdgcn_readfirstlane(nx[0]);
                if (it >= qa.i1) break;
                const int32_t S = __builtin_amdgcn_readfirstlane(nx[1]);
                const int32_t k0 = __builtin_amdgcn_readfirstlane(nx[2]);
                const int32_t k1 = __builtin_amdgcn_readfirstlane(nx[3]);
                const int32_t swp = __builtin_amdgcn_readfirstlane(nx[4]);
                if (threadIdx.x == 0) claim_flag = 0;
                // every wave has read nx (the claim during this item
                // rewrites it), and the previous item is done with the LDS
                // sums
                __syncthreads();
                n3_item<F32, DTC>(a, e, S, l.list + (size_t)S * (half + 1), k0, k1, it, swp, n_near, n_tested,
                                  smin, red, scr, ClaimAtBarrier{qarr, qa, &claim_flag, nx, false});
            }
        } else {
            for (;;) {
                __syncthreads();  // the previous item is done with nx and the LDS sums
                if (wave == 0) claim_item(qarr, qa, nx);
                __syncthreads();
                const int32_t it = __builtin_amdgcn_readfirstlane(nx[0]);
                if (it >= qa.i1) break;
                const int32_t S = __builtin_amdgcn_readfirstlane(nx[1]);
                const int32_t k0 = __builtin_amdgcn_readfirstlane(nx[2]);
                const int32_t k1 = __builtin_amdgcn_readfirstlane(nx[3]);
                const int32_t swp = __builtin_amdgcn_readfirstlane(nx[4]);
                n3_item<F32, DTC>(a, e, S, l.list + (size_t)S * (half + 1), k0, k1, it, swp, n_near, n_tested,
                                  smin, red, scr, ClaimAtBoundary{false});
            }
        }
        // the next queue with items left (a stale read only costs one more
        // round); the choice is wave 0's, broadcast through LDS
        __syncthreads();
        if (threadIdx.x < kN3Queues) {
            const int32_t t = (int32_t)threadIdx.x;
            const int32_t q = 8 * ((home + (t >> 3)) & 7) + ((sub + t) & 7);
            const int32_t q0 = l.first[q * nsb / kN3Queues], q1 = l.first[(q + 1) * nsb / kN3Queues];
            const bool left = q0 < q1 &&
                q0 + (int32_t)__hip_atomic_load(&l.plan->next[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < q1;
            const uint64_t m = __ballot(left);
            if (threadIdx.x == 0) slot = m ? 8 * ((home + (__builtin_ctzll(m) >> 3)) & 7) + ((sub + __builtin_ctzll(m)) & 7) : -1;
        }
        __syncthreads();
        qk = __builtin_amdgcn_readfirstlane(slot);
        if (qk < 0) break;
    }
    if (DTC) commit_dtmin(smin, &a.sc->dtmin_bits);
    n3_commit_counts(a.sc, (int32_t)blockIdx.x, n_near, n_tested, red);
}

}  // namespace

namespace {

// The launch constants of both kernels, after checking that the buffers fit
// the sizes the kernels index with.
hipError_t n3_row_args(const PairLaunch &a, N3Row *row) {
    const int64_t nsb = a.npad / kSuper;
    const N3List &l = a.lst;
    if (a.npad % kPad || !a.box || !a.react || !a.rflag || a.n > a.npad || (a.fp32pair && !a.qf))
        return hipErrorInvalidValue;
    // (a queue's rows, nsb / 64 + 1 at most, are held one per lane)
    if (!l.list || !l.cnt || !l.first || !l.plan || !l.exptab || l.target < 1 || l.cus < 1 ||
        (int64_t)l.max_items < (int64_t)l.target + nsb || nsb / 2 > kMaxHalf || nsb / kN3Queues + 1 > 63)
        return hipErrorInvalidValue;
    int32_t flags = 0;
    if (a.dtcheck) flags |= kFlagDtCheck;
    if (a.gated) flags |= kFlagGated;
    *row = N3Row{a.xs, a.ys, a.nc, a.q, a.qf, a.orig, a.box, a.part, a.react, a.rflag, a.sc,
                 (int32_t)a.n, (int32_t)a.npad, flags, a.kappa, a.rc2p};
    return hipSuccess;
}

}  // namespace

hipError_t launch_n3_tiles(const PairLaunch &a, hipStream_t s) {
    N3Row row;
    const hipError_t e = n3_row_args(a, &row);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_n3_tiles, dim3((unsigned)(a.npad / kSuper)), dim3(256), 0, s, row, a.lst);
    hipLaunchKernelGGL(k_n3_plan, dim3(1), dim3(256), 0, s, row, a.lst);
    return hipGetLastError();
}

hipError_t launch_pairs_n3(const PairLaunch &a, hipStream_t s) {
    N3Row row;
    const hipError_t e = n3_row_args(a, &row);
    if (e != hipSuccess) return e;
    auto kern = a.fp32pair ? (a.dtcheck ? k_pairs_n3<true, true> : k_pairs_n3<true, false>)
                           : (a.dtcheck ? k_pairs_n3<false, true> : k_pairs_n3<false, false>);
    // a persistent grid: as many workgroups per CU as fit at once (registers
    // and LDS of the variant), counted once per variant
    static int per_cu[2][2] = {};
    int &bpc = per_cu[a.fp32pair ? 1 : 0][a.dtcheck ? 1 : 0];
    if (bpc == 0) {
        int nb = 0;
        const hipError_t oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, 256, 0);
        if (oe != hipSuccess) return oe;
        bpc = nb > 0 ? nb : 1;
    }
    hipLaunchKernelGGL(kern, dim3((unsigned)(bpc * a.lst.cus)), dim3(256), 0, s, row, a.lst, exp_tab_coef());
    return hipGetLastError();
}

}  // namespace demon
