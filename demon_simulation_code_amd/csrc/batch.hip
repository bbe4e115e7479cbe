// batch.hip -- many independent clouds of uneven N stepped together.
//
// Each cloud is a Runge_Kutta4 + Confinement/Drag/ShieldedCoulomb(/Thermal)
// system of the reference (Runge_Kutta4.cpp:23-137) with its own kappa, C,
// gamma, its own dt check (Integrator.cpp:68-182) and its own time.  All
// clouds live in one arena; each cloud's slots are padded to a multiple of 256
// so that a 256-thread block of the O(N) kernels never straddles two clouds.
//
// All-pairs work is split into items.  A cloud in the gather form has items
// (cloud, 64R i-slots, j slice): one persistent launch per tile shape R pulls
// them from a device counter, one wavefront per item.  A cloud in the
// Newton-3 form (every unordered pair once, pairs_n3.h) has items (cloud,
// superblock row, slice of superblock offsets): one persistent launch pulls
// them, one workgroup per item.  Items are taken largest first; every wave
// leaves the loop when the counter passes the item count.  The per-stage
// update runs over all clouds in one launch and adds each Newton-3 cloud's
// reactions as the single-cloud stage kernel does.  Four steps (each a
// boundary and 4 x (pairs, update)) are captured once as a hipGraph and
// replayed.  The form and tile shape per N bucket come from the autotuner,
// demon_batch_tune.
#include <algorithm>
#include <cmath>
#include <limits>
#include <new>
#include <string>
#include <vector>

#include "../../include/demon_gpu.h"
#include "pair_core.h"
#include "pairs_n3.h"
#include "philox.h"

namespace demon {
namespace {

struct BCloud {
    int32_t off, n, npad;        // first slot, particles, padded size (multiple of 256)
    int32_t R;                   // i per lane (gather form), 0: Newton-3 form
    int32_t jslice;              // gather: j per slice; Newton-3: superblock offsets per item
    int32_t slices;              // slices per i block (gather) or per row (Newton-3)
    int64_t part_off;            // partial sums: part + part_off, layout [slices][2][npad]
    int64_t react_off;           // Newton-3: reactions, layout [nsb/2][2][npad]
    int64_t rflag_off;           // Newton-3: tile flags, layout [nsb][nsb/2]
    double kappa, rc2p, confine, gamma;
    uint64_t seed;
};

struct BItem {
    int32_t cloud, ibase, slice, pad_;  // ibase: first i slot (gather) or superblock row (Newton-3)
};

constexpr int kShapes = 3;       // gather forms R = 1, 2, 4
constexpr int kGroups = 4;       // item groups: the three gather shapes, then Newton-3
constexpr int kGroupN3 = 3;
constexpr int kCloudPad = 256;

__device__ __forceinline__ int32_t first_lane(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }

template <int R, bool F32>
__global__ __launch_bounds__(256) void k_bpairs(const double *__restrict__ xs,
                                                const double *__restrict__ ys,
                                                const double *__restrict__ nc,
                                                const double *__restrict__ box,
                                                const double *__restrict__ q,
                                                const int32_t *__restrict__ orig,
                                                double *__restrict__ part,
                                                const BCloud *__restrict__ clouds,
                                                const BItem *__restrict__ items, int32_t n_items,
                                                unsigned int *__restrict__ counter,
                                                Scalars *__restrict__ sc, int32_t base_flags,
                                                pair::ExpCoef ec) {
    const int lane = threadIdx.x & 63;
    for (;;) {
        int32_t it = 0;
        if (lane == 0) it = (int32_t)atomicAdd(counter, 1u);
        it = first_lane(it);
        if (it >= n_items) break;
        const BItem item = items[it];
        const BCloud cl = clouds[item.cloud];
        Scalars *s = sc + item.cloud;
        if ((base_flags & pair::kFlagGated) && !s->active) continue;
        const pair::Radii rad = pair::launch_radii(cl.rc2p, base_flags, s->rmax2_bits);
        const int32_t j_begin = item.slice * cl.jslice;
        const int32_t j_end = j_begin + cl.jslice < cl.npad ? j_begin + cl.jslice : cl.npad;
        const int32_t ibase = item.ibase + lane;
        double ax[R], ay[R];
#pragma unroll
        for (int r = 0; r < R; ++r) ax[r] = ay[r] = 0.0;
        uint64_t n_near = 0, n_tested = 0;
        pair::wave_tile<R, F32>(xs + cl.off, ys + cl.off, nc + cl.off, q + cl.off, orig + cl.off,
                                box + cl.off / 16,
                                cl.n, ibase, j_begin, j_end, base_flags, cl.kappa, cl.rc2p, rad,
                                ec, &s->dtmin_bits, ax, ay, n_near, n_tested);
        // the wave's counts (wave-uniform), one atomic each per item
        if (lane == 0) {
            unsigned long long *slot = s->pair_counts[(it + (int32_t)(threadIdx.x >> 6)) & (kCountSlots - 1)];
            atomicAdd(&slot[0], (unsigned long long)n_near);
            atomicAdd(&slot[1], (unsigned long long)n_tested);
        }
        double *px = part + cl.part_off + (size_t)(2 * item.slice) * cl.npad;
        double *py = px + cl.npad;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            px[ibase + 64 * r] = ax[r];
            py[ibase + 64 * r] = ay[r];
        }
    }
}

// Newton-3 items: one workgroup per item (row S of a cloud against a slice
// of superblock offsets).
template <bool F32, bool DTC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_bpairs_n3(
    const double *__restrict__ xs, const double *__restrict__ ys, const double *__restrict__ nc,
    const double *__restrict__ box, const double *__restrict__ q, const float *__restrict__ qf,
    const int32_t *__restrict__ orig, double *__restrict__ part, double *__restrict__ react,
    uint32_t *__restrict__ rflag, const BCloud *__restrict__ clouds, const BItem *__restrict__ items,
    int32_t n_items, unsigned int *__restrict__ counter, Scalars *__restrict__ sc, int32_t base_flags,
    pair::ExpCoef ec, const double *__restrict__ exptab) {
    __shared__ n3::RedLds red;
    __shared__ __align__(16) n3::ScrLds<F32> scr;
    __shared__ n3::TBoxLds tbox;
    __shared__ double etab[kExpTabN];
    __shared__ int32_t slot;
    // the FP64 exp is exp_neg_tab on this table; the item loop's first
    // barrier orders these stores before any read
    if (!F32)
        for (int k = (int)threadIdx.x; k < kExpTabN; k += 256) etab[k] = exptab[k];
    for (;;) {
        // the previous item is done with the LDS sums and with slot
        __syncthreads();
        if (threadIdx.x == 0) slot = (int32_t)atomicAdd(counter, 1u);
        __syncthreads();
        const int32_t it = first_lane(slot);
        if (it >= n_items) break;
        const BItem item = items[it];
        const BCloud cl = clouds[item.cloud];
        Scalars *s = sc + item.cloud;
        if ((base_flags & pair::kFlagGated) && !s->active) continue;
        const int32_t half = cl.npad / kSuper / 2;
        const int32_t d0 = item.slice * cl.jslice;
        const int32_t d1 = d0 + cl.jslice < half + 1 ? d0 + cl.jslice : half + 1;
        const n3::N3Row row{xs + cl.off, ys + cl.off, nc + cl.off, q + cl.off, qf + cl.off, orig + cl.off,
                            box + cl.off / 16, part + cl.part_off, react + cl.react_off,
                            rflag + cl.rflag_off, s, cl.n, cl.npad, base_flags, cl.kappa, cl.rc2p};
        n3::n3_row<F32, DTC>(row, ec, etab, item.ibase, d0, d1, item.slice, red, scr, tbox);
    }
}

struct BUpdate {
    StageBuf cur, nxt;
    double *x, *y, *vx, *vy;
    const double *q, *m;
    const int32_t *orig, *block_cloud;
    const double *part;
    const double *react;     // Newton-3 clouds' reactions (BCloud::react_off)
    const uint32_t *rflag;   // and tile flags (BCloud::rflag_off)
    double *acc1, *acc23;    // [4][P]
    const BCloud *clouds;
    Scalars *sc;
    unsigned int *counters;  // reset for the next pair launches
    int64_t P;
    int32_t forces, dtcheck;
    double heat;
};

__device__ __forceinline__ double bda(double a1, double a23, double a4) {
    return (a1 + (a23 * 2.0 + a4)) / 6.0;
}

template <int STAGE>
__global__ __launch_bounds__(256) void k_bupdate(BUpdate a) {
    __shared__ double red[4];
    __shared__ unsigned long long fmask[kMaxHalf / 64];
    if (blockIdx.x == 0 && threadIdx.x < kGroups) a.counters[threadIdx.x] = 0u;
    const int32_t c = a.block_cloud[blockIdx.x];
    const BCloud cl = a.clouds[c];
    Scalars *sc = a.sc + c;
    if (!sc->active) return;  // uniform per block
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int32_t i = (int32_t)(p - cl.off);
    // a block is one superblock of one cloud; for a Newton-3 cloud, the flags
    // of the tiles whose reactions land on it, as a bitmask over d - 1
    const bool n3c = cl.R == 0 && (a.forces & 4);
    const int32_t half = cl.npad / kSuper / 2;
    if (n3c && threadIdx.x < 64) {
        const uint32_t *fl = a.rflag + cl.rflag_off + (size_t)(i / kSuper) * half;
        for (int32_t w = 0; w < (half + 63) / 64; ++w) {
            const int32_t k = 64 * w + (int32_t)threadIdx.x;
            const unsigned long long m = __ballot(k < half && fl[k] != 0u);
            if (threadIdx.x == 0) fmask[w] = m;
        }
    }
    __syncthreads();
    double dt;
    if (STAGE == 1) {
        dt = sc->init_dt;
        int err = 0;
        if (a.dtcheck) {
            const float d = sqrtf(__uint_as_float(sc->dtmin_bits));
            float dist = 1.0e-4f;
            int k = 0;
            while (d <= dist) {
                dist = dist / 10.0f;
                dt = dt / 10.0;
                if (++k > 60) {
                    err = 2;
                    break;
                }
            }
        }
        if (i == 0) {
            sc->dt_step = dt;
            if (dt != sc->init_dt) sc->dt_reductions += 1;
            if (err) sc->err = err;
            sc->rmax2_bits = 0;
        }
    } else {
        dt = sc->dt_step;
    }
    double r2new = 0.0, nx = 0.0, ny = 0.0;
    if (i < cl.n) {
        const double xs = a.cur.xs[p], ys = a.cur.ys[p], vxs = a.cur.vxs[p], vys = a.cur.vys[p];
        const double qi = a.q[p], mi = a.m[p];
        double fx = 0.0, fy = 0.0;
        if (a.forces & 1) {
            const double cv = qi * cl.confine;
            fx = fx + cv * xs;
            fy = fy + cv * ys;
        }
        if (a.forces & 2) {
            const double dr = mi * (-cl.gamma);
            fx = fx + dr * vxs;
            fy = fy + dr * vys;
        }
        if (a.forces & 4) {
            // the slices' sums, four slices' loads issued before their adds
            // (the same order of additions, more loads in flight)
            const double *pp = a.part + cl.part_off;
            int32_t s = 0;
            for (; s + 4 <= cl.slices; s += 4) {
                double px[4], py[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    px[k] = pp[(size_t)(2 * (s + k)) * cl.npad + i];
                    py[k] = pp[(size_t)(2 * (s + k) + 1) * cl.npad + i];
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    fx = fx + px[k];
                    fy = fy + py[k];
                }
            }
            for (; s < cl.slices; ++s) {
                fx = fx + pp[(size_t)(2 * s) * cl.npad + i];
                fy = fy + pp[(size_t)(2 * s + 1) * cl.npad + i];
            }
            if (n3c) {
                // reactions in increasing superblock offset d (bit d - 1), up
                // to four offsets' loads in flight
                const double *rr = a.react + cl.react_off;
                for (int32_t w = 0; w < (half + 63) / 64; ++w) {
                    unsigned long long mask = fmask[w];
                    while (mask) {
                        int32_t dd[4];
                        int m = 0;
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            dd[k] = 0;
                            if (mask) {
                                dd[k] = 64 * w + __builtin_ctzll(mask);
                                mask &= mask - 1;
                                m = k + 1;
                            }
                        }
                        double rx[4], ry[4];
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            rx[k] = k < m ? rr[(size_t)dd[k] * 2 * cl.npad + i] : 0.0;
                            ry[k] = k < m ? rr[(size_t)dd[k] * 2 * cl.npad + cl.npad + i] : 0.0;
                        }
#pragma unroll
                        for (int k = 0; k < 4; ++k)
                            if (k < m) {
                                fx = fx + rx[k];
                                fy = fy + ry[k];
                            }
                    }
                }
            }
        }
        if (a.forces & 16) {
            const U32x4 o = draw(cl.seed, (uint64_t)a.orig[p], sc->step, kDomainThermal, (uint32_t)STAGE);
            const double tv = u01(o.v[0], o.v[1]) * a.heat;
            const double th = u01(o.v[2], o.v[3]) * 6.283185307179586476925286766559;
            fx = fx + tv * cos(th);
            fy = fy + tv * sin(th);
        }
        const double k = (dt * fx) / mi, l = dt * vxs, mm = (dt * fy) / mi, nn = dt * vys;
        const double x = a.x[p], y = a.y[p], vx = a.vx[p], vy = a.vy[p];
        const int64_t P = a.P;
        double nvx, nvy;
        if (STAGE == 1 || STAGE == 2) {
            double *A = STAGE == 1 ? a.acc1 : a.acc23;
            A[p] = k;
            A[P + p] = l;
            A[2 * P + p] = mm;
            A[3 * P + p] = nn;
            nx = 0.5 * l + x;
            ny = 0.5 * nn + y;
            nvx = 0.5 * k + vx;
            nvy = 0.5 * mm + vy;
        } else if (STAGE == 3) {
            a.acc23[p] += k;
            a.acc23[P + p] += l;
            a.acc23[2 * P + p] += mm;
            a.acc23[3 * P + p] += nn;
            nx = x + l;
            ny = y + nn;
            nvx = vx + k;
            nvy = vy + mm;
        } else {
            nvx = vx + bda(a.acc1[p], a.acc23[p], k);
            nx = x + bda(a.acc1[P + p], a.acc23[P + p], l);
            nvy = vy + bda(a.acc1[2 * P + p], a.acc23[2 * P + p], mm);
            ny = y + bda(a.acc1[3 * P + p], a.acc23[3 * P + p], nn);
            a.x[p] = nx;
            a.y[p] = ny;
            a.vx[p] = nvx;
            a.vy[p] = nvy;
        }
        const double r2 = nx * nx + ny * ny;
        a.nxt.xs[p] = nx;
        a.nxt.ys[p] = ny;
        a.nxt.vxs[p] = nvx;
        a.nxt.vys[p] = nvy;
        a.nxt.nc[p] = -r2 * kNormShrink;
        r2new = r2;
    }
    // clouds start at multiples of 256, so a 64-block never spans two clouds
    wave_box(i < cl.n, nx, ny, a.nxt.box + 4 * (p / kBoxBlock));
    if (STAGE == 4) {
        double v = r2new;
        for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            const double b = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
            if (b > 0.0) atomicMax(&sc->rmax2_bits, (unsigned long long)__double_as_longlong(b));
        }
    }
}

__global__ void k_bboundary(Scalars *sc, int32_t nclouds, int closing, unsigned int *counters) {
    const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c == 0)
        for (int k = 0; k < kGroups; ++k) counters[k] = 0u;
    if (c >= nclouds) return;
    Scalars *s = sc + c;
    if (s->active) {
        s->t = s->t + s->dt_step;
        s->steps_done += 1;
        s->step += 1;
    }
    const int act = !closing && (s->t < s->t_end) && (s->remaining > 0) && (s->err == 0);
    s->active = act;
    if (act) s->remaining -= 1;
    s->dtmin_bits = 0x7f800000u;
}

__global__ void k_bsetup(Scalars *sc, int32_t nclouds, double init_dt, unsigned long long steps) {
    const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nclouds) return;
    sc[c].t_end = __builtin_inf();
    sc[c].init_dt = init_dt;
    sc[c].remaining = steps;
    sc[c].active = 0;
}

// per-cloud max |r|^2 after an upload
__global__ void k_brmax(const double *xs, const double *ys, const int32_t *block_cloud,
                        const BCloud *clouds, Scalars *sc) {
    __shared__ double red[4];
    const int32_t c = block_cloud[blockIdx.x];
    const BCloud cl = clouds[c];
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double v = 0.0;
    if (p - cl.off < cl.n) v = xs[p] * xs[p] + ys[p] * ys[p];
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double b = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        if (b > 0.0) atomicMax(&sc[c].rmax2_bits, (unsigned long long)__double_as_longlong(b));
    }
}

}  // namespace
}  // namespace demon

using namespace demon;

struct demon_batch {
    std::string err;
    int32_t nclouds = 0, device = 0, forces = 0, dt_check = 1, precision = 0;
    double heat = 0.0;
    uint64_t seed = 0;
    std::vector<int64_t> n;
    std::vector<BCloud> hc;
    std::vector<std::vector<int64_t>> perm;
    std::vector<int32_t> block_cloud;
    int64_t P = 0;
    hipStream_t stream = nullptr;
    double *arena = nullptr;     // 30 arrays of P
    double *part = nullptr;
    size_t part_doubles = 0;
    double *react = nullptr;     // Newton-3 clouds' reactions
    size_t react_doubles = 0;
    uint32_t *rflag = nullptr;   // and tile flags
    size_t rflag_words = 0;
    float *qf = nullptr;         // charges in FP32
    int32_t *orig = nullptr, *d_block_cloud = nullptr;
    BCloud *d_clouds = nullptr;
    Scalars *d_sc = nullptr;
    BItem *d_items = nullptr;
    size_t items_cap = 0;
    unsigned int *d_counters = nullptr;
    double *exptab = nullptr;    // e^{-k/64}, the Newton-3 form's exp table
    int32_t group_off[kGroups]{}, group_count[kGroups]{};
    double *x = nullptr, *y = nullptr, *vx = nullptr, *vy = nullptr, *q = nullptr, *m = nullptr;
    StageBuf sb[2]{};
    double *acc1 = nullptr, *acc23 = nullptr;
    // form and tile shape per N bucket (bucket = floor(log2 npad), 0..31):
    // R = 1, 2, 4 gather form with j slices of js; R = 0 Newton-3 form with
    // js superblock offsets per item
    int32_t bucket_R[32], bucket_js[32];
    float bucket_us[32];
    bool use_graph = true;
    hipGraphExec_t graph = nullptr;
    hipEvent_t ev[2]{};
    double last_ms = 0.0;
    bool uploaded_all = false;
    std::vector<char> uploaded;
    // per-launch HIP events around the pair and update launches of plain
    // (uncaptured) steps, when timing is on
    bool timing = false;
    std::vector<hipEvent_t> tev;
    std::vector<uint8_t> tkind;  // 0 pairs, 1 update, per event pair
    size_t tev_used = 0;
    double pair_ms = 0.0, update_ms = 0.0;
    int64_t pair_launches = 0, update_launches = 0;
};

namespace {

thread_local std::string g_berr;

int bfail(demon_batch *b, int code, const std::string &m) {
    if (b) b->err = m;
    else g_berr = m;
    return code;
}

#define BHIP(b, expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return bfail((b), DEMON_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int bucket_of(int64_t npad) {
    int k = 0;
    while ((int64_t(1) << (k + 1)) <= npad) ++k;
    return k;
}

int shape_index(int R) { return R == 0 ? kGroupN3 : R == 1 ? 0 : R == 2 ? 1 : 2; }

// Items of one cloud, with their cost (pair evaluations) for the ordering.
void cloud_items(const BCloud &cl, int32_t c, std::vector<std::pair<int64_t, BItem>> &out) {
    if (cl.R == 0) {
        const int32_t nsb = cl.npad / kSuper, half = nsb / 2;
        for (int32_t S = 0; S < nsb; ++S)
            for (int32_t sl = 0; sl < cl.slices; ++sl) {
                const int32_t d0 = sl * cl.jslice, d1 = std::min(d0 + cl.jslice, half + 1);
                out.push_back({(int64_t)(d1 - d0) * kSuper * kSuper, BItem{c, S, sl, 0}});
            }
        return;
    }
    for (int32_t ib = 0; ib < cl.npad; ib += 64 * cl.R)
        for (int32_t sl = 0; sl < cl.slices; ++sl) {
            const int32_t jl = std::min(cl.jslice, cl.npad - sl * cl.jslice);
            out.push_back({(int64_t)jl * 64 * cl.R, BItem{c, ib, sl, 0}});
        }
}

// Recomputes forms, shapes, buffer offsets and the item lists from the
// bucket table.
int rebuild_items(demon_batch *b) {
    int64_t part_off = 0, react_off = 0, rflag_off = 0;
    for (int32_t c = 0; c < b->nclouds; ++c) {
        BCloud &cl = b->hc[(size_t)c];
        const int bk = bucket_of(cl.npad);
        cl.R = b->bucket_R[bk];
        const int32_t nsb = cl.npad / kSuper, half = nsb / 2;
        if (cl.R == 0 && half > kMaxHalf) cl.R = 1;  // the update's flag mask holds kMaxHalf tiles
        if (cl.R == 0) {
            cl.jslice = std::max<int32_t>(1, std::min<int32_t>(b->bucket_js[bk], half + 1));
            cl.slices = (half + 1 + cl.jslice - 1) / cl.jslice;
            cl.react_off = react_off;
            cl.rflag_off = rflag_off;
            react_off += (int64_t)half * 2 * cl.npad;
            rflag_off += (int64_t)nsb * half;
        } else {
            cl.jslice = std::min<int32_t>(b->bucket_R[bk] == 0 ? 512 : b->bucket_js[bk], cl.npad);
            cl.slices = (cl.npad + cl.jslice - 1) / cl.jslice;
            cl.react_off = cl.rflag_off = 0;
        }
        cl.part_off = part_off;
        part_off += (int64_t)cl.slices * 2 * cl.npad;
    }
    std::vector<std::pair<int64_t, BItem>> all[kGroups];
    for (int32_t c = 0; c < b->nclouds; ++c) {
        const BCloud &cl = b->hc[(size_t)c];
        cloud_items(cl, c, all[shape_index(cl.R)]);
    }
    std::vector<BItem> items;
    for (int g = 0; g < kGroups; ++g) {
        std::stable_sort(all[g].begin(), all[g].end(),
                         [](const std::pair<int64_t, BItem> &a, const std::pair<int64_t, BItem> &b2) {
                             return a.first > b2.first;
                         });
        b->group_off[g] = (int32_t)items.size();
        b->group_count[g] = (int32_t)all[g].size();
        for (auto &e : all[g]) items.push_back(e.second);
    }
    BHIP(b, hipStreamSynchronize(b->stream));
    if (items.size() > b->items_cap) {
        if (b->d_items) (void)hipFree(b->d_items);
        b->d_items = nullptr;
        if (hipMalloc((void **)&b->d_items, sizeof(BItem) * items.size()) != hipSuccess)
            return bfail(b, DEMON_ERR_ALLOC, "hipMalloc items failed");
        b->items_cap = items.size();
    }
    if ((size_t)part_off > b->part_doubles) {
        if (b->part) (void)hipFree(b->part);
        b->part = nullptr;
        if (hipMalloc((void **)&b->part, sizeof(double) * (size_t)part_off) != hipSuccess)
            return bfail(b, DEMON_ERR_ALLOC, "hipMalloc partial sums failed");
        b->part_doubles = (size_t)part_off;
    }
    // at least one element each, so that the kernels always get valid pointers
    if ((size_t)std::max<int64_t>(react_off, 1) > b->react_doubles) {
        if (b->react) (void)hipFree(b->react);
        b->react = nullptr;
        const size_t cnt = (size_t)std::max<int64_t>(react_off, 1);
        if (hipMalloc((void **)&b->react, sizeof(double) * cnt) != hipSuccess)
            return bfail(b, DEMON_ERR_ALLOC, "hipMalloc reaction sums failed");
        b->react_doubles = cnt;
    }
    if ((size_t)std::max<int64_t>(rflag_off, 1) > b->rflag_words) {
        if (b->rflag) (void)hipFree(b->rflag);
        b->rflag = nullptr;
        const size_t cnt = (size_t)std::max<int64_t>(rflag_off, 1);
        if (hipMalloc((void **)&b->rflag, sizeof(uint32_t) * cnt) != hipSuccess)
            return bfail(b, DEMON_ERR_ALLOC, "hipMalloc tile flags failed");
        b->rflag_words = cnt;
    }
    BHIP(b, hipMemcpyAsync(b->d_items, items.data(), sizeof(BItem) * items.size(), hipMemcpyHostToDevice,
                           b->stream));
    BHIP(b, hipMemcpyAsync(b->d_clouds, b->hc.data(), sizeof(BCloud) * b->hc.size(), hipMemcpyHostToDevice,
                           b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));  // items is a local
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    return DEMON_OK;
}

int grid_for(int32_t count, int32_t items_per_block) {
    // persistent grid: enough waves to fill 256 CUs several times over, never
    // more blocks than needed for one item per wave (gather) or per
    // workgroup (Newton-3)
    const int32_t blocks = std::min<int32_t>(2048, (count + items_per_block - 1) / items_per_block);
    return std::max(blocks, 1);
}

int enqueue_bpairs(demon_batch *b, const StageBuf &st, int32_t flags, int only_group = -1,
                   const BItem *items = nullptr, int32_t count = -1) {
    const bool f32 = b->precision == DEMON_PRECISION_FP32_PAIR;
    for (int g = 0; g < kGroups; ++g) {
        if (only_group >= 0 && g != only_group) continue;
        const int32_t cnt = count >= 0 ? count : b->group_count[g];
        if (cnt == 0) continue;
        const BItem *it = items ? items : b->d_items + b->group_off[g];
        unsigned int *ctr = b->d_counters + g;
        const pair::ExpCoef ec = g == kGroupN3 ? pair::exp_tab_coef() : pair::exp_coef();
        if (g == kGroupN3) {
            const dim3 grid((unsigned)grid_for(cnt, 1)), blk(256);
            const bool dtc = (flags & pair::kFlagDtCheck) != 0;
            auto kern = f32 ? (dtc ? k_bpairs_n3<true, true> : k_bpairs_n3<true, false>)
                            : (dtc ? k_bpairs_n3<false, true> : k_bpairs_n3<false, false>);
            hipLaunchKernelGGL(kern, grid, blk, 0, b->stream, st.xs, st.ys, st.nc, st.box, b->q, b->qf, b->orig,
                               b->part, b->react, b->rflag, b->d_clouds, it, cnt, ctr, b->d_sc, flags, ec, b->exptab);
            BHIP(b, hipGetLastError());
            continue;
        }
        const dim3 grid((unsigned)grid_for(cnt, 4)), blk(256);
#define BLAUNCH(RR)                                                                                        \
    if (f32)                                                                                               \
        hipLaunchKernelGGL((k_bpairs<RR, true>), grid, blk, 0, b->stream, st.xs, st.ys, st.nc, st.box, b->q,  \
                           b->orig, b->part, b->d_clouds, it, cnt, ctr, b->d_sc, flags, ec);             \
    else                                                                                                   \
        hipLaunchKernelGGL((k_bpairs<RR, false>), grid, blk, 0, b->stream, st.xs, st.ys, st.nc, st.box, b->q, \
                           b->orig, b->part, b->d_clouds, it, cnt, ctr, b->d_sc, flags, ec);
        if (g == 0) { BLAUNCH(1) } else if (g == 1) { BLAUNCH(2) } else { BLAUNCH(4) }
#undef BLAUNCH
        BHIP(b, hipGetLastError());
    }
    return DEMON_OK;
}

BUpdate bupdate(demon_batch *b, int stage) {
    BUpdate u{};
    const int cur = (stage - 1) & 1;
    u.cur = b->sb[cur];
    u.nxt = b->sb[cur ^ 1];
    u.x = b->x; u.y = b->y; u.vx = b->vx; u.vy = b->vy; u.q = b->q; u.m = b->m;
    u.orig = b->orig; u.block_cloud = b->d_block_cloud; u.part = b->part;
    u.react = b->react; u.rflag = b->rflag;
    u.acc1 = b->acc1; u.acc23 = b->acc23; u.clouds = b->d_clouds; u.sc = b->d_sc;
    u.counters = b->d_counters; u.P = b->P; u.forces = b->forces; u.dtcheck = b->dt_check;
    u.heat = b->heat;
    return u;
}

// Event pair around one plain launch (timed: not under capture).
int tev_open(demon_batch *b, bool timed) {
    if (!timed) return DEMON_OK;
    if (b->tev_used + 2 > b->tev.size())
        for (int k = 0; k < 64; ++k) {
            hipEvent_t e;
            BHIP(b, hipEventCreate(&e));
            b->tev.push_back(e);
        }
    BHIP(b, hipEventRecord(b->tev[b->tev_used], b->stream));
    return DEMON_OK;
}

int tev_close(demon_batch *b, bool timed, uint8_t kind) {
    if (!timed) return DEMON_OK;
    BHIP(b, hipEventRecord(b->tev[b->tev_used + 1], b->stream));
    b->tev_used += 2;
    b->tkind.push_back(kind);
    return DEMON_OK;
}

int enqueue_bstep(demon_batch *b, bool timed) {
    hipLaunchKernelGGL(k_bboundary, dim3((b->nclouds + 255) / 256), dim3(256), 0, b->stream, b->d_sc,
                       b->nclouds, 0, b->d_counters);
    BHIP(b, hipGetLastError());
    const dim3 grid((unsigned)(b->P / 256)), blk(256);
    int rc;
    for (int s = 1; s <= 4; ++s) {
        int32_t flags = pair::kFlagGated;
        if (s == 1 && b->dt_check) flags |= pair::kFlagDtCheck;
        if (b->forces & DEMON_FORCE_SHIELDED_COULOMB) {
            if ((rc = tev_open(b, timed)) || (rc = enqueue_bpairs(b, b->sb[(s - 1) & 1], flags)) ||
                (rc = tev_close(b, timed, 0)))
                return rc;
        }
        const BUpdate u = bupdate(b, s);
        if ((rc = tev_open(b, timed))) return rc;
        switch (s) {
            case 1: hipLaunchKernelGGL(k_bupdate<1>, grid, blk, 0, b->stream, u); break;
            case 2: hipLaunchKernelGGL(k_bupdate<2>, grid, blk, 0, b->stream, u); break;
            case 3: hipLaunchKernelGGL(k_bupdate<3>, grid, blk, 0, b->stream, u); break;
            default: hipLaunchKernelGGL(k_bupdate<4>, grid, blk, 0, b->stream, u); break;
        }
        BHIP(b, hipGetLastError());
        if ((rc = tev_close(b, timed, 1))) return rc;
    }
    return DEMON_OK;
}

}  // namespace

extern "C" {

const char *demon_batch_last_error(const demon_batch *b) { return b ? b->err.c_str() : g_berr.c_str(); }

int demon_batch_create(demon_batch **out, int32_t nclouds, const int64_t *n, const double *kappa,
                       const double *confine, const double *gamma, const demon_params *p) {
    if (!out || !n || !kappa || !confine || !gamma || !p || nclouds < 1)
        return bfail(nullptr, DEMON_ERR_ARG, "bad batch arguments");
    *out = nullptr;
    const int32_t known = DEMON_FORCE_CONFINEMENT | DEMON_FORCE_DRAG | DEMON_FORCE_SHIELDED_COULOMB | DEMON_FORCE_THERMAL;
    if (p->forces & ~known) return bfail(nullptr, DEMON_ERR_UNSUPPORTED, "unsupported force flags");
    if (p->precision != DEMON_PRECISION_FP64 && p->precision != DEMON_PRECISION_FP32_PAIR)
        return bfail(nullptr, DEMON_ERR_ARG, "unknown precision");
    if (p->dt_check && !(p->forces & DEMON_FORCE_SHIELDED_COULOMB))
        return bfail(nullptr, DEMON_ERR_UNSUPPORTED, "the dt check runs inside the all-pairs pass");
    demon_batch *b = new (std::nothrow) demon_batch();
    if (!b) return bfail(nullptr, DEMON_ERR_ALLOC, "host allocation failed");
    b->nclouds = nclouds;
    b->device = p->device;
    b->forces = p->forces;
    b->dt_check = p->dt_check;
    b->precision = p->precision;
    b->heat = p->heat_val;
    b->seed = p->seed;
    auto bail = [&](int code) {
        g_berr = b->err;
        demon_batch_destroy(b);
        return code;
    };
    int64_t off = 0;
    for (int32_t c = 0; c < nclouds; ++c) {
        if (n[c] < 2 || n[c] > (1 << 24)) return bail(bfail(b, DEMON_ERR_ARG, "cloud size out of [2, 2^24]"));
        if ((p->forces & DEMON_FORCE_SHIELDED_COULOMB) && !(kappa[c] > 0.0))
            return bail(bfail(b, DEMON_ERR_ARG, "shielding must be positive"));
        BCloud cl{};
        cl.off = (int32_t)off;
        cl.n = (int32_t)n[c];
        cl.npad = (int32_t)((n[c] + kCloudPad - 1) / kCloudPad * kCloudPad);
        cl.kappa = kappa[c];
        const double rc = 10.0 / kappa[c];
        cl.rc2p = rc * rc * (1.0 + kCutoffSlack);
        cl.confine = confine[c];
        cl.gamma = gamma[c];
        cl.seed = p->seed + (uint64_t)c * 0x9E3779B97F4A7C15ull;
        b->hc.push_back(cl);
        b->n.push_back(n[c]);
        for (int32_t k = 0; k < cl.npad / 256; ++k) b->block_cloud.push_back(c);
        off += cl.npad;
        if (off > (int64_t)1 << 30) return bail(bfail(b, DEMON_ERR_ARG, "batch too large"));
    }
    b->P = off;
    b->perm.resize((size_t)nclouds);
    b->uploaded.assign((size_t)nclouds, 0);
    for (int k = 0; k < 32; ++k) {
        b->bucket_R[k] = 1;
        b->bucket_js[k] = 512;
        b->bucket_us[k] = 0.0f;
    }
    if (hipSetDevice(p->device) != hipSuccess) return bail(bfail(b, DEMON_ERR_HIP, "hipSetDevice failed"));
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(bfail(b, DEMON_ERR_HIP, "hipStreamCreate failed"));
    const size_t P = (size_t)b->P;
    if (hipMalloc((void **)&b->arena, sizeof(double) * P * 25) != hipSuccess ||
        hipMalloc((void **)&b->orig, sizeof(int32_t) * P) != hipSuccess ||
        hipMalloc((void **)&b->d_block_cloud, sizeof(int32_t) * b->block_cloud.size()) != hipSuccess ||
        hipMalloc((void **)&b->d_clouds, sizeof(BCloud) * (size_t)nclouds) != hipSuccess ||
        hipMalloc((void **)&b->d_sc, sizeof(Scalars) * (size_t)nclouds) != hipSuccess ||
        hipMalloc((void **)&b->d_counters, sizeof(unsigned int) * kGroups) != hipSuccess ||
        hipMalloc((void **)&b->exptab, sizeof(double) * kExpTabN) != hipSuccess)
        return bail(bfail(b, DEMON_ERR_ALLOC, "hipMalloc failed"));
    // every fill and copy goes on the batch's own stream: it is non-blocking,
    // so work on the null stream would not be ordered before its kernels
    if (hipMemsetAsync(b->arena, 0, sizeof(double) * P * 25, b->stream) != hipSuccess ||
        hipMemsetAsync(b->d_sc, 0, sizeof(Scalars) * (size_t)nclouds, b->stream) != hipSuccess ||
        hipMemsetAsync(b->d_counters, 0, sizeof(unsigned int) * kGroups, b->stream) != hipSuccess)
        return bail(bfail(b, DEMON_ERR_HIP, "hipMemset failed"));
    double *a = b->arena;
    auto take = [&]() { double *r = a; a += P; return r; };
    b->x = take(); b->y = take(); b->vx = take(); b->vy = take(); b->q = take(); b->m = take();
    for (StageBuf *s : {&b->sb[0], &b->sb[1]}) {
        s->xs = take(); s->ys = take(); s->vxs = take(); s->vys = take(); s->nc = take();
    }
    b->acc1 = a; a += 4 * P;
    b->acc23 = a; a += 4 * P;
    for (StageBuf *s : {&b->sb[0], &b->sb[1]}) {  // box tables, P/16 doubles each
        s->box = a;
        a += P / 16;
    }
    b->qf = reinterpret_cast<float *>(a);  // P floats
    a += P / 2;
    if (hipMemcpyAsync(b->d_block_cloud, b->block_cloud.data(), sizeof(int32_t) * b->block_cloud.size(),
                       hipMemcpyHostToDevice, b->stream) != hipSuccess)
        return bail(bfail(b, DEMON_ERR_HIP, "hipMemcpy failed"));
    {
        double tab[kExpTabN];
        exp_table_host(tab);
        if (hipMemcpyAsync(b->exptab, tab, sizeof tab, hipMemcpyHostToDevice, b->stream) != hipSuccess ||
            hipStreamSynchronize(b->stream) != hipSuccess)  // tab is a local
            return bail(bfail(b, DEMON_ERR_HIP, "hipMemcpy failed"));
    }
    for (int k = 0; k < 2; ++k)
        if (hipEventCreate(&b->ev[k]) != hipSuccess) return bail(bfail(b, DEMON_ERR_HIP, "hipEventCreate failed"));
    // padding: nc = -inf everywhere until clouds are uploaded
    for (int k = 0; k < 2; ++k)
        if (launch_prepare_stage(b->sb[k], 0, b->P, b->stream) != hipSuccess)
            return bail(bfail(b, DEMON_ERR_HIP, "prepare failed"));
    const int rc = rebuild_items(b);
    if (rc) return bail(rc);
    *out = b;
    return DEMON_OK;
}

void demon_batch_destroy(demon_batch *b) {
    if (!b) return;
    // drain the stream before the captured step goes away
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    for (hipEvent_t e : b->ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->tev) (void)hipEventDestroy(e);
    for (void *ptr : {(void *)b->arena, (void *)b->part, (void *)b->react, (void *)b->rflag, (void *)b->orig,
                      (void *)b->d_block_cloud,
                      (void *)b->d_clouds, (void *)b->d_sc, (void *)b->d_items, (void *)b->d_counters,
                      (void *)b->exptab})
        if (ptr) (void)hipFree(ptr);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int demon_batch_upload(demon_batch *b, int32_t cloud, const double *x, const double *y,
                       const double *vx, const double *vy, const double *charge, const double *mass) {
    if (!b || cloud < 0 || cloud >= b->nclouds || !x || !y || !charge || !mass)
        return bfail(b, DEMON_ERR_ARG, "bad upload arguments");
    BHIP(b, hipSetDevice(b->device));
    BCloud &cl = b->hc[(size_t)cloud];
    const int64_t n = cl.n;
    // Morton order within the cloud (as the single-cloud context)
    double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
    for (int64_t i = 0; i < n; ++i) {
        x0 = std::min(x0, x[i]); x1 = std::max(x1, x[i]);
        y0 = std::min(y0, y[i]); y1 = std::max(y1, y[i]);
    }
    const double sx = x1 > x0 ? 65535.0 / (x1 - x0) : 0.0, sy = y1 > y0 ? 65535.0 / (y1 - y0) : 0.0;
    auto spread = [](uint64_t v) {
        v = (v | (v << 8)) & 0x00FF00FFull;
        v = (v | (v << 4)) & 0x0F0F0F0Full;
        v = (v | (v << 2)) & 0x33333333ull;
        v = (v | (v << 1)) & 0x55555555ull;
        return v;
    };
    std::vector<std::pair<uint64_t, int64_t>> key((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        uint32_t ax = 0, ay = 0;
        if (std::isfinite(x[i]) && std::isfinite(y[i])) {
            ax = (uint32_t)std::min(65535.0, std::max(0.0, (x[i] - x0) * sx));
            ay = (uint32_t)std::min(65535.0, std::max(0.0, (y[i] - y0) * sy));
        }
        key[(size_t)i] = {spread(ax) | (spread(ay) << 1), i};
    }
    std::sort(key.begin(), key.end());
    std::vector<int64_t> &perm = b->perm[(size_t)cloud];
    perm.resize((size_t)n);
    for (int64_t k = 0; k < n; ++k) perm[(size_t)k] = key[(size_t)k].second;
    const int64_t np = cl.npad;
    std::vector<double> h((size_t)np * 6, 0.0);
    std::vector<int32_t> o((size_t)np);
    const double *src[6] = {x, y, vx, vy, charge, mass};
    for (int a = 0; a < 6; ++a)
        for (int64_t k = 0; k < n; ++k) h[(size_t)(a * np + k)] = src[a] ? src[a][perm[(size_t)k]] : 0.0;
    for (int64_t k = 0; k < np; ++k) o[(size_t)k] = k < n ? (int32_t)perm[(size_t)k] : (int32_t)(0x40000000 + k);
    BHIP(b, hipStreamSynchronize(b->stream));
    // Every copy is ordered on the batch's stream.  A device-to-device
    // hipMemcpy runs on the null stream and returns before it completes; the
    // batch's stream is non-blocking, so the box pass below could otherwise
    // read the stage record before it lands (boxes of stale or zero
    // coordinates make far block pairs look fully inside the cutoff).
    double *dst[6] = {b->x, b->y, b->vx, b->vy, b->q, b->m};
    for (int a = 0; a < 6; ++a)
        BHIP(b, hipMemcpyAsync(dst[a] + cl.off, h.data() + (size_t)a * np, sizeof(double) * np,
                               hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipMemcpyAsync(b->orig + cl.off, o.data(), sizeof(int32_t) * np, hipMemcpyHostToDevice, b->stream));
    BHIP(b, launch_round_fp32(b->q + cl.off, b->qf + cl.off, np, b->stream));
    // stage-1 record of this cloud
    const double *bs[4] = {b->x, b->y, b->vx, b->vy};
    double *rec[4] = {b->sb[0].xs, b->sb[0].ys, b->sb[0].vxs, b->sb[0].vys};
    for (int a = 0; a < 4; ++a)
        BHIP(b, hipMemcpyAsync(rec[a] + cl.off, bs[a] + cl.off, sizeof(double) * np, hipMemcpyDeviceToDevice,
                               b->stream));
    StageBuf view{b->sb[0].xs + cl.off, b->sb[0].ys + cl.off, b->sb[0].vxs + cl.off, b->sb[0].vys + cl.off,
                  b->sb[0].nc + cl.off};
    view.box = b->sb[0].box + cl.off / 16;  // 4 doubles per 64-particle block
    BHIP(b, launch_prepare_stage(view, n, np, b->stream));
    Scalars s{};
    s.dtmin_bits = 0x7f800000u;
    BHIP(b, hipMemcpyAsync(b->d_sc + cloud, &s, sizeof(s), hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    b->uploaded[(size_t)cloud] = 1;
    b->uploaded_all = std::all_of(b->uploaded.begin(), b->uploaded.end(), [](char v) { return v != 0; });
    if (b->uploaded_all) {
        hipLaunchKernelGGL(k_brmax, dim3((unsigned)(b->P / 256)), dim3(256), 0, b->stream, b->sb[0].xs,
                           b->sb[0].ys, b->d_block_cloud, b->d_clouds, b->d_sc);
        BHIP(b, hipGetLastError());
        BHIP(b, hipStreamSynchronize(b->stream));
    }
    return DEMON_OK;
}

int demon_batch_download(demon_batch *b, int32_t cloud, double *x, double *y, double *vx, double *vy) {
    if (!b || cloud < 0 || cloud >= b->nclouds) return bfail(b, DEMON_ERR_ARG, "bad cloud index");
    if (!b->uploaded[(size_t)cloud]) return bfail(b, DEMON_ERR_ARG, "cloud not uploaded");
    BHIP(b, hipSetDevice(b->device));
    const BCloud &cl = b->hc[(size_t)cloud];
    std::vector<double> h((size_t)cl.npad);
    double *dst[4] = {x, y, vx, vy};
    const double *src[4] = {b->x, b->y, b->vx, b->vy};
    BHIP(b, hipStreamSynchronize(b->stream));
    for (int a = 0; a < 4; ++a) {
        if (!dst[a]) continue;
        BHIP(b, hipMemcpy(h.data(), src[a] + cl.off, sizeof(double) * cl.npad, hipMemcpyDeviceToHost));
        for (int64_t k = 0; k < cl.n; ++k) dst[a][b->perm[(size_t)cloud][(size_t)k]] = h[(size_t)k];
    }
    return DEMON_OK;
}

// Steps per captured graph: a call of s steps launches ceil(s / 4) graphs;
// the steps beyond s are no-ops (k_bboundary leaves every cloud inactive once
// its remaining steps are used up, and every kernel is gated on that).
constexpr int kBatchGraphSteps = 4;

int demon_batch_steps(demon_batch *b, int64_t steps, double dt) {
    if (!b || steps < 0 || !(dt > 0.0)) return bfail(b, DEMON_ERR_ARG, "bad steps/dt");
    if (!b->uploaded_all) return bfail(b, DEMON_ERR_ARG, "upload every cloud first");
    BHIP(b, hipSetDevice(b->device));
    b->tev_used = 0;
    b->tkind.clear();
    b->pair_ms = b->update_ms = 0.0;
    b->pair_launches = b->update_launches = 0;
    hipLaunchKernelGGL(k_bsetup, dim3((b->nclouds + 255) / 256), dim3(256), 0, b->stream, b->d_sc,
                       b->nclouds, dt, (unsigned long long)steps);
    BHIP(b, hipGetLastError());
    BHIP(b, hipEventRecord(b->ev[0], b->stream));
    if (b->use_graph) {
        if (!b->graph) {
            hipGraph_t g = nullptr;
            BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            int rc = DEMON_OK;
            for (int k = 0; k < kBatchGraphSteps && !rc; ++k) rc = enqueue_bstep(b, false);
            const hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return bfail(b, DEMON_ERR_HIP, "graph capture failed");
            const hipError_t e2 = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e2 != hipSuccess) return bfail(b, DEMON_ERR_HIP, "graph instantiate failed");
        }
        for (int64_t s = 0; s < steps; s += kBatchGraphSteps) BHIP(b, hipGraphLaunch(b->graph, b->stream));
    } else {
        for (int64_t s = 0; s < steps; ++s) {
            const int rc = enqueue_bstep(b, b->timing);
            if (rc) return rc;
        }
    }
    hipLaunchKernelGGL(k_bboundary, dim3((b->nclouds + 255) / 256), dim3(256), 0, b->stream, b->d_sc,
                       b->nclouds, 1, b->d_counters);
    BHIP(b, hipGetLastError());
    BHIP(b, hipEventRecord(b->ev[1], b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    float ms = 0.0f;
    BHIP(b, hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    b->last_ms = ms;
    for (size_t k = 0; k + 1 < b->tev_used; k += 2) {
        float km = 0.0f;
        BHIP(b, hipEventElapsedTime(&km, b->tev[k], b->tev[k + 1]));
        if (b->tkind[k / 2] == 0) {
            b->pair_ms += km;
            b->pair_launches += 1;
        } else {
            b->update_ms += km;
            b->update_launches += 1;
        }
    }
    std::vector<Scalars> s((size_t)b->nclouds);
    BHIP(b, hipMemcpy(s.data(), b->d_sc, sizeof(Scalars) * s.size(), hipMemcpyDeviceToHost));
    for (const Scalars &v : s)
        if (v.err == 2)
            return bfail(b, DEMON_ERR_STATE, "dt check: two tested particles of a cloud coincide");
    return DEMON_OK;
}

int demon_batch_cloud_stats(const demon_batch *b, int32_t cloud, demon_stats *out) {
    if (!b || !out || cloud < 0 || cloud >= b->nclouds) return bfail(nullptr, DEMON_ERR_ARG, "bad arguments");
    Scalars s{};
    if (hipMemcpy(&s, b->d_sc + cloud, sizeof(s), hipMemcpyDeviceToHost) != hipSuccess)
        return bfail(const_cast<demon_batch *>(b), DEMON_ERR_HIP, "hipMemcpy failed");
    out->steps = (int64_t)s.steps_done;
    out->dt_reductions = (int64_t)s.dt_reductions;
    out->force_evals = 4 * (int64_t)s.steps_done;
    out->current_time = s.t;
    out->last_dt = s.dt_step;
    return DEMON_OK;
}

int demon_batch_use_graph(demon_batch *b, int32_t enable) {
    if (!b) return bfail(nullptr, DEMON_ERR_ARG, "batch is NULL");
    b->use_graph = enable != 0;
    if (b->graph) {
        (void)hipStreamSynchronize(b->stream);
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    return DEMON_OK;
}

int demon_batch_set_timing(demon_batch *b, int32_t enable) {
    if (!b) return bfail(nullptr, DEMON_ERR_ARG, "batch is NULL");
    b->timing = enable != 0;
    return DEMON_OK;
}

int demon_batch_kernel_ms(const demon_batch *b, double *pair_ms, int64_t *pair_launches, double *update_ms,
                          int64_t *update_launches) {
    if (!b) return bfail(nullptr, DEMON_ERR_ARG, "batch is NULL");
    if (pair_ms) *pair_ms = b->pair_ms;
    if (pair_launches) *pair_launches = b->pair_launches;
    if (update_ms) *update_ms = b->update_ms;
    if (update_launches) *update_launches = b->update_launches;
    return DEMON_OK;
}

int demon_batch_pair_counts(const demon_batch *b, int64_t *near_pairs, int64_t *tested_pairs) {
    if (!b) return bfail(nullptr, DEMON_ERR_ARG, "batch is NULL");
    std::vector<Scalars> s((size_t)b->nclouds);
    if (hipMemcpy(s.data(), b->d_sc, sizeof(Scalars) * s.size(), hipMemcpyDeviceToHost) != hipSuccess)
        return bfail(const_cast<demon_batch *>(b), DEMON_ERR_HIP, "hipMemcpy pair counts failed");
    unsigned long long near = 0, tested = 0;
    for (const Scalars &v : s)
        for (int k = 0; k < kCountSlots; ++k) {
            near += v.pair_counts[k][0];
            tested += v.pair_counts[k][1];
        }
    if (near_pairs) *near_pairs = (int64_t)near;
    if (tested_pairs) *tested_pairs = (int64_t)tested;
    return DEMON_OK;
}

int demon_batch_last_ms(const demon_batch *b, double *ms) {
    if (!b || !ms) return bfail(nullptr, DEMON_ERR_ARG, "NULL argument");
    *ms = b->last_ms;
    return DEMON_OK;
}

int demon_batch_set_shape(demon_batch *b, int32_t bucket, int32_t i_per_lane, int32_t jslice) {
    if (!b || bucket < 0 || bucket > 31) return bfail(b, DEMON_ERR_ARG, "bad bucket");
    if (i_per_lane == DEMON_BATCH_NEWTON3) {
        if (jslice < 1) return bfail(b, DEMON_ERR_ARG, "Newton-3: jslice (superblock offsets per item) >= 1");
    } else if ((i_per_lane != 1 && i_per_lane != 2 && i_per_lane != 4) || jslice < 64 || jslice % 64) {
        return bfail(b, DEMON_ERR_ARG, "i_per_lane in {0 (Newton-3), 1, 2, 4}, gather jslice a multiple of 64");
    }
    b->bucket_R[bucket] = i_per_lane;
    b->bucket_js[bucket] = jslice;
    return rebuild_items(b);
}

int demon_batch_get_shape(const demon_batch *b, int32_t bucket, int32_t *i_per_lane, int32_t *jslice,
                          double *tuned_us) {
    if (!b || bucket < 0 || bucket > 31) return bfail(nullptr, DEMON_ERR_ARG, "bad bucket");
    if (i_per_lane) *i_per_lane = b->bucket_R[bucket];
    if (jslice) *jslice = b->bucket_js[bucket];
    if (tuned_us) *tuned_us = b->bucket_us[bucket];
    return DEMON_OK;
}

// Autotuner: for every N bucket present, times the all-pairs pass of that
// bucket's clouds (ungated, on the uploaded state) for each candidate: the
// gather form (R in {1,2,4}, j slice 64..8192) and the Newton-3 form (1, 2,
// 4, ... superblock offsets per item), and keeps the fastest.  The Newton-3
// form also adds its reactions in the update kernel (one read per evaluated
// tile row), which the timing leaves out.
int demon_batch_tune(demon_batch *b, int32_t reps) {
    if (!b || reps < 1) return bfail(b, DEMON_ERR_ARG, "bad tune arguments");
    if (!b->uploaded_all) return bfail(b, DEMON_ERR_ARG, "upload every cloud first");
    BHIP(b, hipSetDevice(b->device));
    std::vector<int> buckets;
    for (const BCloud &cl : b->hc) {
        const int bk = bucket_of(cl.npad);
        if (std::find(buckets.begin(), buckets.end(), bk) == buckets.end()) buckets.push_back(bk);
    }
    for (int bk : buckets) {
        std::vector<std::pair<int, int>> cand;  // (R, js)
        for (int R : {1, 2, 4})
            for (int js = 64; js <= 8192; js *= 2) {
                if (js > (int64_t(1) << (bk + 1)) * 2) break;
                cand.push_back({R, js});
            }
        const int32_t half_max = (int32_t)((int64_t(1) << (bk + 1)) / kSuper / 2);
        for (int span = 1; span <= half_max + 1 && half_max <= kMaxHalf; span *= 2) {
            cand.push_back({0, span});
            if (span > half_max) break;
        }
        float best = std::numeric_limits<float>::infinity();
        int bestR = b->bucket_R[bk], bestJ = b->bucket_js[bk];
        for (const auto &cj : cand) {
            const int R = cj.first, js = cj.second;
            b->bucket_R[bk] = R;
            b->bucket_js[bk] = js;
            int rc = rebuild_items(b);
            if (rc) return rc;
            // the items of this bucket's clouds only
            std::vector<std::pair<int64_t, BItem>> all;
            for (int32_t c = 0; c < b->nclouds; ++c) {
                const BCloud &cl = b->hc[(size_t)c];
                if (bucket_of(cl.npad) == bk) cloud_items(cl, c, all);
            }
            std::vector<BItem> items;
            for (auto &e : all) items.push_back(e.second);
            BItem *d_it = nullptr;
            BHIP(b, hipMalloc((void **)&d_it, sizeof(BItem) * items.size()));
            BHIP(b, hipMemcpyAsync(d_it, items.data(), sizeof(BItem) * items.size(), hipMemcpyHostToDevice,
                                   b->stream));
            // one warm-up pass, then `reps` passes back to back between two
            // events: the GPU's time per pass, without the host's launch
            // latency between an event and its kernel
            float ms = 0.0f;
            int ok = DEMON_OK;
            for (int r = 0; r <= reps && !ok; ++r) {
                if (r == 1) {
                    if (hipStreamSynchronize(b->stream) != hipSuccess) ok = bfail(b, DEMON_ERR_HIP, "tune sync");
                    hipEventRecord(b->ev[0], b->stream);
                }
                // the persistent kernels' item counters start at zero
                hipMemsetAsync(b->d_counters, 0, sizeof(unsigned int) * kGroups, b->stream);
                if (!ok) ok = enqueue_bpairs(b, b->sb[0], 0, shape_index(R), d_it, (int32_t)items.size());
            }
            hipEventRecord(b->ev[1], b->stream);
            if (hipStreamSynchronize(b->stream) != hipSuccess && !ok) ok = bfail(b, DEMON_ERR_HIP, "tune sync");
            if (!ok) hipEventElapsedTime(&ms, b->ev[0], b->ev[1]);
            (void)hipFree(d_it);
            if (ok) return ok;
            const float us = ms / reps * 1000.0f;
            if (us < best) {
                best = us;
                bestR = R;
                bestJ = js;
            }
        }
        b->bucket_R[bk] = bestR;
        b->bucket_js[bk] = bestJ;
        b->bucket_us[bk] = best;
    }
    BHIP(b, hipMemsetAsync(b->d_counters, 0, sizeof(unsigned int) * kGroups, b->stream));
    return rebuild_items(b);
}

}  // extern "C"
