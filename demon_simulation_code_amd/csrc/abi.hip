// abi.hip -- the C ABI of libdemon_gpu.so (declared in include/demon_gpu.h):
// context lifetime, host<->device transfers in the caller's particle order,
// and the launch sequence of Runge_Kutta4::moveParticles.
//
// Device layout (all FP64 SoA of npad = roundup(n, 1024) entries, in the
// device order -- a Morton order of the uploaded positions):
//   x y vx vy q m              base state and per-particle constants
//   sb[0], sb[1]               ping-pong stage records {xs ys vxs vys nc} and
//                              their bounding-box tables (npad/16 doubles)
//   sx                         scratch stage record for demon_force_stage
//   acc1[4], acc23[4]          RK increments (k l m n) of stage 1, of 2 (+3)
//   fout[2]                    force output of demon_force_stage
//   part                       all-pairs partial sums of the i side: per j
//                              slice [slices][2][npad] (gather kernel) or per
//                              work item [items][2][256] (Newton-3 kernel)
//   react[nsb/2][2], rflag     Newton-3 reactions and tile flags (own buffers,
//                              with the tile lists and the work plan)
// plus orig[npad] (caller index of each device slot) and the Scalars block.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/demon_gpu.h"
#include "internal.h"
#include "philox.h"

using namespace demon;

struct demon_ctx {
    std::string err;
    demon_params p{};
    int64_t n = 0, npad = 0;
    hipStream_t stream = nullptr;
    double *d_arena = nullptr;   // 32 arrays of npad doubles
    double *d_part = nullptr;    // part_doubles(npad)
    int32_t *d_orig = nullptr;
    Scalars *d_sc = nullptr;
    double *x = nullptr, *y = nullptr, *vx = nullptr, *vy = nullptr, *q = nullptr, *m = nullptr;
    float *qf = nullptr;  // the charges rounded to FP32 (FP32 pair arithmetic)
    StageBuf sb[2]{}, sx{};
    double *acc1 = nullptr, *acc23 = nullptr, *fout = nullptr;
    std::vector<int64_t> perm;   // perm[k] = caller index stored in device slot k
    std::vector<double> h_tmp;   // host staging (npad * 4)
    bool uploaded = false;
    int32_t i_per_lane = 1;
    int32_t slices_req = 0;      // 0 = automatic
    int32_t slices = 1, jslice = 0;  // gather kernel
    int32_t adj_slices = 1, adj_jslice = 0;  // adjoint pair kernel
    int32_t pair_mode = DEMON_PAIRS_GATHER;
    int32_t n3_slices = 1;       // Newton-3 kernel: work items per superblock row when every tile is live
    double *d_react = nullptr;   // Newton-3 reactions [nsb/2][2][npad]
    uint32_t *d_rflag = nullptr; // Newton-3 tile flags [nsb][nsb/2]
    int32_t *d_n3i = nullptr;    // Newton-3 tile lists and item table (n3_ints)
    N3Plan *d_plan = nullptr;    // Newton-3 work plan
    double *d_exptab = nullptr;  // Newton-3 kernel's exp table e^{-k/64}
    int32_t cus = 256;           // compute units of the device
    bool use_graph = false;
    int32_t order = 4;           // integrator of the captured step: 4 (RK4) or 2 (RK2)
    hipGraphExec_t graph = nullptr;
    bool timing = true;
    std::vector<hipEvent_t> ev;  // pairs of events around timed launches (plain launches only)
    std::vector<uint8_t> ev_kind;// what each pair brackets (TimedKind)
    size_t ev_used = 0;
    hipEvent_t ev_call[2]{};
    double last_pair_ms = 0.0, last_total_ms = 0.0;
    double last_update_ms = 0.0, last_list_ms = 0.0;
    int64_t last_pair_launches = 0, last_update_launches = 0;
    int64_t force_evals = 0;
    uint64_t rng_step = 0;       // thermal RNG step counter; survives re-uploads
    // trajectory fit (demon_fit_loss_grad): one arena, grown on demand
    double *d_fit = nullptr;
    size_t fit_doubles = 0;
    double last_fwd_ms = 0.0, last_bwd_ms = 0.0;
    double last_adj_ms = 0.0;         // the last fit's adjoint pair launches (HIP events)
    int64_t last_adj_launches = 0;
    hipEvent_t ev_fit[3]{};
    // analysis scratch (demon_kinetic_energy, demon_pair_histogram), grown on demand
    double *d_an = nullptr;
    size_t an_doubles = 0;
};

namespace {

thread_local std::string g_err;

int fail(demon_ctx *c, int code, const std::string &msg) {
    if (c)
        c->err = msg;
    else
        g_err = msg;
    return code;
}

int hip_fail(demon_ctx *c, hipError_t e, const char *what) {
    return fail(c, DEMON_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(c, expr)                                        \
    do {                                                        \
        hipError_t e_ = (expr);                                 \
        if (e_ != hipSuccess) return hip_fail((c), e_, #expr);  \
    } while (0)

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

// Every ForceFlag of the reference (Force.h:75-91).
constexpr int32_t kAllForces = 0x7FFF;

// The force constants of a parameter block as the kernels read them.
ForceParams force_params(const demon_params &p) {
    ForceParams f{};
    f.confine = p.confine; f.gamma = p.drag_gamma; f.heat = p.heat_val;
    f.bfield = p.magnetic_field; f.drag_scale = p.drag_scale;
    f.confine_x = p.confine_x; f.confine_y = p.confine_y; f.void_decay = p.void_decay;
    f.rot_inner = p.rot_inner; f.rot_outer = p.rot_outer; f.rot_const = p.rot_const;
    f.drive_const = p.drive_const; f.drive_amplitude = p.drive_amplitude; f.drive_shift = p.drive_shift;
    f.electric_field = p.electric_field; f.plasma_radius = p.plasma_radius;
    f.vert_electric_field = p.vert_electric_field; f.vert_decay = p.vert_decay;
    f.gravity = p.gravity;
    f.heat_in = p.heat_in; f.heat_out = p.heat_out; f.heat_radius = p.heat_radius;
    f.heat_scale = p.heat_scale; f.heat_offset = p.heat_offset;
    f.seed = p.seed;
    return f;
}

int check_params(demon_ctx *c, const demon_params *p) {
    if (!p) return fail(c, DEMON_ERR_ARG, "params is NULL");
    if (p->forces & ~kAllForces)
        return fail(c, DEMON_ERR_UNSUPPORTED, "unknown force flags (the reference's ForceFlag bits are 1..16384)");
    const double consts[] = {p->magnetic_field, p->drag_scale, p->confine_x, p->confine_y, p->void_decay,
                             p->rot_inner, p->rot_outer, p->rot_const, p->drive_const, p->drive_amplitude,
                             p->drive_shift, p->electric_field, p->plasma_radius, p->vert_electric_field,
                             p->vert_decay, p->gravity, p->heat_in, p->heat_out, p->heat_radius,
                             p->heat_scale, p->heat_offset};
    for (double v : consts)
        if (std::isnan(v)) return fail(c, DEMON_ERR_ARG, "force constants must not be NaN");
    if (!std::isfinite(p->magnetic_field) || !std::isfinite(p->drag_scale))
        return fail(c, DEMON_ERR_ARG, "magnetic field and drag scale must be finite");
    if ((p->forces & DEMON_FORCE_SHIELDED_COULOMB) && !(p->shielding > 0.0 && std::isfinite(p->shielding)))
        return fail(c, DEMON_ERR_ARG, "shielding constant must be positive and finite");
    if (p->precision != DEMON_PRECISION_FP64 && p->precision != DEMON_PRECISION_FP32_PAIR)
        return fail(c, DEMON_ERR_ARG, "unknown precision");
    return DEMON_OK;
}

// j slices of at least min_len particles, about `want` of them: whole
// 64-particle box blocks, or a quarter or a half of one (a slice never
// straddles two blocks, whose boxes the kernels test one at a time).
void slice_shape(int64_t npad, int64_t want, int64_t min_len, int32_t *slices, int32_t *jslice) {
    const int64_t s = std::max<int64_t>(
        1, std::min<int64_t>({want, (int64_t)kMaxSlices, std::max<int64_t>(1, npad / min_len)}));
    const int64_t need = (npad + s - 1) / s;
    const int64_t js = need <= kBoxBlock / 4 ? kBoxBlock / 4 : need <= kBoxBlock / 2 ? kBoxBlock / 2
                                                                                      : round_up(need, kBoxBlock);
    *slices = (int32_t)((npad + js - 1) / js);
    *jslice = (int32_t)js;
}

// Number of j slices: about 32768 waves (32 per SIMD of the 256 CUs), so
// that the uneven per-wave work left by the box culling evens out.  The
// gather kernel's slices go down to 16 j: in a cloud of few blocks a wave
// walks its chunks one scalar-load round trip after another with its SIMD
// to itself, and shorter slices put more waves on the chip (N = 1024: 27 us
// per launch with 16 slices of 64 j, 43 us with 8 of 128).  The adjoint
// pair kernel walks whole box blocks: its slices are at least 64 j long.
void choose_slices(demon_ctx *c) {
    const int64_t waves_per_slice = c->npad / (64 * c->i_per_lane);
    const int64_t want = c->slices_req > 0 ? c->slices_req : (32768 + waves_per_slice - 1) / waves_per_slice;
    slice_shape(c->npad, want, kBoxBlock / 4, &c->slices, &c->jslice);
    slice_shape(c->npad, want, kBoxBlock, &c->adj_slices, &c->adj_jslice);
    // Newton-3: a launch is cut into about n3_slices * nsb work items of
    // consecutive live tiles (one workgroup of four waves each), i.e.
    // n3_slices items per row when every one of its nsb/2 + 1 tiles is live;
    // about 32768 items (32 per workgroup of the persistent grid), so that
    // the last items of a dense launch are short
    const int64_t nsb = c->npad / kSuper, offsets = nsb / 2 + 1;
    int64_t s3 = c->slices_req > 0 ? c->slices_req : (32768 + nsb - 1) / nsb;
    c->n3_slices = (int32_t)std::max<int64_t>(1, std::min<int64_t>({s3, (int64_t)kMaxSlices, offsets}));
}

// Newton-3 items of one launch: at most target + nsb (every row's last item
// may be short), target = n3_slices * nsb <= kMaxSlices * nsb.
int64_t n3_max_items(int64_t npad) { return (int64_t)(kMaxSlices + 1) * (npad / kSuper); }

// Doubles of the partial-sum buffer: j slices of the gather kernel or work
// items of the Newton-3 kernel.
size_t part_doubles(int64_t npad) {
    return std::max<size_t>((size_t)npad * 2 * kMaxSlices, (size_t)n3_max_items(npad) * 2 * kSuper);
}

// int32 entries of the Newton-3 tile lists and item table: list
// [nsb][nsb/2 + 1], cnt [nsb], first [nsb + 1].
size_t n3_ints(int64_t npad) {
    const size_t nsb = (size_t)(npad / kSuper);
    return nsb * (nsb / 2 + 1) + nsb + (nsb + 1);
}

// Device memory of the Newton-3 reaction buffers, tile lists and plan.
size_t n3_bytes(int64_t npad) {
    const size_t nsb = (size_t)(npad / kSuper), half = nsb / 2;
    return sizeof(double) * half * 2 * (size_t)npad + sizeof(uint32_t) * half * nsb + sizeof(int32_t) * n3_ints(npad) +
           sizeof(N3Plan);
}

// Newton-3 is the default while its reaction buffers stay under this size,
// from this many padded particles on: below it a cloud has too few
// superblock rows to fill the chip, and the gather kernel, with its j
// slices, is faster (measured with graphs on one MI355X: 2.1x at N = 1024
// and 2048, 1.35x at 4096; Newton-3 1.27x faster at 8192, 1.75x at 16384).
constexpr size_t kN3DefaultBytes = (size_t)16 << 30;
constexpr int64_t kN3MinNpad = 6144;

int ensure_n3(demon_ctx *c) {
    if (c->d_react) return DEMON_OK;
    const size_t nsb = (size_t)(c->npad / kSuper), half = nsb / 2;
    if (half > (size_t)kMaxHalf)
        return fail(c, DEMON_ERR_UNSUPPORTED, "the Newton-3 kernel supports n <= 524288; use DEMON_PAIRS_GATHER");
    const size_t rbytes = sizeof(double) * half * 2 * (size_t)c->npad;
    if (hipMalloc((void **)&c->d_react, rbytes) != hipSuccess) {
        c->d_react = nullptr;
        return fail(c, DEMON_ERR_ALLOC, "hipMalloc of the Newton-3 reaction buffers failed");
    }
    if (hipMalloc((void **)&c->d_rflag, sizeof(uint32_t) * half * nsb) != hipSuccess ||
        hipMalloc((void **)&c->d_n3i, sizeof(int32_t) * n3_ints(c->npad)) != hipSuccess ||
        hipMalloc((void **)&c->d_plan, sizeof(N3Plan)) != hipSuccess ||
        hipMalloc((void **)&c->d_exptab, sizeof(double) * kExpTabN) != hipSuccess) {
        for (void *ptr : {(void *)c->d_react, (void *)c->d_rflag, (void *)c->d_n3i, (void *)c->d_plan,
                          (void *)c->d_exptab})
            if (ptr) (void)hipFree(ptr);
        c->d_react = nullptr;
        c->d_rflag = nullptr;
        c->d_n3i = nullptr;
        c->d_plan = nullptr;
        c->d_exptab = nullptr;
        return fail(c, DEMON_ERR_ALLOC, "hipMalloc of the Newton-3 tile flags, lists or plan failed");
    }
    // no tile evaluated yet: the stage kernels must read no reaction and no
    // item; the plan's counters start at zero
    // (on the context's stream, which is non-blocking: null-stream fills would
    // not be ordered before its kernels)
    HIP_TRY(c, hipMemsetAsync(c->d_rflag, 0, sizeof(uint32_t) * half * nsb, c->stream));
    HIP_TRY(c, hipMemsetAsync(c->d_n3i, 0, sizeof(int32_t) * n3_ints(c->npad), c->stream));
    HIP_TRY(c, hipMemsetAsync(c->d_plan, 0, sizeof(N3Plan), c->stream));
    double tab[kExpTabN];
    exp_table_host(tab);
    HIP_TRY(c, hipMemcpyAsync(c->d_exptab, tab, sizeof tab, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));  // tab is a local
    return DEMON_OK;
}

N3List n3_list(const demon_ctx *c) {
    const int64_t nsb = c->npad / kSuper;
    N3List l{};
    l.list = c->d_n3i;
    l.cnt = l.list + nsb * (nsb / 2 + 1);
    l.first = l.cnt + nsb;
    l.plan = c->d_plan;
    l.target = (int32_t)(c->n3_slices * nsb);
    l.max_items = (int32_t)n3_max_items(c->npad);
    l.cus = c->cus;
    l.exptab = c->d_exptab;
    return l;
}

PairSums pair_sums(const demon_ctx *c) {
    PairSums ps{};
    ps.part = c->d_part;
    if (c->pair_mode == DEMON_PAIRS_NEWTON3) {
        ps.slices = 0;
        ps.react = c->d_react;
        ps.rflag = c->d_rflag;
        ps.nsb = (int32_t)(c->npad / kSuper);
        ps.first = n3_list(c).first;
    } else {
        ps.slices = c->slices;
    }
    return ps;
}

uint64_t morton2(uint32_t a, uint32_t b) {
    auto spread = [](uint64_t v) {
        v &= 0xFFFFFFFFull;
        v = (v | (v << 16)) & 0x0000FFFF0000FFFFull;
        v = (v | (v << 8)) & 0x00FF00FF00FF00FFull;
        v = (v | (v << 4)) & 0x0F0F0F0F0F0F0F0Full;
        v = (v | (v << 2)) & 0x3333333333333333ull;
        v = (v | (v << 1)) & 0x5555555555555555ull;
        return v;
    };
    return spread(a) | (spread(b) << 1);
}

// Device order: sort by the Morton code of the positions on a 2^16 grid over
// their bounding box (ties and non-finite positions keep caller order).
void build_perm(demon_ctx *c, const double *x, const double *y) {
    const int64_t n = c->n;
    double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
    for (int64_t i = 0; i < n; ++i) {
        if (std::isfinite(x[i])) { x0 = std::min(x0, x[i]); x1 = std::max(x1, x[i]); }
        if (std::isfinite(y[i])) { y0 = std::min(y0, y[i]); y1 = std::max(y1, y[i]); }
    }
    const double sx = (x1 > x0) ? 65535.0 / (x1 - x0) : 0.0;
    const double sy = (y1 > y0) ? 65535.0 / (y1 - y0) : 0.0;
    std::vector<std::pair<uint64_t, int64_t>> key((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        uint32_t a = 0, b = 0;
        if (std::isfinite(x[i]) && std::isfinite(y[i])) {
            a = (uint32_t)std::min(65535.0, std::max(0.0, (x[i] - x0) * sx));
            b = (uint32_t)std::min(65535.0, std::max(0.0, (y[i] - y0) * sy));
        }
        key[(size_t)i] = {morton2(a, b), i};
    }
    std::sort(key.begin(), key.end());
    c->perm.resize((size_t)n);
    for (int64_t k = 0; k < n; ++k) c->perm[(size_t)k] = key[(size_t)k].second;
}

// host caller-order array -> device slot order (into h_tmp + off) -> device
int put(demon_ctx *c, double *dst, const double *src, double fill, size_t slot) {
    double *h = c->h_tmp.data() + slot * (size_t)c->npad;
    for (int64_t k = 0; k < c->n; ++k) h[k] = src ? src[c->perm[(size_t)k]] : fill;
    for (int64_t k = c->n; k < c->npad; ++k) h[k] = 0.0;
    HIP_TRY(c, hipMemcpyAsync(dst, h, sizeof(double) * c->npad, hipMemcpyHostToDevice, c->stream));
    return DEMON_OK;
}

// Launches timed with HIP events on the context stream: the all-pairs
// kernel, the Newton-3 tile-list kernel and the stage update.  Events are
// recorded around plain launches only, not into a captured graph.
enum TimedKind : uint8_t { kTimedPairs = 0, kTimedList = 1, kTimedUpdate = 2, kTimedAdjPairs = 3 };

void ev_reset(demon_ctx *c) {
    c->ev_used = 0;
    c->ev_kind.clear();
}

// Opens a timed launch when want is set (the event pair is closed by ev_close).
int ev_open_if(demon_ctx *c, bool want, bool *on) {
    *on = want;
    if (!*on) return DEMON_OK;
    if (c->ev_used + 2 > c->ev.size()) {
        for (int k = 0; k < 64; ++k) {
            hipEvent_t e;
            HIP_TRY(c, hipEventCreate(&e));
            c->ev.push_back(e);
        }
    }
    HIP_TRY(c, hipEventRecord(c->ev[c->ev_used], c->stream));
    return DEMON_OK;
}

int ev_open(demon_ctx *c, bool *on) { return ev_open_if(c, c->timing && !c->use_graph, on); }

int ev_close(demon_ctx *c, bool on, TimedKind kind) {
    if (!on) return DEMON_OK;
    HIP_TRY(c, hipEventRecord(c->ev[c->ev_used + 1], c->stream));
    c->ev_used += 2;
    c->ev_kind.push_back((uint8_t)kind);
    return DEMON_OK;
}

int enqueue_pairs(demon_ctx *c, const StageBuf &st, int gated, int dtcheck) {
    PairLaunch a{};
    a.xs = st.xs; a.ys = st.ys; a.nc = st.nc; a.q = c->q; a.qf = c->qf; a.orig = c->d_orig; a.box = st.box;
    a.part = c->d_part; a.sc = c->d_sc; a.n = c->n; a.npad = c->npad;
    a.jslice = c->jslice; a.slices = c->slices; a.i_per_lane = c->i_per_lane;
    a.dtcheck = dtcheck; a.gated = gated;
    a.fp32pair = c->p.precision == DEMON_PRECISION_FP32_PAIR;
    a.kappa = c->p.shielding;
    const double rc = 10.0 / c->p.shielding;
    a.rc2p = rc * rc * (1.0 + kCutoffSlack);
    bool on = false;
    int err;
    if (c->pair_mode == DEMON_PAIRS_NEWTON3) {
        a.react = c->d_react;
        a.rflag = c->d_rflag;
        a.lst = n3_list(c);
        if ((err = ev_open(c, &on))) return err;
        HIP_TRY(c, launch_n3_tiles(a, c->stream));
        if ((err = ev_close(c, on, kTimedList)) || (err = ev_open(c, &on))) return err;
        HIP_TRY(c, launch_pairs_n3(a, c->stream));
    } else {
        if ((err = ev_open(c, &on))) return err;
        HIP_TRY(c, launch_pairs(a, c->stream));
    }
    return ev_close(c, on, kTimedPairs);
}

int timed_update(demon_ctx *c, int stage, const UpdateLaunch &u) {
    bool on = false;
    int rc = ev_open(c, &on);
    if (rc) return rc;
    HIP_TRY(c, launch_update(stage, u, c->stream));
    return ev_close(c, on, kTimedUpdate);
}

UpdateLaunch make_update(demon_ctx *c, int stage, int dtcheck, double *store) {
    UpdateLaunch u{};
    // stage records: 1 reads sb0, 2 reads sb1, 3 reads sb0, 4 reads sb1;
    // RK2's second stage reads sb1
    const int cur = stage == kRk2Final ? 1 : (stage - 1) & 1;
    u.cur = c->sb[cur];
    u.nxt = c->sb[cur ^ 1];
    u.x = c->x; u.y = c->y; u.vx = c->vx; u.vy = c->vy; u.q = c->q; u.m = c->m;
    u.orig = c->d_orig; u.pairs = pair_sums(c);
    u.acc1 = c->acc1; u.acc23 = c->acc23; u.sc = c->d_sc;
    u.n = c->n; u.npad = c->npad; u.forces = c->p.forces; u.dtcheck = dtcheck;
    u.fp = force_params(c->p);
    u.store = store;
    return u;
}

// One step of the context's integrator (c->order): RK4 stages 1..4, or RK2's
// stage 1 and kRk2Final.  store (optional, RK4 only): 4 consecutive stage
// slots of 5*npad doubles.
int enqueue_step_ex(demon_ctx *c, int dtcheck, double *store) {
    HIP_TRY(c, launch_boundary(c->d_sc, 0, c->stream));
    const bool yk = c->p.forces & DEMON_FORCE_SHIELDED_COULOMB;
    if (c->order == 2) {
        for (int s : {1, (int)kRk2Final}) {
            if (yk) {
                const int rc = enqueue_pairs(c, c->sb[s == 1 ? 0 : 1], 1, s == 1 && dtcheck);
                if (rc) return rc;
            } else if (s == 1 && dtcheck) {
                return fail(c, DEMON_ERR_UNSUPPORTED,
                            "the dt check runs inside the all-pairs pass; enable ShieldedCoulomb");
            }
            const int rc = timed_update(c, s, make_update(c, s, dtcheck, nullptr));
            if (rc) return rc;
        }
        return DEMON_OK;
    }
    for (int s = 1; s <= 4; ++s) {
        const StageBuf &st = c->sb[(s - 1) & 1];
        if (yk) {
            const int rc = enqueue_pairs(c, st, 1, s == 1 && dtcheck);
            if (rc) return rc;
        } else if (s == 1 && dtcheck) {
            return fail(c, DEMON_ERR_UNSUPPORTED,
                        "the dt check runs inside the all-pairs pass; enable ShieldedCoulomb");
        }
        double *slot = store ? store + (size_t)(s - 1) * store_doubles(c->npad) : nullptr;
        const int rc = timed_update(c, s, make_update(c, s, dtcheck, slot));
        if (rc) return rc;
    }
    return DEMON_OK;
}

int enqueue_step(demon_ctx *c) { return enqueue_step_ex(c, c->p.dt_check, nullptr); }

// The captured step is destroyed only once the stream has drained: a replay
// still in flight must not lose its executable graph.
void drop_graph(demon_ctx *c) {
    if (c->graph) {
        if (c->stream) (void)hipStreamSynchronize(c->stream);
        (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
    }
}

// Steps per captured graph.  Steps beyond a call's count are no-ops (the
// kernels are gated on Scalars::active, which k_boundary clears when no step
// is left or t >= t_end), so a call of s steps launches ceil(s / kGraphSteps)
// graphs; for small clouds the per-launch host cost is the step's cost.
constexpr int kGraphSteps = 8;

// Device bytes the fit may spend on stage records kept from its forward sweep
// (the environment variable DEMON_FIT_KEEP_BYTES overrides it).
constexpr size_t kFitKeepBytes = (size_t)16 << 30;

int ensure_graph(demon_ctx *c) {
    if (c->graph) return DEMON_OK;
    hipGraph_t g = nullptr;
    HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = DEMON_OK;
    for (int k = 0; k < kGraphSteps && !rc; ++k) rc = enqueue_step(c);
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(c, e, "hipGraphInstantiate");
    return DEMON_OK;
}

__global__ void k_call_setup(Scalars *sc, double t, double t_end, double init_dt,
                             unsigned long long remaining) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        sc->t = t;
        sc->t_end = t_end;
        sc->init_dt = init_dt;
        sc->remaining = remaining;
        sc->active = 0;
    }
}

// Queues `steps` steps plus the closing boundary, then reads the scalars back.
int run_steps(demon_ctx *c, int64_t steps, Scalars *out) {
    ev_reset(c);
    HIP_TRY(c, hipEventRecord(c->ev_call[0], c->stream));
    if (c->use_graph) {
        const int rc = ensure_graph(c);
        if (rc) return rc;
        for (int64_t s = 0; s < steps; s += kGraphSteps) HIP_TRY(c, hipGraphLaunch(c->graph, c->stream));
    } else {
        for (int64_t s = 0; s < steps; ++s) {
            const int rc = enqueue_step(c);
            if (rc) return rc;
        }
    }
    HIP_TRY(c, launch_boundary(c->d_sc, 1, c->stream));
    HIP_TRY(c, hipEventRecord(c->ev_call[1], c->stream));
    HIP_TRY(c, hipMemcpyAsync(out, c->d_sc, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    HIP_TRY(c, hipEventElapsedTime(&ms, c->ev_call[0], c->ev_call[1]));
    c->last_total_ms += ms;
    for (size_t k = 0; k + 1 < c->ev_used; k += 2) {
        float pm = 0.0f;
        HIP_TRY(c, hipEventElapsedTime(&pm, c->ev[k], c->ev[k + 1]));
        switch (c->ev_kind[k / 2]) {
            case kTimedPairs: c->last_pair_ms += pm; c->last_pair_launches += 1; break;
            case kTimedList: c->last_list_ms += pm; break;
            default: c->last_update_ms += pm; c->last_update_launches += 1; break;
        }
    }
    c->force_evals += c->order * steps;
    c->rng_step = out->step;
    if (out->err == 2)
        return fail(c, DEMON_ERR_STATE,
                    "dt check: two tested particles coincide (distance 0); the reference's "
                    "modifyTimeStep would divide dt by 10 forever");
    return DEMON_OK;
}


// Device copies of the base state into / out of a 4*npad checkpoint slot.
int save_state(demon_ctx *c, double *slot) {
    const size_t b = sizeof(double) * c->npad;
    const double *src[4] = {c->x, c->y, c->vx, c->vy};
    for (int a = 0; a < 4; ++a)
        HIP_TRY(c, hipMemcpyAsync(slot + (size_t)a * c->npad, src[a], b, hipMemcpyDeviceToDevice, c->stream));
    return DEMON_OK;
}

// base state <- slot, and the stage-1 record rebuilt from it.
int restore_state(demon_ctx *c, const double *slot) {
    const size_t b = sizeof(double) * c->npad;
    double *dst[4] = {c->x, c->y, c->vx, c->vy};
    double *rec[4] = {c->sb[0].xs, c->sb[0].ys, c->sb[0].vxs, c->sb[0].vys};
    for (int a = 0; a < 4; ++a) {
        HIP_TRY(c, hipMemcpyAsync(dst[a], slot + (size_t)a * c->npad, b, hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(rec[a], slot + (size_t)a * c->npad, b, hipMemcpyDeviceToDevice, c->stream));
    }
    HIP_TRY(c, launch_prepare_stage(c->sb[0], c->n, c->npad, c->stream));
    HIP_TRY(c, hipMemsetAsync(&c->d_sc->rmax2_bits, 0, sizeof(unsigned long long), c->stream));
    HIP_TRY(c, launch_rmax(c->sb[0], c->n, c->d_sc, c->stream));
    return DEMON_OK;
}

int fit_arena(demon_ctx *c, size_t doubles) {
    if (doubles <= c->fit_doubles) return DEMON_OK;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (c->d_fit) (void)hipFree(c->d_fit);
    c->d_fit = nullptr;
    c->fit_doubles = 0;
    if (hipMalloc((void **)&c->d_fit, sizeof(double) * doubles) != hipSuccess)
        return fail(c, DEMON_ERR_ALLOC, "hipMalloc of the fit buffers failed");
    c->fit_doubles = doubles;
    return DEMON_OK;
}

double host_sum(demon_ctx *c, const double *dev, std::vector<double> &tmp, int *rc) {
    tmp.resize((size_t)c->npad);
    hipError_t e = hipMemcpyAsync(tmp.data(), dev, sizeof(double) * c->npad, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        *rc = hip_fail(c, e, "hipMemcpy");
        return 0.0;
    }
    double acc = 0.0;
    for (int64_t k = 0; k < c->n; ++k) acc += tmp[(size_t)k];
    return acc;
}

// Selects the integrator whose step is queued (and captured).
void set_order(demon_ctx *c, int32_t order) {
    if (c->order != order) {
        drop_graph(c);
        c->order = order;
    }
}

void reset_timing(demon_ctx *c) {
    c->last_pair_ms = c->last_total_ms = c->last_update_ms = c->last_list_ms = 0.0;
    c->last_pair_launches = c->last_update_launches = 0;
}

int rk_steps(demon_ctx *c, int32_t order, int64_t steps, double init_dt, double *current_time) {
    if (!c) return fail(nullptr, DEMON_ERR_ARG, "ctx is NULL");
    if (!c->uploaded) return fail(c, DEMON_ERR_ARG, "no state uploaded");
    if (!current_time || steps < 0 || !(init_dt > 0.0)) return fail(c, DEMON_ERR_ARG, "bad steps/dt/time");
    HIP_TRY(c, hipSetDevice(c->p.device));
    set_order(c, order);
    reset_timing(c);
    hipLaunchKernelGGL(k_call_setup, dim3(1), dim3(64), 0, c->stream, c->d_sc, *current_time,
                       std::numeric_limits<double>::infinity(), init_dt,
                       (unsigned long long)steps);
    HIP_TRY(c, hipGetLastError());
    Scalars s{};
    const int rc = run_steps(c, steps, &s);
    if (rc) return rc;
    *current_time = s.t;
    return DEMON_OK;
}

int rk_move(demon_ctx *c, int32_t order, double *current_time, double end_time, double init_dt,
            int64_t *steps_done) {
    if (!c) return fail(nullptr, DEMON_ERR_ARG, "ctx is NULL");
    if (!c->uploaded) return fail(c, DEMON_ERR_ARG, "no state uploaded");
    if (!current_time || !(init_dt > 0.0)) return fail(c, DEMON_ERR_ARG, "bad dt/time");
    HIP_TRY(c, hipSetDevice(c->p.device));
    set_order(c, order);
    reset_timing(c);
    Scalars s{};
    HIP_TRY(c, hipMemcpy(&s, c->d_sc, sizeof(Scalars), hipMemcpyDeviceToHost));
    const unsigned long long before = s.steps_done;
    hipLaunchKernelGGL(k_call_setup, dim3(1), dim3(64), 0, c->stream, c->d_sc, *current_time,
                       end_time, init_dt, ~0ull);
    HIP_TRY(c, hipGetLastError());
    double t = *current_time;
    for (;;) {
        // steps the while loop needs if dt stays init_dt (same accumulation)
        int64_t est = 0;
        for (double tt = t; tt < end_time && est < 4096; tt += init_dt) ++est;
        if (est == 0) break;
        const int rc = run_steps(c, est, &s);
        if (rc) return rc;
        if (!(s.t > t)) return fail(c, DEMON_ERR_STATE, "time does not advance (dt underflow)");
        t = s.t;
    }
    *current_time = t;
    if (steps_done) *steps_done = (int64_t)(s.steps_done - before);
    return DEMON_OK;
}

}  // namespace

extern "C" {

const char *demon_version(void) { return "demon-mi355x 0.1 (gfx950)"; }
int demon_abi_version(void) { return DEMON_ABI_VERSION; }

const char *demon_last_error(const demon_ctx *ctx) {
    return ctx ? ctx->err.c_str() : g_err.c_str();
}

int demon_create(demon_ctx **out, int64_t n, const demon_params *p) {
    if (!out) return fail(nullptr, DEMON_ERR_ARG, "ctx out-pointer is NULL");
    *out = nullptr;
    if (n < 2 || n > (int64_t)1 << 30) return fail(nullptr, DEMON_ERR_ARG, "n must be in [2, 2^30]");
    int rc = check_params(nullptr, p);
    if (rc) return rc;
    demon_ctx *c = new (std::nothrow) demon_ctx();
    if (!c) return fail(nullptr, DEMON_ERR_ALLOC, "host allocation failed");
    c->p = *p;
    c->n = n;
    c->npad = round_up(n, kPad);
    auto bail = [&](int code) {
        g_err = c->err;
        demon_destroy(c);
        return code;
    };
    hipError_t e = hipSetDevice(p->device);
    if (e != hipSuccess) return bail(hip_fail(c, e, "hipSetDevice"));
    e = hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, p->device);
    if (e != hipSuccess) return bail(hip_fail(c, e, "hipDeviceGetAttribute"));
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return bail(hip_fail(c, e, "hipStreamCreate"));
    const size_t P = (size_t)c->npad;
    e = hipMalloc((void **)&c->d_arena, sizeof(double) * P * 32);
    if (e != hipSuccess) return bail(fail(c, DEMON_ERR_ALLOC, "hipMalloc state arena failed"));
    e = hipMalloc((void **)&c->d_part, sizeof(double) * part_doubles(c->npad));
    if (e != hipSuccess) return bail(fail(c, DEMON_ERR_ALLOC, "hipMalloc partial sums failed"));
    e = hipMalloc((void **)&c->d_orig, sizeof(int32_t) * P);
    if (e != hipSuccess) return bail(fail(c, DEMON_ERR_ALLOC, "hipMalloc index map failed"));
    e = hipMalloc((void **)&c->d_sc, sizeof(Scalars));
    if (e != hipSuccess) return bail(fail(c, DEMON_ERR_ALLOC, "hipMalloc scalars failed"));
    // fills on the context's own (non-blocking) stream, ordered before its kernels
    e = hipMemsetAsync(c->d_arena, 0, sizeof(double) * P * 32, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->d_part, 0, sizeof(double) * part_doubles(c->npad), c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(c->d_sc, 0, sizeof(Scalars), c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return bail(hip_fail(c, e, "hipMemset"));
    double *b = c->d_arena;
    auto take = [&]() { double *r = b; b += P; return r; };
    c->x = take(); c->y = take(); c->vx = take(); c->vy = take(); c->q = take(); c->m = take();
    for (StageBuf *s : {&c->sb[0], &c->sb[1], &c->sx}) {
        s->xs = take(); s->ys = take(); s->vxs = take(); s->vys = take(); s->nc = take();
    }
    c->acc1 = b; b += 4 * P;
    c->acc23 = b; b += 4 * P;
    c->fout = b; b += 2 * P;
    // one more array holds the three stages' box tables (P/16 doubles each)
    // and the FP32 charges (P/2 doubles)
    for (StageBuf *s : {&c->sb[0], &c->sb[1], &c->sx}) {
        s->box = b;
        b += P / 16;
    }
    c->qf = reinterpret_cast<float *>(b);
    b += P / 2;
    for (int k = 0; k < 2; ++k) {
        e = hipEventCreate(&c->ev_call[k]);
        if (e != hipSuccess) return bail(hip_fail(c, e, "hipEventCreate"));
    }
    c->h_tmp.assign(P * 4, 0.0);
    choose_slices(c);
    if (c->npad >= kN3MinNpad && n3_bytes(c->npad) <= kN3DefaultBytes && c->npad / kSuper / 2 <= kMaxHalf) {
        if (ensure_n3(c)) return bail(DEMON_ERR_ALLOC);
        c->pair_mode = DEMON_PAIRS_NEWTON3;
    }
    *out = c;
    return DEMON_OK;
}

void demon_destroy(demon_ctx *c) {
    if (!c) return;
    drop_graph(c);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->ev_call)
        if (e) (void)hipEventDestroy(e);
    if (c->d_arena) (void)hipFree(c->d_arena);
    if (c->d_part) (void)hipFree(c->d_part);
    if (c->d_orig) (void)hipFree(c->d_orig);
    if (c->d_sc) (void)hipFree(c->d_sc);
    if (c->d_fit) (void)hipFree(c->d_fit);
    if (c->d_react) (void)hipFree(c->d_react);
    if (c->d_rflag) (void)hipFree(c->d_rflag);
    if (c->d_n3i) (void)hipFree(c->d_n3i);
    if (c->d_plan) (void)hipFree(c->d_plan);
    if (c->d_exptab) (void)hipFree(c->d_exptab);
    if (c->d_an) (void)hipFree(c->d_an);
    for (hipEvent_t e : c->ev_fit)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    (void)hipGetLastError();  // the teardown's errors are not the next call's
    delete c;
}

int demon_set_params(demon_ctx *c, const demon_params *p) {
    if (!c) return fail(nullptr, DEMON_ERR_ARG, "ctx is NULL");
    int rc = check_params(c, p);
    if (rc) return rc;
    if (p->device != c->p.device) return fail(c, DEMON_ERR_ARG, "device cannot change");
    c->p = *p;
    drop_graph(c);
    return DEMON_OK;
}

int demon_get_params(const demon_ctx *c, demon_params *p) {
    if (!c || !p) return fail(nullptr, DEMON_ERR_ARG, "NULL argument");
    *p = c->p;
    return DEMON_OK;
}

int demon_upload(demon_ctx *c, const double *x, const double *y, const double *vx,
                 const double *vy, const double *charge, const double *mass) {
    if (!c) return fail(nullptr, DEMON_ERR_ARG, "ctx is NULL");
    if (!x || !y || !charge || !mass) return fail(c, DEMON_ERR_ARG, "x, y, charge and mass are required");
    HIP_TRY(c, hipSetDevice(c->p.device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    build_perm(c, x, y);
    std::vector<int32_t> orig((size_t)c->npad);
    for (int64_t k = 0; k < c->npad; ++k)
        orig[(size_t)k] = k < c->n ? (int32_t)c->perm[(size_t)k] : (int32_t)(0x40000000 + k);
    HIP_TRY(c, hipMemcpyAsync(c->d_orig, orig.data(), sizeof(int32_t) * c->npad, hipMemcpyHostToDevice, c->stream));
    int rc;
    // each put uses its own staging slot until the stream is synchronised
    if ((rc = put(c, c->x, x, 0.0, 0)) || (rc = put(c, c->y, y, 0.0, 1)) ||
        (rc = put(c, c->vx, vx, 0.0, 2)) || (rc = put(c, c->vy, vy, 0.0, 3)))
        return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if ((rc = put(c, c->q, charge, 0.0, 0)) || (rc = put(c, c->m, mass, 1.0, 1))) return rc;
    HIP_TRY(c, launch_round_fp32(c->q, c->qf, c->npad, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    // stage-1 record = base state
    const size_t bytes = sizeof(double) * c->npad;
    HIP_TRY(c, hipMemcpyAsync(c->sb[0].xs, c->x, bytes, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->sb[0].ys, c->y, bytes, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->sb[0].vxs, c->vx, bytes, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->sb[0].vys, c->vy, bytes, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, launch_prepare_stage(c->sb[0], c->n, c->npad, c->stream));
    HIP_TRY(c, launch_prepare_stage(c->sb[1], c->n, c->npad, c->stream));
    HIP_TRY(c, launch_prepare_stage(c->sx, 0, c->npad, c->stream));
    Scalars s{};
    s.dtmin_bits = 0x7f800000u;
    s.step = c->rng_step;  // a re-upload never replays earlier thermal draws
    HIP_TRY(c, hipMemcpyAsync(c->d_sc, &s, sizeof(s), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, launch_rmax(c->sb[0], c->n, c->d_sc, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->uploaded = true;
    c->force_evals = 0;
    return DEMON_OK;
}

int demon_download(demon_ctx *c, double *x, double *y, double *vx, double *vy) {
    if (!c) return fail(nullptr, DEMON_ERR_ARG, "ctx is NULL");
    if (!c->uploaded) return fail(c, DEMON_ERR_ARG, "no state uploaded");
    HIP_TRY(c, hipSetDevice(c->p.device));
    double *dst[4] = {x, y, vx, vy};
    const double *src[4] = {c->x, c->y, c->vx, c->vy};
    for (int a = 0; a < 4; ++a) {
        if (!dst[a]) continue;
        HIP_TRY(c, hipMemcpyAsync(c->h_tmp.data(), src[a], sizeof(double) * c->npad,
                                  hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        for (int64_t k = 0; k < c->n; ++k) dst[a][c->perm[(size_t)k]] = c->h_tmp[(size_t)k];
    }
    return DEMON_OK;
}

int demon_force_stage_at(demon_ctx *c, int32_t forces, int32_t stage, uint64_t step, double t,
                         const double *xs, const double *ys, const double *vxs,
                         const double *vys, double *fx, double *fy) {
    if (!c) return fail(nullptr, DEMON_ERR_ARG, "ctx is NULL");
    if (!c->uploaded) return fail(c, DEMON_ERR_ARG, "demon_upload must provide charge and mass first");
    if (!xs || !ys || !fx || !fy) return fail(c, DEMON_ERR_ARG, "NULL array");
    const int32_t needs_v = DEMON_FORCE_DRAG | DEMON_FORCE_MAGNETIC | DEMON_FORCE_TIME_VARYING_DRAG;
    if ((forces & needs_v) && (!vxs || !vys))
        return fail(c, DEMON_ERR_ARG, "drag and magnetic forces need velocities");
    if (!std::isfinite(t)) return fail(c, DEMON_ERR_ARG, "stage time must be finite");
    if (forces & ~c->p.forces) return fail(c, DEMON_ERR_ARG, "force not enabled in the context");
    if (stage < 1 || stage > 4) return fail(c, DEMON_ERR_ARG, "stage must be 1..4");
    HIP_TRY(c, hipSetDevice(c->p.device));
    int rc;
    if ((rc = put(c, c->sx.xs, xs, 0.0, 0)) || (rc = put(c, c->sx.ys, ys, 0.0, 1)) ||
        (rc = put(c, c->sx.vxs, vxs, 0.0, 2)) || (rc = put(c, c->sx.vys, vys, 0.0, 3)))
        return rc;
    HIP_TRY(c, launch_prepare_stage(c->sx, c->n, c->npad, c->stream));
    const bool yk = forces & DEMON_FORCE_SHIELDED_COULOMB;
    ev_reset(c);
    if (yk && (rc = enqueue_pairs(c, c->sx, 0, 0))) return rc;
    ForceOutLaunch f{};
    f.st = c->sx; f.q = c->q; f.m = c->m; f.orig = c->d_orig; f.pairs = pair_sums(c);
    f.fx = c->fout; f.fy = c->fout + c->npad; f.n = c->n; f.npad = c->npad;
    f.forces = forces; f.fp = force_params(c->p); f.t = t;
    f.step = step; f.stage = stage;
    HIP_TRY(c, launch_force_out(f, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->h_tmp.data(), c->fout, sizeof(double) * 2 * c->npad,
                              hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const double *gx = c->h_tmp.data(), *gy = gx + c->npad;
    for (int64_t k = 0; k < c->n; ++k) {
        const int64_t o = c->perm[(size_t)k];
        fx[o] += gx[k];
        fy[o] += gy[k];
    }
    return DEMON_OK;
}

int demon_force_stage(demon_ctx *c, int32_t forces, int32_t stage, uint64_t step,
                      const double *xs, const double *ys, const double *vxs,
                      const double *vys, double *fx, double *fy) {
    if (forces & (DEMON_FORCE_TIME_VARYING_DRAG | DEMON_FORCE_TIME_VARYING_THERMAL | DEMON_FORCE_DRIVING))
        return fail(c, DEMON_ERR_ARG,
                    "TimeVaryingDrag, TimeVaryingThermal and Driving need the stage time: use demon_force_stage_at");
    return demon_force_stage_at(c, forces, stage, step, 0.0, xs, ys, vxs, vys, fx, fy);
}

int demon_rk4_steps(demon_ctx *c, int64_t steps, double init_dt, double *current_time) {
    return rk_steps(c, 4, steps, init_dt, current_time);
}

int demon_rk4_move(demon_ctx *c, double *current_time, double end_time, double init_dt,
                   int64_t *steps_done) {
    return rk_move(c, 4, current_time, end_time, init_dt, steps_done);
}

int demon_rk2_steps(demon_ctx *c, int64_t steps, double init_dt, double *current_time) {
    return rk_steps(c, 2, steps, init_dt, current_time);
}

int demon_rk2_move(demon_ctx *c, double *current_time, double end_time, double init_dt,
                   int64_t *steps_done) {
    return rk_move(c, 2, current_time, end_time, init_dt, steps_done);
}

int demon_use_graph(demon_ctx *c, int32_t enable) {
    if (!c) return fail(nullptr, DEMON_ERR_ARG, "ctx is NULL");
    c->use_graph = enable != 0;
    if (!c->use_graph) drop_graph(c);
    return DEMON_OK;
}

int demon_get_stats(const demon_ctx *c, demon_stats *out) {
    if (!c || !out) return fail(nullptr, DEMON_ERR_ARG, "NULL argument");
    Scalars s{};
    hipError_t e = hipMemcpy(&s, c->d_sc, sizeof(Scalars), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(const_cast<demon_ctx *>(c), e, "hipMemcpy stats");
    out->steps = (int64_t)s.steps_done;
    out->dt_reductions = (int64_t)s.dt_reductions;
    out->force_evals = c->force_evals;
    out->current_time = s.t;
    out->last_dt = s.dt_step;
    return DEMON_OK;
}

int demon_synchronize(demon_ctx *c) {
    if (!c) return fail(nullptr, DEMON_ERR_ARG, "ctx is NULL");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return DEMON_OK;
}

int demon_set_pair_shape(demon_ctx *c, int32_t i_per_lane, int32_t j_slices) {
    if (!c) return fail(nullptr, DEMON_ERR_ARG, "ctx is NULL");
    if (i_per_lane != 1 && i_per_lane != 2 && i_per_lane != 4)
        return fail(c, DEMON_ERR_ARG, "i_per_lane must be 1, 2 or 4");
    if (j_slices < 0 || j_slices > kMaxSlices) return fail(c, DEMON_ERR_ARG, "j_slices out of range");
    c->i_per_lane = i_per_lane;
    c->slices_req = j_slices;
    choose_slices(c);
    drop_graph(c);
    return DEMON_OK;
}

int demon_get_pair_shape(const demon_ctx *c, int32_t *i_per_lane, int32_t *j_slices) {
    if (!c) return fail(nullptr, DEMON_ERR_ARG, "ctx is NULL");
    const bool n3 = c->pair_mode == DEMON_PAIRS_NEWTON3;
    if (i_per_lane) *i_per_lane = n3 ? 1 : c->i_per_lane;
    if (j_slices) *j_slices = n3 ? c->n3_slices : c->slices;
    return DEMON_OK;
}

int demon_set_pair_mode(demon_ctx *c, int32_t mode) {
    if (!c) return fail(nullptr, DEMON_ERR_ARG, "ctx is NULL");
    if (mode != DEMON_PAIRS_GATHER && mode != DEMON_PAIRS_NEWTON3)
        return fail(c, DEMON_ERR_ARG, "mode must be DEMON_PAIRS_GATHER or DEMON_PAIRS_NEWTON3");
    if (mode == DEMON_PAIRS_NEWTON3) {
        HIP_TRY(c, hipSetDevice(c->p.device));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        const int rc = ensure_n3(c);
        if (rc) return rc;
    }
    c->pair_mode = mode;
    drop_graph(c);
    return DEMON_OK;
}

int demon_get_pair_mode(const demon_ctx *c, int32_t *mode) {
    if (!c || !mode) return fail(nullptr, DEMON_ERR_ARG, "NULL argument");
    *mode = c->pair_mode;
    return DEMON_OK;
}

int demon_set_timing(demon_ctx *c, int32_t enable) {
    if (!c) return fail(nullptr, DEMON_ERR_ARG, "ctx is NULL");
    c->timing = enable != 0;
    return DEMON_OK;
}

int demon_last_timing(const demon_ctx *c, double *pair_ms, double *total_ms) {
    if (!c) return fail(nullptr, DEMON_ERR_ARG, "ctx is NULL");
    if (pair_ms) *pair_ms = c->last_pair_ms;
    if (total_ms) *total_ms = c->last_total_ms;
    return DEMON_OK;
}

int demon_last_kernel_ms(const demon_ctx *c, double *pair_ms, int64_t *pair_launches, double *list_ms,
                         double *update_ms, int64_t *update_launches) {
    if (!c) return fail(nullptr, DEMON_ERR_ARG, "ctx is NULL");
    if (pair_ms) *pair_ms = c->last_pair_ms;
    if (pair_launches) *pair_launches = c->last_pair_launches;
    if (list_ms) *list_ms = c->last_list_ms;
    if (update_ms) *update_ms = c->last_update_ms;
    if (update_launches) *update_launches = c->last_update_launches;
    return DEMON_OK;
}

int demon_pair_tiles(const demon_ctx *c, int64_t *tiles, int64_t *items) {
    if (!c) return fail(nullptr, DEMON_ERR_ARG, "ctx is NULL");
    Scalars s{};
    hipError_t e = hipMemcpy(&s, c->d_sc, sizeof(Scalars), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(const_cast<demon_ctx *>(c), e, "hipMemcpy tile count");
    if (tiles) *tiles = (int64_t)s.n3_tiles;
    if (items) *items = (int64_t)s.n3_items;
    return DEMON_OK;
}

int demon_pair_counts(const demon_ctx *c, int64_t *near_pairs, int64_t *tested_pairs) {
    if (!c) return fail(nullptr, DEMON_ERR_ARG, "ctx is NULL");
    Scalars s{};
    hipError_t e = hipMemcpy(&s, c->d_sc, sizeof(Scalars), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(const_cast<demon_ctx *>(c), e, "hipMemcpy pair counts");
    unsigned long long near = 0, tested = 0;
    for (int k = 0; k < kCountSlots; ++k) {
        near += s.pair_counts[k][0];
        tested += s.pair_counts[k][1];
    }
    if (near_pairs) *near_pairs = (int64_t)near;
    if (tested_pairs) *tested_pairs = (int64_t)tested;
    return DEMON_OK;
}

int demon_fit_loss_grad(demon_ctx *c, int64_t steps, double dt, int64_t obs_every,
                        int64_t ckpt_every, const double *target_x, const double *target_y,
                        double *loss, double *grad_q, double *grad_kappa, double *grad_confine,
                        double *grad_gamma, double *grad_x0, double *grad_y0, double *grad_vx0,
                        double *grad_vy0) {
    if (!c) return fail(nullptr, DEMON_ERR_ARG, "ctx is NULL");
    if (!c->uploaded) return fail(c, DEMON_ERR_ARG, "no state uploaded");
    if (steps < 1 || obs_every < 1 || ckpt_every < 1 || !(dt > 0.0) || !loss)
        return fail(c, DEMON_ERR_ARG, "bad steps/obs_every/ckpt_every/dt/loss");
    const int32_t allowed = DEMON_FORCE_CONFINEMENT | DEMON_FORCE_DRAG | DEMON_FORCE_SHIELDED_COULOMB;
    if ((c->p.forces & ~allowed) || !(c->p.forces & DEMON_FORCE_SHIELDED_COULOMB))
        return fail(c, DEMON_ERR_UNSUPPORTED, "the adjoint covers Confinement, Drag and ShieldedCoulomb only");
    set_order(c, 4);  // the fit differentiates the RK4 map
    if (c->p.precision != DEMON_PRECISION_FP64)
        return fail(c, DEMON_ERR_UNSUPPORTED, "the adjoint is FP64 only");
    const int64_t n_obs = steps / obs_every;
    if (n_obs > 0 && (!target_x || !target_y)) return fail(c, DEMON_ERR_ARG, "targets required");
    const bool want = grad_q || grad_kappa || grad_confine || grad_gamma || grad_x0 || grad_y0 ||
                      grad_vx0 || grad_vy0;
    HIP_TRY(c, hipSetDevice(c->p.device));
    const size_t P = (size_t)c->npad;
    const int64_t n_ckpt = (steps + ckpt_every - 1) / ckpt_every;
    // arena: ckpt[n_ckpt][4] | targets[n_obs][2] | seeds[n_obs][2] | lossp | store[K][4] (SD each)
    //        | bar[4] | ibar[16] | u[2] | g[4] | apart[slices][4]
    const size_t off_ck = 0, off_tg = off_ck + (size_t)n_ckpt * 4 * P;
    const size_t off_sd = off_tg + (size_t)n_obs * 2 * P, off_lp = off_sd + (size_t)n_obs * 2 * P;
    const size_t SD = store_doubles(c->npad);  // one stored stage record
    // The forward sweep keeps the stage records of the last kseg segments, as
    // many as fit in kFitKeepBytes and a quarter of the free device memory (at
    // least one); the backward sweep recomputes only the segments before them,
    // into the first segment's slots, which it has consumed by then.
    int64_t kseg = 1;
    if (want) {
        size_t free_b = 0, total_b = 0;
        HIP_TRY(c, hipMemGetInfo(&free_b, &total_b));
        const size_t seg_bytes = (size_t)ckpt_every * 4 * SD * sizeof(double);
        size_t cap = kFitKeepBytes;
        if (const char *env = std::getenv("DEMON_FIT_KEEP_BYTES")) cap = (size_t)std::strtoull(env, nullptr, 10);
        const size_t fit = std::min(cap, free_b / 4) / seg_bytes;
        kseg = (int64_t)std::max<size_t>(1, std::min<size_t>(fit, (size_t)n_ckpt));
    }
    const int64_t kept0 = (n_ckpt - kseg) * ckpt_every;  // first step whose records are kept
    const int64_t store_steps = std::max(std::min(ckpt_every, steps), steps - kept0);
    const size_t off_st = off_lp + P, off_bar = off_st + (size_t)store_steps * 4 * SD;
    const size_t off_ib = off_bar + 4 * P, off_u = off_ib + 16 * P, off_g = off_u + 2 * P;
    const size_t off_ap = off_g + 4 * P, off_dt = off_ap + (size_t)c->adj_slices * 4 * P;
    const size_t total = off_dt + (size_t)steps;
    int rc = fit_arena(c, total);
    if (rc) return rc;
    double *A = c->d_fit;
    for (int k = 0; k < 3; ++k)
        if (!c->ev_fit[k]) HIP_TRY(c, hipEventCreate(&c->ev_fit[k]));
    // targets in device order
    for (int64_t o = 0; o < n_obs; ++o) {
        if ((rc = put(c, A + off_tg + (size_t)(2 * o) * P, target_x + o * c->n, 0.0, 0)) ||
            (rc = put(c, A + off_tg + (size_t)(2 * o + 1) * P, target_y + o * c->n, 0.0, 1)))
            return rc;
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    HIP_TRY(c, hipMemsetAsync(A + off_lp, 0, sizeof(double) * P, c->stream));
    // store slots need valid padding (nc = -inf) for the adjoint prefilter
    if (want) HIP_TRY(c, launch_prepare_records(A + off_st, store_steps * 4, c->npad, c->stream));
    const double norm = 1.0 / (2.0 * (double)c->n * (double)(n_obs > 0 ? n_obs : 1));
    const double inf = std::numeric_limits<double>::infinity();

    HIP_TRY(c, hipMemsetAsync(&c->d_sc->adj_counts, 0, sizeof(c->d_sc->adj_counts), c->stream));
    // forward sweep with checkpoints and observations
    HIP_TRY(c, hipEventRecord(c->ev_fit[0], c->stream));
    hipLaunchKernelGGL(k_call_setup, dim3(1), dim3(64), 0, c->stream, c->d_sc, 0.0, inf, dt,
                       (unsigned long long)steps);
    HIP_TRY(c, hipGetLastError());
    ev_reset(c);
    const bool timing = c->timing;
    c->timing = false;
    for (int64_t s = 0; s < steps; ++s) {
        if (s % ckpt_every == 0 && (rc = save_state(c, A + off_ck + (size_t)(s / ckpt_every) * 4 * P))) break;
        double *store = (want && s >= kept0) ? A + off_st + (size_t)(s - kept0) * 4 * SD : nullptr;
        if ((rc = enqueue_step_ex(c, c->p.dt_check, store))) break;
        // the step's dt (after the dt check), replayed by the adjoint sweep
        HIP_TRY(c, hipMemcpyAsync(A + off_dt + s, &c->d_sc->dt_step, sizeof(double),
                                  hipMemcpyDeviceToDevice, c->stream));
        if ((s + 1) % obs_every == 0 && (s + 1) / obs_every <= n_obs) {
            // the step's last update wrote the new state; close it before observing
            const int64_t o = (s + 1) / obs_every - 1;
            HIP_TRY(c, launch_observe(c->x, c->y, A + off_tg + (size_t)(2 * o) * P,
                                      want ? A + off_sd + (size_t)(2 * o) * P : nullptr, A + off_lp,
                                      norm, c->n, c->npad, c->stream));
        }
    }
    if (!rc) rc = launch_boundary(c->d_sc, 1, c->stream) == hipSuccess ? 0 : fail(c, DEMON_ERR_HIP, "boundary");
    HIP_TRY(c, hipEventRecord(c->ev_fit[1], c->stream));
    std::vector<double> tmp;
    if (!rc) *loss = host_sum(c, A + off_lp, tmp, &rc) * norm;
    std::vector<double> dts((size_t)steps);
    if (!rc) {
        hipError_t e = hipMemcpyAsync(dts.data(), A + off_dt, sizeof(double) * steps, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = hip_fail(c, e, "hipMemcpy dt");
        Scalars sc{};
        if (!rc && hipMemcpy(&sc, c->d_sc, sizeof(sc), hipMemcpyDeviceToHost) == hipSuccess && sc.err == 2)
            rc = fail(c, DEMON_ERR_STATE, "dt check: two tested particles coincide");
    }
    if (!rc && want) {
        double *bar = A + off_bar, *ibar = A + off_ib, *u = A + off_u, *g = A + off_g, *ap = A + off_ap;
        HIP_TRY(c, hipMemsetAsync(bar, 0, sizeof(double) * 4 * P, c->stream));
        HIP_TRY(c, hipMemsetAsync(g, 0, sizeof(double) * 4 * P, c->stream));
        const double rc2p = (10.0 / c->p.shielding) * (10.0 / c->p.shielding) * (1.0 + kCutoffSlack);
        for (int64_t seg = n_ckpt - 1; seg >= 0 && !rc; --seg) {
            const int64_t s0 = seg * ckpt_every, s1 = std::min(steps, s0 + ckpt_every);
            const bool kept = s0 >= kept0;  // records stored by the forward sweep
            const int64_t base = kept ? s0 - kept0 : 0;
            if (!kept) {
                if ((rc = restore_state(c, A + off_ck + (size_t)seg * 4 * P))) break;
                hipLaunchKernelGGL(k_call_setup, dim3(1), dim3(64), 0, c->stream, c->d_sc, 0.0, inf, dt,
                                   (unsigned long long)(s1 - s0));
                for (int64_t s = s0; s < s1 && !rc; ++s)
                    rc = enqueue_step_ex(c, c->p.dt_check, A + off_st + (size_t)(s - s0) * 4 * SD);
                if (rc) break;
                HIP_TRY(c, launch_boundary(c->d_sc, 1, c->stream));
            }
            for (int64_t s = s1 - 1; s >= s0 && !rc; --s) {
                const bool obs = (s + 1) % obs_every == 0 && (s + 1) / obs_every <= n_obs;
                const double *seed = obs ? A + off_sd + (size_t)(2 * ((s + 1) / obs_every - 1)) * P : nullptr;
                HIP_TRY(c, launch_adj_begin(bar, ibar, u, g, seed, c->m, c->n, c->npad, c->stream));
                for (int k = 4; k >= 1; --k) {
                    const StageBuf st = store_view(A + off_st + ((size_t)(s - s0 + base) * 4 + (k - 1)) * SD, c->npad);
                    // timed with events when the context times its launches
                    bool on = false;
                    if ((rc = ev_open_if(c, timing, &on))) break;
                    HIP_TRY(c, launch_adj_pairs(st, c->q, u, u + P, ap, c->n, c->npad, c->adj_jslice, c->adj_slices,
                                                c->p.shielding, rc2p, c->d_sc, c->stream));
                    if ((rc = ev_close(c, on, kTimedAdjPairs))) break;
                    HIP_TRY(c, launch_adj_stage(bar, ibar, u, g, k, st, c->q, c->m, ap, c->adj_slices,
                                                c->p.forces, dts[(size_t)s], c->p.confine,
                                                c->p.drag_gamma, c->n, c->npad, c->stream));
                }
            }
        }
        if (!rc) {
            double gk = host_sum(c, g + P, tmp, &rc);
            double gc = rc ? 0 : host_sum(c, g + 2 * P, tmp, &rc);
            double gg = rc ? 0 : host_sum(c, g + 3 * P, tmp, &rc);
            if (grad_kappa) *grad_kappa = gk;
            if (grad_confine) *grad_confine = gc;
            if (grad_gamma) *grad_gamma = gg;
            double *outs[5] = {grad_q, grad_x0, grad_y0, grad_vx0, grad_vy0};
            const double *srcs[5] = {g, bar, bar + P, bar + 2 * P, bar + 3 * P};
            for (int a = 0; a < 5 && !rc; ++a) {
                if (!outs[a]) continue;
                tmp.resize(P);
                hipError_t e = hipMemcpyAsync(tmp.data(), srcs[a], sizeof(double) * P, hipMemcpyDeviceToHost, c->stream);
                if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
                if (e != hipSuccess) rc = hip_fail(c, e, "hipMemcpy gradient");
                for (int64_t k = 0; k < c->n && !rc; ++k) outs[a][c->perm[(size_t)k]] = tmp[(size_t)k];
            }
        }
    }
    // the context's state is the initial state again
    if (!rc) rc = restore_state(c, A + off_ck);
    HIP_TRY(c, hipEventRecord(c->ev_fit[2], c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->timing = timing;
    float f0 = 0, f1 = 0;
    HIP_TRY(c, hipEventElapsedTime(&f0, c->ev_fit[0], c->ev_fit[1]));
    HIP_TRY(c, hipEventElapsedTime(&f1, c->ev_fit[1], c->ev_fit[2]));
    c->last_fwd_ms = f0;
    c->last_bwd_ms = f1;
    c->last_adj_ms = 0.0;
    c->last_adj_launches = 0;
    for (size_t k = 0; k + 1 < c->ev_used; k += 2) {
        if (c->ev_kind[k / 2] != kTimedAdjPairs) continue;
        float pm = 0.0f;
        HIP_TRY(c, hipEventElapsedTime(&pm, c->ev[k], c->ev[k + 1]));
        c->last_adj_ms += pm;
        c->last_adj_launches += 1;
    }
    ev_reset(c);
    return rc;
}

int demon_fit_adjoint_kernel(const demon_ctx *c, double *pair_ms, int64_t *launches, int64_t *near_pairs,
                             int64_t *tested_pairs) {
    if (!c) return fail(nullptr, DEMON_ERR_ARG, "ctx is NULL");
    Scalars s{};
    hipError_t e = hipMemcpy(&s, c->d_sc, sizeof(Scalars), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(const_cast<demon_ctx *>(c), e, "hipMemcpy adjoint pair counts");
    unsigned long long near = 0, tested = 0;
    for (int k = 0; k < kCountSlots; ++k) {
        near += s.adj_counts[k][0];
        tested += s.adj_counts[k][1];
    }
    if (pair_ms) *pair_ms = c->last_adj_ms;
    if (launches) *launches = c->last_adj_launches;
    if (near_pairs) *near_pairs = (int64_t)near;
    if (tested_pairs) *tested_pairs = (int64_t)tested;
    return DEMON_OK;
}

int demon_fit_timing(const demon_ctx *c, double *forward_ms, double *backward_ms) {
    if (!c) return fail(nullptr, DEMON_ERR_ARG, "ctx is NULL");
    if (forward_ms) *forward_ms = c->last_fwd_ms;
    if (backward_ms) *backward_ms = c->last_bwd_ms;
    return DEMON_OK;
}

int demon_hex_lattice(int64_t n, double spacing, double jitter, uint64_t seed, double *x,
                      double *y) {
    if (n < 1 || !x || !y || !(spacing > 0.0)) return fail(nullptr, DEMON_ERR_ARG, "bad lattice arguments");
    // radius (in spacings) of a disc holding n sites of area sqrt(3)/2 each, plus margin
    const double rho = std::sqrt((double)n * std::sqrt(3.0) / (2.0 * M_PI)) + 3.0;
    const int64_t L = (int64_t)std::ceil(2.0 * rho) + 2;
    struct Site { int64_t r2, j, i; };
    std::vector<Site> sites;
    sites.reserve((size_t)((2 * L + 1) * (2 * L + 1)));
    for (int64_t j = -L; j <= L; ++j)
        for (int64_t i = -L; i <= L; ++i) sites.push_back({i * i + i * j + j * j, j, i});
    if ((int64_t)sites.size() < n) return fail(nullptr, DEMON_ERR_ARG, "lattice too small");
    std::partial_sort(sites.begin(), sites.begin() + n, sites.end(), [](const Site &a, const Site &b) {
        if (a.r2 != b.r2) return a.r2 < b.r2;
        if (a.j != b.j) return a.j < b.j;
        return a.i < b.i;
    });
    const double h = 0.8660254037844386;  // sqrt(3)/2
    double cx = 0.0, cy = 0.0;
    for (int64_t k = 0; k < n; ++k) {
        x[k] = spacing * ((double)sites[(size_t)k].i + 0.5 * (double)sites[(size_t)k].j);
        y[k] = spacing * (h * (double)sites[(size_t)k].j);
        cx += x[k];
        cy += y[k];
    }
    cx /= (double)n;
    cy /= (double)n;
    for (int64_t k = 0; k < n; ++k) {
        const U32x4 o = draw(seed, (uint64_t)k, 0, kDomainLattice, 0);
        x[k] = (x[k] - cx) + jitter * (2.0 * u01(o.v[0], o.v[1]) - 1.0);
        y[k] = (y[k] - cy) + jitter * (2.0 * u01(o.v[2], o.v[3]) - 1.0);
    }
    return DEMON_OK;
}

// ---------------------------------------------------------------------------
// Analysis of the current state.

namespace {

int analysis_scratch(demon_ctx *c, size_t doubles) {
    if (c->an_doubles >= doubles) return DEMON_OK;
    if (c->d_an) {
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        HIP_TRY(c, hipFree(c->d_an));
        c->d_an = nullptr;
        c->an_doubles = 0;
    }
    HIP_TRY(c, hipMalloc(&c->d_an, sizeof(double) * doubles));
    c->an_doubles = doubles;
    return DEMON_OK;
}

}  // namespace

// The scratch of demon_kinetic_energy_noalloc: the per-block sums and the
// totals of the kinetic-energy reduction for clouds of npad <= 2^22, in a
// file-scope device array instead of the context's analysis scratch.  Every
// GPU the library runs on holds its own copy (512 KB) until the process
// ends; calls that overlap on one GPU share it.
constexpr int64_t kNoAllocMaxNpad = (int64_t)1 << 22;
constexpr size_t kNoAllocDoubles = (size_t)(kNoAllocMaxNpad / 256) * 4 + 8;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {

// Device address of `scratch` per device ordinal, looked up (with that device
// current) on the first call on the device.  In host code `scratch` names a
// host-side stand-in, never the device memory.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};

int scratch_address(demon_ctx *c, double **out) {
    const int dev = c->p.device;
    if (dev < 0 || dev >= kMaxDevices) return fail(c, DEMON_ERR_UNSUPPORTED, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(c, e, "hipGetSymbolAddress of the kinetic-energy scratch");
        g_scratch[dev] = static_cast<double *>(ptr);
    }
    *out = g_scratch[dev];
    return DEMON_OK;
}

}  // namespace

int demon_kinetic_energy_noalloc(demon_ctx *c, double *total, double *thermal) {
    if (!c) return fail(nullptr, DEMON_ERR_ARG, "ctx is NULL");
    if (!c->uploaded) return fail(c, DEMON_ERR_ARG, "no state uploaded");
    if (!total && !thermal) return fail(c, DEMON_ERR_ARG, "no output");
    if (c->npad > kNoAllocMaxNpad)
        return fail(c, DEMON_ERR_UNSUPPORTED,
                    "demon_kinetic_energy_noalloc: n beyond its fixed scratch (at most 4194304 particles); "
                    "use demon_kinetic_energy");
    HIP_TRY(c, hipSetDevice(c->p.device));
    double *base = nullptr;
    int rc = scratch_address(c, &base);
    if (rc) return rc;
    const size_t part = (size_t)(c->npad / 256) * 4;
    double *totals = base + part;
    // every block sum and total is written before it is read: nothing to clear
    HIP_TRY(c, launch_kinetic(c->vx, c->vy, c->m, c->n, c->npad, base, totals, c->stream));
    double h[8];
    HIP_TRY(c, hipMemcpyAsync(h, totals, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (total) *total = 0.5 * h[3];
    if (thermal) *thermal = 0.5 * h[4];
    return DEMON_OK;
}

int demon_kinetic_energy(demon_ctx *c, double *total, double *thermal) {
    if (!c) return fail(nullptr, DEMON_ERR_ARG, "ctx is NULL");
    if (!c->uploaded) return fail(c, DEMON_ERR_ARG, "no state uploaded");
    if (!total && !thermal) return fail(c, DEMON_ERR_ARG, "no output");
    HIP_TRY(c, hipSetDevice(c->p.device));
    const size_t part = (size_t)(c->npad / 256) * 4;
    int rc = analysis_scratch(c, part + 8);
    if (rc) return rc;
    double *totals = c->d_an + part;
    HIP_TRY(c, launch_kinetic(c->vx, c->vy, c->m, c->n, c->npad, c->d_an, totals, c->stream));
    double h[8];
    HIP_TRY(c, hipMemcpyAsync(h, totals, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (total) *total = 0.5 * h[3];
    if (thermal) *thermal = 0.5 * h[4];
    return DEMON_OK;
}

int demon_pair_histogram(demon_ctx *c, double r_max, int32_t nbins, uint64_t *counts) {
    if (!c) return fail(nullptr, DEMON_ERR_ARG, "ctx is NULL");
    if (!c->uploaded) return fail(c, DEMON_ERR_ARG, "no state uploaded");
    if (!counts) return fail(c, DEMON_ERR_ARG, "counts is NULL");
    if (!(r_max > 0.0) || !std::isfinite(r_max)) return fail(c, DEMON_ERR_ARG, "r_max must be positive and finite");
    if (nbins < 1 || nbins > kMaxHistBins) return fail(c, DEMON_ERR_ARG, "nbins must be in [1, 2048]");
    HIP_TRY(c, hipSetDevice(c->p.device));
    // edges k * (r_max / nbins) with the last one r_max, as numpy.linspace
    std::vector<double> e2((size_t)nbins + 1);
    const double w = r_max / (double)nbins;
    for (int32_t k = 0; k < nbins; ++k) {
        const double e = (double)k * w;
        e2[(size_t)k] = e * e;
    }
    e2[(size_t)nbins] = r_max * r_max;
    const size_t need = (size_t)nbins + 1 + (size_t)nbins;  // edges2, counts
    int rc = analysis_scratch(c, need);
    if (rc) return rc;
    double *d_e2 = c->d_an;
    unsigned long long *d_counts = reinterpret_cast<unsigned long long *>(c->d_an + nbins + 1);
    HIP_TRY(c, hipMemcpyAsync(d_e2, e2.data(), sizeof(double) * e2.size(), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemsetAsync(d_counts, 0, sizeof(unsigned long long) * (size_t)nbins, c->stream));
    PairHistLaunch a{};
    a.x = c->x;
    a.y = c->y;
    a.n = c->n;
    a.nb = (int32_t)((c->n + 255) / 256);
    a.nbins = nbins;
    a.edges2 = d_e2;
    a.r2max = e2[(size_t)nbins];
    a.inv_width = (float)((double)nbins / r_max);
    a.counts = d_counts;
    HIP_TRY(c, launch_pair_hist(a, c->stream));
    static_assert(sizeof(unsigned long long) == sizeof(uint64_t), "64-bit counts");
    HIP_TRY(c, hipMemcpyAsync(counts, d_counts, sizeof(uint64_t) * (size_t)nbins, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return DEMON_OK;
}

int demon_trajectory_spectrum(const double *series, int64_t rows, int64_t stride, int64_t nuse, double *amp,
                              int32_t device) {
    if (!series || !amp) return fail(nullptr, DEMON_ERR_ARG, "NULL array");
    if (rows < 1 || nuse < 1 || stride < nuse) return fail(nullptr, DEMON_ERR_ARG, "need rows >= 1, 1 <= nuse <= stride");
    // the transform is O(rows^2): 2^22 snapshots is about 10^13 operations
    if (rows > ((int64_t)1 << 22) || nuse > ((int64_t)1 << 31))
        return fail(nullptr, DEMON_ERR_ARG, "at most 2^22 snapshots and 2^31 particles");
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(nullptr, e, "hipSetDevice");
    std::vector<double> packed;
    const double *src = series;
    if (stride != nuse) {
        packed.resize((size_t)(rows * nuse));
        for (int64_t t = 0; t < rows; ++t)
            std::memcpy(&packed[(size_t)(t * nuse)], series + (size_t)(t * stride), sizeof(double) * (size_t)nuse);
        src = packed.data();
    }
    const size_t nser = (size_t)(rows * nuse), half = (size_t)(rows / 2);
    // a stream of its own, so that the call waits for nothing else on the device
    hipStream_t st = nullptr;
    e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(nullptr, e, "hipStreamCreateWithFlags");
    double *d = nullptr;
    e = hipMalloc(&d, sizeof(double) * (nser + 3 * (size_t)rows + half + 1));
    if (e != hipSuccess) {
        (void)hipStreamDestroy(st);
        return hip_fail(nullptr, e, "hipMalloc");
    }
    double *d_ser = d, *d_s = d + nser, *d_wre = d_s + rows, *d_wim = d_wre + rows, *d_amp = d_wim + rows;
    e = hipMemcpyAsync(d_ser, src, sizeof(double) * nser, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_spectrum(d_ser, rows, nuse, d_s, d_wre, d_wim, d_amp, st);
    if (e == hipSuccess && half) e = hipMemcpyAsync(amp, d_amp, sizeof(double) * half, hipMemcpyDeviceToHost, st);
    const hipError_t es = hipStreamSynchronize(st);
    if (e == hipSuccess) e = es;
    (void)hipFree(d);
    (void)hipStreamDestroy(st);
    if (e != hipSuccess) return hip_fail(nullptr, e, "trajectory spectrum");
    return DEMON_OK;
}

}  // extern "C"
