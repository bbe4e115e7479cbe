// adjoint.hip -- discrete adjoint of the RK4 step (Runge_Kutta4.cpp:23-151)
// with Confinement, Drag and ShieldedCoulomb forces, for gradients of a
// trajectory-matching loss with respect to the per-particle charges, the
// shielding constant kappa, the confinement constant C, the drag gamma and
// the initial state.  The reference has no differentiation; the CPU
// reference gradient is the oracle's restatement of the same adjoint.
//
// Notation (per particle i, stage state S = (x, y, vx, vy)):
//   F_i = q_i C r_i - gamma m_i v_i + sum_j phi_ij d_ij,
//   phi_ij = Cc q_i q_j (1 + u) e^-u / r^3,  u = kappa r,  d_ij = r_i - r_j.
// Given the adjoint u_i = kbar_i / m_i of the stage's velocity increment
// k = dt F / m, the stage adjoint is
//   rbar_i = dt [ C q_i u_i + sum_j M_ij (u_i - u_j) ],
//   M_ij = phi I + (phi'/r) d d^T,  phi'/r = -Cc q_i q_j e^-u (u^2+3u+3) / r^5,
//   vbar_i = dt [ lbar_i - gamma kbar_i ],
// and the parameter adjoints
//   qbar_i     += dt [ C u_i.r_i + sum_j (phi_ij / q_i) d_ij.(u_i - u_j) ],
//   kappabar   += dt sum_ij (-Cc q_i q_j kappa e^-u / r) u_i.d_ij,
//   Cbar       += dt sum_i q_i u_i.r_i,      gammabar += -dt sum_i m_i u_i.v_i.
// The pair sums run in k_adj_pairs with the forward kernel's box culling,
// prefilter and contribution test.
#include "pair_core.h"

namespace demon {

namespace {

using namespace pair;

constexpr int kChunkAdj = 4;  // j per chunk: x, y, q, ux, uy in 40 SGPRs

// Adjoint pair sums of one chunk for the wave's i (lane), gather form:
//   rx, ry += phi (u_i - u_j) + (phi'/r) ((u_i - u_j).d) d
//   dq     += (phi / q_i) (u_i - u_j).d
//   dk     += (-Cc q_i q_j kappa e^-u / r) u_i.d
// with phi = Cc q_i q_j (1 + u) e^-u / r^3, phi'/r = -Cc q_i q_j e^-u (u^2 + 3u + 3) / r^5.
template <bool CHECK, bool SELF>
__device__ __forceinline__ void chunk_adj(const ITile<1> &t, double uxi, double uyi, double qi,
                                          const double (&cx)[kChunkAdj], const double (&cy)[kChunkAdj],
                                          const double (&cq)[kChunkAdj], const double (&cux)[kChunkAdj],
                                          const double (&cuy)[kChunkAdj], int32_t j0, const PairEnv &e,
                                          uint64_t &n_near, double &rx, double &ry, double &dq, double &dk) {
#pragma unroll
    for (int u = 0; u < kChunkAdj; ++u) {
        const double dx = t.xi[0] - cx[u];
        const double dy = t.yi[0] - cy[u];
        const double r2 = fma(dx, dx, dy * dy);
        bool in = pair_in<CHECK>(dx, dy, r2, e);
        if (SELF) in = in && j0 + u != t.ibase;
        n_near += (uint64_t)__builtin_popcountll(__ballot(in));
        if (in) {
            const double y = rsqrt_fp64(r2);
            const double y2 = y * y;
            const double y3 = y2 * y;
            const double uu = (r2 * y) * e.cut.kappa;
            const double cqj = kCoulomb * cq[u];
            const double ex = exp_neg(uu, e.ec);
            const double g = (cqj * ex) * (1.0 + uu) * y3;  // phi / q_i
            const double ce = (qi * cqj) * ex;
            const double phi = qi * g;
            const double dphi = -ce * ((uu + 3.0) * uu + 3.0) * (y3 * y2);
            const double dux = uxi - cux[u], duy = uyi - cuy[u];
            const double dot = fma(dx, dux, dy * duy);
            const double w = dphi * dot;
            rx = rx + fma(phi, dux, w * dx);
            ry = ry + fma(phi, duy, w * dy);
            dq = fma(g, dot, dq);
            dk = fma(-ce * e.cut.kappa * y, fma(uxi, dx, uyi * dy), dk);
        }
    }
}

// Pair sums of the stage adjoint: for each i the sums above over all j of
// slice blockIdx.y, with the forward kernel's box culling and contribution
// test (the same cutoff set; the derivative ignores the jump at u = 10,
// which has measure zero).  Gather form, deterministic.  The wave's pair
// counts (ordered pairs inside the cutoff, and not culled by the boxes) go to
// sc->adj_counts with one atomic per counter per workgroup.
__global__ __launch_bounds__(256) void k_adj_pairs(const double *__restrict__ xs,
                                                   const double *__restrict__ ys,
                                                   const double *__restrict__ nc,
                                                   const double *__restrict__ box,
                                                   const double *__restrict__ q,
                                                   const double *__restrict__ ux,
                                                   const double *__restrict__ uy,
                                                   double *__restrict__ part, int32_t n, int32_t npad,
                                                   int32_t jslice, double kappa, double rc2p, ExpCoef ec,
                                                   Scalars *__restrict__ sc) {
    const Radii rad{rc2p, -1.0};
    const PairEnv e{make_cutoff(kappa, rc2p), rad, fp32_kl2(kappa), false, nullptr, ec};
    const int32_t ibase = (int32_t)blockIdx.x * 256 + (int32_t)threadIdx.x;
    const ITile<1> t = make_itile<1>(xs, ys, nc, q, n, ibase, rad.rf2);
    const double qi = q[ibase], uxi = ux[ibase], uyi = uy[ibase];
    const Box ib = box_union(empty_box(), box + 4 * (t.wlo / kBoxBlock));
    const bool ireal = t.whi <= n;
    double rx = 0.0, ry = 0.0, dq = 0.0, dk = 0.0;
    uint64_t n_near = 0, n_tested = 0;
    const int32_t j_begin = (int32_t)blockIdx.y * jslice;
    const int32_t j_end = j_begin + jslice < npad ? j_begin + jslice : npad;
    for (int32_t jb = j_begin; jb < j_end; jb += kBoxBlock) {
        const BlockClass cls = classify(ib, box + 4 * (jb / kBoxBlock), ireal, jb + kBoxBlock <= n, rad, e.cut);
        if (cls == kSkip) continue;
        n_tested += 64 * kBoxBlock;
        for (int32_t j0 = jb; j0 < jb + kBoxBlock; j0 += kChunkAdj) {
            double cx[kChunkAdj], cy[kChunkAdj], cq[kChunkAdj], cux[kChunkAdj], cuy[kChunkAdj];
#pragma unroll
            for (int u = 0; u < kChunkAdj; ++u) {
                cx[u] = xs[j0 + u];
                cy[u] = ys[j0 + u];
            }
            if (cls != kDense && !chunk_hit<1, kChunkAdj>(t, cx, cy, nc, j0)) continue;
#pragma unroll
            for (int u = 0; u < kChunkAdj; ++u) {
                cq[u] = q[j0 + u];
                cux[u] = ux[j0 + u];
                cuy[u] = uy[j0 + u];
            }
            if (j0 < t.whi && j0 + kChunkAdj > t.wlo)
                chunk_adj<true, true>(t, uxi, uyi, qi, cx, cy, cq, cux, cuy, j0, e, n_near, rx, ry, dq, dk);
            else if (cls == kDense)
                chunk_adj<false, false>(t, uxi, uyi, qi, cx, cy, cq, cux, cuy, j0, e, n_near, rx, ry, dq, dk);
            else
                chunk_adj<true, false>(t, uxi, uyi, qi, cx, cy, cq, cux, cuy, j0, e, n_near, rx, ry, dq, dk);
        }
    }
    // every thread reaches this point
    __shared__ unsigned long long cnt[4][2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        cnt[wave][0] = n_near;
        cnt[wave][1] = n_tested;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long *slot = sc->adj_counts[(blockIdx.x + blockIdx.y) & (kCountSlots - 1)];
        atomicAdd(&slot[0], ((cnt[0][0] + cnt[1][0]) + cnt[2][0]) + cnt[3][0]);
        atomicAdd(&slot[1], ((cnt[0][1] + cnt[1][1]) + cnt[2][1]) + cnt[3][1]);
    }
    double *p0 = part + (size_t)(4 * blockIdx.y) * npad;
    p0[ibase] = rx;
    p0[npad + ibase] = ry;
    p0[2 * npad + ibase] = dq;
    p0[3 * npad + ibase] = dk;
}

// Adjoint state of the backward sweep, per particle (device order):
//   bar[4][npad]        adjoint of the step-start state (x, y, vx, vy)
//   ibar[4][4][npad]    adjoint of the increments (l, n, k, mm) of stage k
//   u[2][npad]          kbar/m of the stage being processed
//   g[4][npad]          accumulated qbar, kappabar, Cbar, gammabar partials
struct AdjState {
    double *bar, *ibar, *u, *g;
};

// Start of one step's backward sweep: add the observation seed, split the
// step-output adjoint over the four increments (x' = x + (a1 + 2(a2+a3) + a4)/6)
// and form u for stage 4.
__global__ __launch_bounds__(256) void k_adj_begin(AdjState a, const double *seed,
                                                   const double *m, int64_t n, int64_t npad) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    if (seed) {
        a.bar[i] += seed[i];
        a.bar[npad + i] += seed[npad + i];
    }
    const double w[4] = {1.0 / 6.0, 2.0 / 6.0, 2.0 / 6.0, 1.0 / 6.0};
    for (int k = 0; k < 4; ++k)
        for (int c = 0; c < 4; ++c) a.ibar[(size_t)(4 * k + c) * npad + i] = w[k] * a.bar[(size_t)c * npad + i];
    const double mi = m[i];
    a.u[i] = a.ibar[(size_t)(4 * 3 + 2) * npad + i] / mi;
    a.u[npad + i] = a.ibar[(size_t)(4 * 3 + 3) * npad + i] / mi;
}

// Backward through stage k (1..4): stage adjoint, parameter partials, the
// stage-state dependence S_k = s + c_k inc_{k-1}, and u for stage k-1.
__global__ __launch_bounds__(256) void k_adj_stage(AdjState a, int stage, const double *xs,
                                                   const double *ys, const double *vxs,
                                                   const double *vys, const double *q,
                                                   const double *m, const double *part,
                                                   int32_t slices, int32_t forces, double dt,
                                                   double confine, double gamma, int64_t n,
                                                   int64_t npad) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int k = stage - 1;
    const double qi = q[i], mi = m[i];
    const double ux = a.u[i], uy = a.u[npad + i];
    const double *ib = a.ibar + (size_t)(4 * k) * npad;
    const double lb = ib[i], nb = ib[npad + i], kb = ib[2 * npad + i], mb = ib[3 * npad + i];
    double rx = 0.0, ry = 0.0, dq = 0.0, dk = 0.0;
    for (int32_t s = 0; s < slices; ++s) {
        const double *p0 = part + (size_t)(4 * s) * npad;
        rx += p0[i];
        ry += p0[npad + i];
        dq += p0[2 * npad + i];
        dk += p0[3 * npad + i];
    }
    double gc = 0.0, gg = 0.0;
    if (forces & 1) {
        rx += confine * qi * ux;
        ry += confine * qi * uy;
        const double ur = ux * xs[i] + uy * ys[i];
        dq += confine * ur;
        gc = qi * ur;
    }
    double vbx = lb, vby = nb;
    if (forces & 2) {
        vbx -= gamma * kb;
        vby -= gamma * mb;
        gg = -mi * (ux * vxs[i] + uy * vys[i]);
    }
    const double sb[4] = {dt * rx, dt * ry, dt * vbx, dt * vby};
    a.g[i] += dt * dq;
    a.g[npad + i] += dt * dk;
    a.g[2 * npad + i] += dt * gc;
    a.g[3 * npad + i] += dt * gg;
    for (int c = 0; c < 4; ++c) a.bar[(size_t)c * npad + i] += sb[c];
    if (k > 0) {
        const double cst = (k == 3) ? 1.0 : 0.5;
        double *pb = a.ibar + (size_t)(4 * (k - 1)) * npad;
        for (int c = 0; c < 4; ++c) pb[(size_t)c * npad + i] += cst * sb[c];
        a.u[i] = pb[2 * npad + i] / mi;
        a.u[npad + i] = pb[3 * npad + i] / mi;
    }
}

// Observation at the end of a forward step: seed = 2 (r - r*) norm and the
// per-particle squared error.
__global__ __launch_bounds__(256) void k_observe(const double *x, const double *y,
                                                 const double *tgt, double *seed, double *lossp,
                                                 double norm, int64_t n, int64_t npad) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double ex = x[i] - tgt[i], ey = y[i] - tgt[npad + i];
    lossp[i] += ex * ex + ey * ey;
    if (seed) {
        seed[i] = 2.0 * ex * norm;
        seed[npad + i] = 2.0 * ey * norm;
    }
}

unsigned blocks(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

hipError_t launch_adj_pairs(const StageBuf &st, const double *q, const double *ux, const double *uy,
                            double *part, int64_t n, int64_t npad, int32_t jslice, int32_t slices,
                            double kappa, double rc2p, Scalars *sc, hipStream_t s) {
    if (npad % kPad || jslice % kBoxBlock || (int64_t)slices * jslice < npad || !st.box || n > npad || !sc)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_adj_pairs, dim3((unsigned)(npad / 256), (unsigned)slices), dim3(256), 0, s, st.xs,
                       st.ys, st.nc, st.box, q, ux, uy, part, (int32_t)n, (int32_t)npad, jslice, kappa, rc2p,
                       exp_coef(), sc);
    return hipGetLastError();
}

hipError_t launch_adj_begin(double *bar, double *ibar, double *u, double *g, const double *seed,
                            const double *m, int64_t n, int64_t npad, hipStream_t s) {
    AdjState a{bar, ibar, u, g};
    hipLaunchKernelGGL(k_adj_begin, dim3(blocks(n)), dim3(256), 0, s, a, seed, m, n, npad);
    return hipGetLastError();
}

hipError_t launch_adj_stage(double *bar, double *ibar, double *u, double *g, int stage,
                            const StageBuf &st, const double *q, const double *m, const double *part,
                            int32_t slices, int32_t forces, double dt, double confine, double gamma,
                            int64_t n, int64_t npad, hipStream_t s) {
    AdjState a{bar, ibar, u, g};
    hipLaunchKernelGGL(k_adj_stage, dim3(blocks(n)), dim3(256), 0, s, a, stage, st.xs, st.ys, st.vxs,
                       st.vys, q, m, part, slices, forces, dt, confine, gamma, n, npad);
    return hipGetLastError();
}

hipError_t launch_observe(const double *x, const double *y, const double *tgt, double *seed,
                          double *lossp, double norm, int64_t n, int64_t npad, hipStream_t s) {
    hipLaunchKernelGGL(k_observe, dim3(blocks(n)), dim3(256), 0, s, x, y, tgt, seed, lossp, norm, n, npad);
    return hipGetLastError();
}

}  // namespace demon
