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
// The pair sums run in k_adj_pairs with the same candidate prefilter as the
// forward kernel (the cutoff set is identical; the derivative ignores the
// jump at u = 10, which has measure zero).
#include "internal.h"

namespace demon {

namespace {

constexpr int kChunk = 8;

template <int R>
__global__ __launch_bounds__(256) void k_adj_pairs(const double *__restrict__ xs,
                                                   const double *__restrict__ ys,
                                                   const double *__restrict__ nc,
                                                   const double *__restrict__ q,
                                                   const double *__restrict__ ux,
                                                   const double *__restrict__ uy,
                                                   double *__restrict__ part, int64_t npad,
                                                   int32_t jslice, double kappa, double rc2p) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t ibase = (int64_t)blockIdx.x * (256 * R) + wave * (64 * R) + lane;

    double xi[R], yi[R], xm[R], ym[R], nT[R], qi[R], uxi[R], uyi[R];
    double rx[R], ry[R], dq[R], dk[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int64_t i = ibase + 64 * r;
        xi[r] = xs[i];
        yi[r] = ys[i];
        xm[r] = -2.0 * xi[r];
        ym[r] = -2.0 * yi[r];
        nT[r] = -(rc2p + nc[i]);
        qi[r] = q[i];
        uxi[r] = ux[i];
        uyi[r] = uy[i];
        rx[r] = ry[r] = dq[r] = dk[r] = 0.0;
    }
    const int32_t j_begin = (int32_t)blockIdx.y * jslice;
    const int32_t j_end = j_begin + jslice < (int32_t)npad ? j_begin + jslice : (int32_t)npad;
    for (int32_t j0 = j_begin; j0 < j_end; j0 += kChunk) {
        double cx[kChunk], cy[kChunk], cn[kChunk];
#pragma unroll
        for (int u = 0; u < kChunk; ++u) {
            cx[u] = xs[j0 + u];
            cy[u] = ys[j0 + u];
            cn[u] = nc[j0 + u];
        }
        uint64_t hit = 0;
#pragma unroll
        for (int u = 0; u < kChunk; ++u)
#pragma unroll
            for (int r = 0; r < R; ++r)
                hit |= __ballot(fma(xm[r], cx[u], fma(ym[r], cy[u], nT[r])) < cn[u]);
        if (!hit) continue;
#pragma unroll 1
        for (int u = 0; u < kChunk; ++u) {
            const int32_t j = j0 + u;
            const double xj = xs[j], yj = ys[j], nj = nc[j];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const double w = fma(xm[r], xj, fma(ym[r], yj, nT[r]));
                if (!(w < nj) || (int64_t)j == ibase + 64 * r) continue;
                const double dx = xi[r] - xj, dy = yi[r] - yj;
                const double rr = sqrt(dx * dx + dy * dy);
                const double uu = rr * kappa;
                if (!(uu < 10.0)) continue;
                const double e = exp(-uu);
                const double qj = q[j];
                const double r3 = rr * rr * rr;
                const double g = kCoulomb * qj * (1.0 + uu) * e / r3;     // phi / q_i
                const double phi = qi[r] * g;
                const double cqq = kCoulomb * (qi[r] * qj) * e;
                const double dphi = -cqq * (uu * uu + 3.0 * uu + 3.0) / (r3 * rr * rr);
                const double dux = uxi[r] - ux[j], duy = uyi[r] - uy[j];
                const double dot = dx * dux + dy * duy;
                rx[r] = rx[r] + (phi * dux + dphi * dot * dx);
                ry[r] = ry[r] + (phi * duy + dphi * dot * dy);
                dq[r] = dq[r] + g * dot;
                dk[r] = dk[r] + (-cqq * kappa / rr) * (uxi[r] * dx + uyi[r] * dy);
            }
        }
    }
    double *p0 = part + (size_t)(4 * blockIdx.y) * npad;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int64_t i = ibase + 64 * r;
        p0[i] = rx[r];
        p0[npad + i] = ry[r];
        p0[2 * npad + i] = dq[r];
        p0[3 * npad + i] = dk[r];
    }
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

hipError_t launch_adj_pairs(const double *xs, const double *ys, const double *nc, const double *q,
                            const double *ux, const double *uy, double *part, int64_t npad,
                            int32_t jslice, int32_t slices, int32_t i_per_lane, double kappa,
                            double rc2p, hipStream_t s) {
    if (npad % kPad || jslice % kChunk || (int64_t)slices * jslice < npad) return hipErrorInvalidValue;
    switch (i_per_lane) {
        case 1:
            hipLaunchKernelGGL(k_adj_pairs<1>, dim3((unsigned)(npad / 256), slices), dim3(256), 0, s, xs, ys, nc,
                               q, ux, uy, part, npad, jslice, kappa, rc2p);
            break;
        case 2:
            hipLaunchKernelGGL(k_adj_pairs<2>, dim3((unsigned)(npad / 512), slices), dim3(256), 0, s, xs, ys, nc,
                               q, ux, uy, part, npad, jslice, kappa, rc2p);
            break;
        default:
            hipLaunchKernelGGL(k_adj_pairs<4>, dim3((unsigned)(npad / 1024), slices), dim3(256), 0, s, xs, ys,
                               nc, q, ux, uy, part, npad, jslice, kappa, rc2p);
            break;
    }
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
