// stage.hip -- the O(N) per-stage work of Runge_Kutta4::moveParticles, fused
// into one pass per stage.
//
// For stage s the update kernel, per particle:
//   * sums the stage force in the driver's construction order
//     (driver.cpp:372-402), each plugin present when its flag is set:
//       Confinement  F  = (q C) r_s            ConfinementForce.cpp:48-54
//       ConfinementVoid  the confinement term, then
//                    F += ((q (-d)) (e^{-d r}/r)) r_s  ConfinementForceVoid.cpp:13-55
//       Drag         F += (m (-gamma)) v_s     DragForce.cpp:49-55
//       Driving      Fx += (sin(k x - w t_s) e^0) A      DrivingForce.cpp:14-61
//       Magnetic     F += (q B) (vy_s, -vx_s)  MagneticForce.cpp:44-49
//       RectConfinement F += ((q cX) x, (q cY) y)      RectConfinementForce.cpp:50-59
//       Rotational   F += c (-y, x)/r for rmin < r < rmax  RotationalForce.cpp:50-66
//       Yukawa       F += sum of the all-pairs partial sums (fixed order)
//       Thermal      F += heat U (cos t, sin t) ThermalForce.cpp:118-122
//       ThermalLocalized the same with heat = r < R ? in : out
//                    (ThermalForceLocalized.cpp:124-133)
//       TimeVaryingDrag F += (m (-(scale t_s + gamma))) v_s
//       TimeVaryingThermal the thermal force with heat = scale t_s + offset
//                    (TimeVaryingThermalForce.cpp:14-42)
//       Electric     F += ((q E) e^{r/(-R)}) r_s   ElectricForce.cpp:48-62
//       Gravitational Fy += m g                GravitationalForce.cpp:46-50
//       VertElectric Fy += (q E) e^{y/d}       VertElectricForce.cpp:47-53
//     with the stage time t_s = t, t + dt/2, t + dt/2, t + dt
//     (Runge_Kutta4.cpp:31-88); each thermal plugin draws from its own
//     Philox counter domain
//   * forms the RK increments k = (dt F)/m, l = dt v_s (Runge_Kutta4.cpp:32-104),
//   * builds the next stage state (CacheOperator.cpp:28-73: x + l/2, x + l)
//     or, after stage 4, the combine x += (a1 + ((a2 + a3) 2 + a4)) / 6
//     (Runge_Kutta4.cpp:107-151),
//   * writes the next stage's pair records (x, y, v, nc).
// Stage 1 also turns the device-wide FP32 minimum distance into the step's dt
// (Integrator::tryToReduceTimeStep, Integrator.cpp:144-155).
// Runge_Kutta2::moveParticles (Runge_Kutta2.cpp:23-78) shares stage 1 (its
// operation2 is the same x + l/2); its second stage (kRk2Final) evaluates the
// force at t + dt/2 on that record and ends the step with x += l2, v += k2.
// The force arrays of the reference are never materialised between stages,
// so nothing needs zeroing (Runge_Kutta4.cpp:44-46).
#include "internal.h"
#include "philox.h"

namespace demon {

namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;
// DrivingForce::waveNum and angFreq (DrivingForce.cpp:14-15), the same doubles
constexpr double kPiD = 3.14159265358979323846;
constexpr double kWaveNum = 2.0 * kPiD / 0.002;
constexpr double kAngFreq = 2.0 * kPiD * 10.0;

// One thermal kick (ThermalForce.cpp:118-122): U in [0,1), theta in [0,2pi)
// from the Philox block of (particle, step, domain, stage).
__device__ __forceinline__ void thermal_kick(double heat, uint64_t seed, int32_t oi, uint64_t step,
                                             uint32_t domain, int32_t stage, double &fx, double &fy) {
    const U32x4 o = draw(seed, (uint64_t)oi, step, domain, (uint32_t)stage);
    const double uu = u01(o.v[0], o.v[1]);
    const double th = u01(o.v[2], o.v[3]) * kTwoPi;
    const double tv = uu * heat;
    fx = fx + tv * cos(th);
    fy = fy + tv * sin(th);
}

__device__ __forceinline__ double da(double a1, double a23, double a4) {
    // Runge_Kutta4::da with a2 + a3 already summed (same rounding: a2 + a3 first).
    return (a1 + (a23 * 2.0 + a4)) / 6.0;
}

// Stage force on particle i (caller index oi) in the reference's summation order.
__device__ __forceinline__ void stage_force(int32_t forces, double xs, double ys, double vxs,
                                            double vys, double qi, double mi,
                                            const PairSums &ps, const unsigned long long *fmask,
                                            int64_t npad, int64_t i, int32_t sb, const ForceParams &fp,
                                            double t_stage, uint64_t step, int32_t stage,
                                            int32_t oi, double &fx, double &fy) {
    fx = 0.0;
    fy = 0.0;
    if (forces & 1) {
        const double cv = qi * fp.confine;
        fx = fx + cv * xs;
        fy = fy + cv * ys;
    }
    if (forces & 2048) {
        const double cv = qi * fp.confine;
        fx = fx + cv * xs;
        fy = fy + cv * ys;
        const double dv = qi * (-fp.void_decay);
        const double r = sqrt(xs * xs + ys * ys);
        const double er = exp(r * (-fp.void_decay)) / r;
        fx = fx + (dv * er) * xs;
        fy = fy + (dv * er) * ys;
    }
    if (forces & 2) {
        const double dr = mi * (-fp.gamma);
        fx = fx + dr * vxs;
        fy = fy + dr * vys;
    }
    if (forces & 64) {
        const double dist = xs - fp.drive_shift;
        const double arg = xs * kWaveNum - t_stage * kAngFreq;
        const double earg = (dist - dist) / fp.drive_const;  // -driveConst = dampConst
        fx = fx + (sin(arg) * exp(earg)) * fp.drive_amplitude;
    }
    if (forces & 1024) {
        const double qb = qi * fp.bfield;
        fx = fx + qb * vys;
        fy = fy - qb * vxs;
    }
    if (forces & 8192) {
        fx = fx + (qi * fp.confine_x) * xs;
        fy = fy + (qi * fp.confine_y) * ys;
    }
    if (forces & 128) {
        const double r = sqrt(xs * xs + ys * ys);
        if (r > fp.rot_inner && r < fp.rot_outer) {
            fx = fx - (fp.rot_const * ys) / r;
            fy = fy + (fp.rot_const * xs) / r;
        }
    }
    if (forces & 4) {
        if (ps.first) {
            // Newton-3: the work items of superblock sb, in item order; eight
            // items' loads are issued before their adds (the same order of
            // additions, more loads in flight)
            const int32_t it1 = ps.first[sb + 1], li = (int32_t)(i - (int64_t)sb * kSuper);
            int32_t it = ps.first[sb];
            for (; it + 8 <= it1; it += 8) {
                double px[8], py[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    px[k] = ps.part[(size_t)(it + k) * (2 * kSuper) + li];
                    py[k] = ps.part[(size_t)(it + k) * (2 * kSuper) + kSuper + li];
                }
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    fx = fx + px[k];
                    fy = fy + py[k];
                }
            }
            for (; it < it1; ++it) {
                fx = fx + ps.part[(size_t)it * (2 * kSuper) + li];
                fy = fy + ps.part[(size_t)it * (2 * kSuper) + kSuper + li];
            }
        }
        {
            // gather kernel: its j slices in order, eight slices' loads issued
            // before their adds
            int32_t s = 0;
            for (; s + 8 <= ps.slices; s += 8) {
                double px[8], py[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    px[k] = ps.part[(size_t)(2 * (s + k)) * npad + i];
                    py[k] = ps.part[(size_t)(2 * (s + k) + 1) * npad + i];
                }
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    fx = fx + px[k];
                    fy = fy + py[k];
                }
            }
            for (; s < ps.slices; ++s) {
                fx = fx + ps.part[(size_t)(2 * s) * npad + i];
                fy = fy + ps.part[(size_t)(2 * s + 1) * npad + i];
            }
        }
        if (ps.react) {
            // Newton-3 reactions of the evaluated tiles, in increasing
            // superblock offset d (bit d - 1 of the block's flag mask)
            const int32_t words = (ps.nsb / 2 + 63) / 64;
            for (int32_t w = 0; w < words; ++w) {
                unsigned long long mask = fmask[w];
                // up to eight offsets' loads in flight, added in increasing d
                while (mask) {
                    int32_t dd[8];
                    int m = 0;
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        dd[k] = 0;
                        if (mask) {
                            dd[k] = 64 * w + __builtin_ctzll(mask);
                            mask &= mask - 1;
                            m = k + 1;
                        }
                    }
                    double rx[8], ry[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        const double *rr = ps.react + (size_t)dd[k] * 2 * npad;
                        rx[k] = k < m ? rr[i] : 0.0;
                        ry[k] = k < m ? rr[npad + i] : 0.0;
                    }
#pragma unroll
                    for (int k = 0; k < 8; ++k)
                        if (k < m) {
                            fx = fx + rx[k];
                            fy = fy + ry[k];
                        }
                }
            }
        }
    }
    if (forces & 16) thermal_kick(fp.heat, fp.seed, oi, step, kDomainThermal, stage, fx, fy);
    if (forces & 32) {
        const double r = sqrt(xs * xs + ys * ys);
        thermal_kick(r < fp.heat_radius ? fp.heat_in : fp.heat_out, fp.seed, oi, step,
                     kDomainThermalLocalized, stage, fx, fy);
    }
    if (forces & 256) {
        const double dr = mi * (-(fp.drag_scale * t_stage + fp.gamma));
        fx = fx + dr * vxs;
        fy = fy + dr * vys;
    }
    if (forces & 512)
        thermal_kick(fp.heat_scale * t_stage + fp.heat_offset, fp.seed, oi, step,
                     kDomainTimeVaryingThermal, stage, fx, fy);
    if (forces & 4096) {
        const double r = sqrt(xs * xs + ys * ys);
        const double ev = (qi * fp.electric_field) * exp(r / (fp.plasma_radius * (-1)));
        fx = fx + ev * xs;
        fy = fy + ev * ys;
    }
    if (forces & 8) fy = fy + mi * fp.gravity;
    if (forces & 16384) fy = fy + (qi * fp.vert_electric_field) * exp(ys / fp.vert_decay);
}

// The Newton-3 tile flags of this block's superblock (one 256-thread block
// = one superblock) as a bitmask in LDS: bit d - 1 of fmask is set when the
// reactions react[d - 1] were written for it.  Every thread of the block
// must call it (it ends with a barrier).
__device__ __forceinline__ void load_flag_mask(const PairSums &ps, unsigned long long *fmask) {
    if (ps.react && threadIdx.x < 64) {
        const int32_t half = ps.nsb / 2, words = (half + 63) / 64;
        const uint32_t *fl = ps.rflag + (size_t)blockIdx.x * half;
        // every word's flags in one round trip, without branches (a flag
        // beyond the row is read from the row's last slot and not
        // counted), then the ballots
        uint32_t f[kMaxHalf / 64];
#pragma unroll
        for (int w = 0; w < kMaxHalf / 64; ++w) {
            const int32_t k = 64 * w + (int32_t)threadIdx.x;
            const uint32_t v = fl[k < half ? k : half - 1];
            f[w] = k < half ? v : 0u;
        }
#pragma unroll
        for (int w = 0; w < kMaxHalf / 64; ++w)
            if (w < words) {
                const unsigned long long m = __ballot(f[w] != 0u);
                if (threadIdx.x == 0) fmask[w] = m;
            }
    }
    __syncthreads();
}

template <int STAGE>
__global__ __launch_bounds__(256) void k_update(UpdateLaunch a) {
    __shared__ double red[4];
    __shared__ unsigned long long fmask[kMaxHalf / 64];
    Scalars *sc = a.sc;
    if (!sc->active) return;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    load_flag_mask(a.pairs, fmask);

    // the stage whose force this launch evaluates (RK2's second stage is a
    // stage-2 evaluation: record x + l1/2, time t + dt/2, thermal key 2)
    constexpr int FS = STAGE == kRk2Final ? 2 : STAGE;
    double dt;
    if (STAGE == 1) {
        // Integrator::modifyTimeStep: while any tested pair has
        // sqrtf(s) <= dist: dist /= 10, dt /= 10 (Integrator.cpp:144-155).  The
        // outcome depends only on the smallest s, so every thread replays the
        // loop on the device-wide minimum.
        dt = sc->init_dt;
        int err = 0;
        if (a.dtcheck) {
            const float d = sqrtf(__uint_as_float(sc->dtmin_bits));
            float dist = 1.0e-4f;
            int k = 0;
            while (d <= dist) {
                dist = dist / 10.0f;
                dt = dt / 10.0;
                if (++k > 60) {  // coincident particles: the reference loops forever
                    err = 2;
                    break;
                }
            }
        }
        if (i == 0) {
            sc->dt_step = dt;
            if (dt != sc->init_dt) sc->dt_reductions += 1;
            if (err) sc->err = err;
            sc->rmax2_bits = 0;  // refilled by this step's stage 4
        }
    } else {
        dt = sc->dt_step;
    }

    double r2new = 0.0, nx = 0.0, ny = 0.0;
    if (i < a.n) {
        const double xs = a.cur.xs[i], ys = a.cur.ys[i];
        const double vxs = a.cur.vxs[i], vys = a.cur.vys[i];
        const double qi = a.q[i], mi = a.m[i];
        if (a.store) {
            a.store[i] = xs;
            a.store[a.npad + i] = ys;
            a.store[2 * a.npad + i] = vxs;
            a.store[3 * a.npad + i] = vys;
            a.store[4 * a.npad + i] = a.cur.nc[i];
            if ((i & (kBoxBlock - 1)) == 0) {
                double *bx = a.store + 5 * a.npad + 4 * (i / kBoxBlock);
                const double *cb = a.cur.box + 4 * (i / kBoxBlock);
                bx[0] = cb[0];
                bx[1] = cb[1];
                bx[2] = cb[2];
                bx[3] = cb[3];
            }
        }
        double fx, fy;
        // Runge_Kutta4.cpp:31,50,69,88 (Runge_Kutta2.cpp:31,50): the stage times
        const double ts = FS == 1 ? sc->t : FS == 4 ? sc->t + dt : sc->t + dt / 2.0;
        stage_force(a.forces, xs, ys, vxs, vys, qi, mi, a.pairs, fmask, a.npad, i, (int32_t)blockIdx.x, a.fp, ts,
                    sc->step, FS, a.orig[i], fx, fy);
        const double k = (dt * fx) / mi;
        const double l = dt * vxs;
        const double mm = (dt * fy) / mi;
        const double nn = dt * vys;
        const double x = a.x[i], y = a.y[i], vx = a.vx[i], vy = a.vy[i];
        double *A1 = a.acc1, *A23 = a.acc23;
        const int64_t P = a.npad;
        double nvx, nvy;
        if (STAGE == 1 || STAGE == 2) {
            double *A = (STAGE == 1) ? A1 : A23;
            A[i] = k;
            A[P + i] = l;
            A[2 * P + i] = mm;
            A[3 * P + i] = nn;
            nx = 0.5 * l + x;
            ny = 0.5 * nn + y;
            nvx = 0.5 * k + vx;
            nvy = 0.5 * mm + vy;
        } else if (STAGE == kRk2Final) {
            // Runge_Kutta2.cpp:68-74
            nvx = vx + k;
            nx = x + l;
            nvy = vy + mm;
            ny = y + nn;
            a.x[i] = nx;
            a.y[i] = ny;
            a.vx[i] = nvx;
            a.vy[i] = nvy;
        } else if (STAGE == 3) {
            A23[i] = A23[i] + k;
            A23[P + i] = A23[P + i] + l;
            A23[2 * P + i] = A23[2 * P + i] + mm;
            A23[3 * P + i] = A23[3 * P + i] + nn;
            nx = x + l;
            ny = y + nn;
            nvx = vx + k;
            nvy = vy + mm;
        } else {
            nvx = vx + da(A1[i], A23[i], k);
            nx = x + da(A1[P + i], A23[P + i], l);
            nvy = vy + da(A1[2 * P + i], A23[2 * P + i], mm);
            ny = y + da(A1[3 * P + i], A23[3 * P + i], nn);
            a.x[i] = nx;
            a.y[i] = ny;
            a.vx[i] = nvx;
            a.vy[i] = nvy;
        }
        const double r2 = nx * nx + ny * ny;
        a.nxt.xs[i] = nx;
        a.nxt.ys[i] = ny;
        a.nxt.vxs[i] = nvx;
        a.nxt.vys[i] = nvy;
        a.nxt.nc[i] = -r2 * kNormShrink;
        r2new = r2;
    }
    if (a.nxt.box) wave_box(i < a.n, nx, ny, a.nxt.box + 4 * (i / kBoxBlock));
    if (STAGE == 4 || STAGE == kRk2Final) {
        // block max of |r|^2 for the next stage-1 candidate radius
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

// Step boundary: close the previous step (t += dt, Runge_Kutta4.cpp:135) and
// decide whether the next one runs (the while test, :24).
// The closing launch of a call only finishes the last step.
__global__ void k_boundary(Scalars *sc, int closing) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (sc->active) {
        sc->t = sc->t + sc->dt_step;
        sc->steps_done += 1;
        sc->step += 1;
    }
    const int act = !closing && (sc->t < sc->t_end) && (sc->remaining > 0) && (sc->err == 0);
    sc->active = act;
    if (act) sc->remaining -= 1;
    sc->dtmin_bits = 0x7f800000u;  // +inf
}

__global__ __launch_bounds__(256) void k_force_out(ForceOutLaunch a) {
    __shared__ unsigned long long fmask[kMaxHalf / 64];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    load_flag_mask(a.pairs, fmask);
    if (i >= a.n) return;
    double fx, fy;
    stage_force(a.forces, a.st.xs[i], a.st.ys[i], a.st.vxs[i], a.st.vys[i], a.q[i], a.m[i],
                a.pairs, fmask, a.npad, i, (int32_t)blockIdx.x, a.fp, a.t, a.step, a.stage, a.orig[i], fx, fy);
    a.fx[i] = fx;
    a.fy[i] = fy;
}

__device__ inline void prepare_at(const StageBuf &st, int64_t i, int64_t n) {
    double x = 0.0, y = 0.0;
    if (i < n) {
        x = st.xs[i];
        y = st.ys[i];
        st.nc[i] = -(x * x + y * y) * kNormShrink;
    } else {
        // NaN coordinates fail every distance test of the pair kernels
        st.xs[i] = __builtin_nan("");
        st.ys[i] = __builtin_nan("");
        st.vxs[i] = 0.0;
        st.vys[i] = 0.0;
        st.nc[i] = -__builtin_inf();
    }
    if (st.box) wave_box(i < n, x, y, st.box + 4 * (i / kBoxBlock));
}

__global__ __launch_bounds__(256) void k_prepare(StageBuf st, int64_t n, int64_t npad) {
    prepare_at(st, (int64_t)blockIdx.x * 256 + threadIdx.x, n);  // npad is a multiple of 256
}

// Empties `count` consecutive stored stage records (stride store_doubles(npad)):
// block b covers particles [256 (b mod bps), +256) of record b / bps.
__global__ __launch_bounds__(256) void k_prepare_records(double *base, int64_t npad) {
    const int64_t bps = npad / 256, b = blockIdx.x;
    const StageBuf st = store_view(base + (size_t)(b / bps) * store_doubles(npad), npad);
    prepare_at(st, (b % bps) * 256 + threadIdx.x, 0);
}

__global__ __launch_bounds__(256) void k_rmax(StageBuf st, int64_t n, Scalars *sc) {
    __shared__ double red[4];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double v = 0.0;
    if (i < n) v = st.xs[i] * st.xs[i] + st.ys[i] * st.ys[i];
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double b = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        if (b > 0.0) atomicMax(&sc->rmax2_bits, (unsigned long long)__double_as_longlong(b));
    }
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

hipError_t launch_update(int stage, const UpdateLaunch &a, hipStream_t s) {
    const dim3 grid(blocks_for(a.n)), block(256);
    switch (stage) {
        case 1: hipLaunchKernelGGL(k_update<1>, grid, block, 0, s, a); break;
        case 2: hipLaunchKernelGGL(k_update<2>, grid, block, 0, s, a); break;
        case 3: hipLaunchKernelGGL(k_update<3>, grid, block, 0, s, a); break;
        case 4: hipLaunchKernelGGL(k_update<4>, grid, block, 0, s, a); break;
        case kRk2Final: hipLaunchKernelGGL(k_update<kRk2Final>, grid, block, 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_boundary(Scalars *sc, int closing, hipStream_t s) {
    hipLaunchKernelGGL(k_boundary, dim3(1), dim3(64), 0, s, sc, closing);
    return hipGetLastError();
}

hipError_t launch_force_out(const ForceOutLaunch &a, hipStream_t s) {
    if (a.n < 1) return hipSuccess;
    hipLaunchKernelGGL(k_force_out, dim3(blocks_for(a.n)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_prepare_stage(StageBuf st, int64_t n, int64_t npad, hipStream_t s) {
    if (npad % 256) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_prepare, dim3(blocks_for(npad)), dim3(256), 0, s, st, n, npad);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_round_fp32(const double *__restrict__ src, float *__restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    dst[i] = (float)src[i];
}

hipError_t launch_prepare_records(double *base, int64_t count, int64_t npad, hipStream_t s) {
    if (npad % 256 || count < 0) return hipErrorInvalidValue;
    const int64_t blocks = count * (npad / 256);
    if (blocks > 0x7FFFFFFF) return hipErrorInvalidValue;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_prepare_records, dim3((unsigned)blocks), dim3(256), 0, s, base, npad);
    return hipGetLastError();
}

hipError_t launch_round_fp32(const double *src, float *dst, int64_t npad, hipStream_t s) {
    if (npad % 256) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_round_fp32, dim3(blocks_for(npad)), dim3(256), 0, s, src, dst);
    return hipGetLastError();
}

hipError_t launch_rmax(StageBuf st, int64_t n, Scalars *sc, hipStream_t s) {
    hipLaunchKernelGGL(k_rmax, dim3(blocks_for(n)), dim3(256), 0, s, st, n, sc);
    return hipGetLastError();
}

}  // namespace demon
