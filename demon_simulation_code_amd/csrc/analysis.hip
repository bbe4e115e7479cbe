// analysis.hip -- on-device analysis of a cloud: its kinetic energy (the
// kinetic temperature), the pair-distance histogram behind g(r), and the
// trajectory spectrum of the reference's FFTAnalysis -a
// (FFTAnalysis.cpp:213-283).
//
// Every reduction runs in a fixed order (lane-strided partial sums, a fixed
// butterfly, then blocks in index order), so results do not depend on
// scheduling; the histogram counts are integers and exact.
#include "internal.h"

namespace demon {

namespace {

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Sum of K doubles over a 256-thread block, in a fixed order; every thread
// gets the totals.  red: LDS of 4 * K doubles.
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double *red) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0)
        for (int k = 0; k < K; ++k) red[wave * K + k] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = ((red[k] + red[K + k]) + red[2 * K + k]) + red[3 * K + k];
    __syncthreads();
}

// Pass 1 (totals = nullptr): per block {sum m, sum m vx, sum m vy, sum m |v|^2}.
// Pass 2: per block {sum m |v - v_cm|^2} with v_cm from the pass-1 totals.
__global__ __launch_bounds__(256) void k_kinetic_part(const double *vx, const double *vy, const double *m,
                                                      int64_t n, const double *totals, double *part) {
    __shared__ double red[4 * 4];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (i < n) {
        const double mi = m[i], a = vx[i], b = vy[i];
        if (!totals) {
            v[0] = mi;
            v[1] = mi * a;
            v[2] = mi * b;
            v[3] = mi * (a * a + b * b);
        } else {
            const double da = a - totals[1] / totals[0], db = b - totals[2] / totals[0];
            v[0] = mi * (da * da + db * db);
        }
    }
    block_sum<4>(v, red);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) part[(size_t)blockIdx.x * 4 + k] = v[k];
}

// One block: totals[k] = sum over blocks b (in order of b's strided slot) of part[b][k].
__global__ __launch_bounds__(256) void k_kinetic_total(const double *part, int64_t blocks, double *totals) {
    __shared__ double red[4 * 4];
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t b = threadIdx.x; b < blocks; b += 256)
        for (int k = 0; k < 4; ++k) v[k] += part[(size_t)b * 4 + k];
    block_sum<4>(v, red);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) totals[k] = v[k];
}

// ---------------------------------------------------------------------------
// Pair-distance histogram.  Tiles of 256 x 256 particles; tile (I, J = I + d
// mod nb) for d = 0..nb/2 covers each unordered block pair once (for even nb
// the offset nb/2 only from I < nb/2).  A tile whose bounding boxes are at
// least r_max apart is skipped.  Counts go to per-wave LDS histograms, then
// to the global 64-bit counts.

__device__ __forceinline__ void wave_box(double x, double y, double &x0, double &x1, double &y0, double &y1) {
    // fmin / fmax skip the NaN of padding slots
    x0 = x1 = x;
    y0 = y1 = y;
    for (int o = 32; o >= 1; o >>= 1) {
        x0 = fmin(x0, __shfl_xor(x0, o));
        x1 = fmax(x1, __shfl_xor(x1, o));
        y0 = fmin(y0, __shfl_xor(y0, o));
        y1 = fmax(y1, __shfl_xor(y1, o));
    }
}

__global__ __launch_bounds__(256) void k_pair_hist(PairHistLaunch a) {
    extern __shared__ double smem[];
    const int32_t nbins = a.nbins;
    double *e2 = smem;                         // nbins + 1
    double *xj = e2 + (nbins + 2) / 2 * 2;     // 256 (8-byte aligned, even offset)
    double *yj = xj + 256;                     // 256
    double *boxes = yj + 256;                  // [2 tiles][4 waves][4]
    uint32_t *hist = reinterpret_cast<uint32_t *>(boxes + 32);  // [4][nbins]

    const int32_t nb = a.nb, half = nb / 2;
    const int32_t I = blockIdx.x, d = blockIdx.y;
    if (d == half && !(nb & 1) && I >= half && d > 0) return;
    const int32_t J = (I + d) % nb;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    const int64_t i = (int64_t)I * 256 + tid, j = (int64_t)J * 256 + tid;
    const double nan = __builtin_nan("");
    const double xi = i < a.n ? a.x[i] : nan, yi = i < a.n ? a.y[i] : nan;
    const double xjj = j < a.n ? a.x[j] : nan, yjj = j < a.n ? a.y[j] : nan;
    xj[tid] = xjj;
    yj[tid] = yjj;
    for (int32_t b = tid; b <= nbins; b += 256) e2[b] = a.edges2[b];
    for (int32_t b = tid; b < 4 * nbins; b += 256) hist[b] = 0u;
    double bx[4];
    wave_box(xi, yi, bx[0], bx[1], bx[2], bx[3]);
    if (lane == 0)
        for (int k = 0; k < 4; ++k) boxes[wave * 4 + k] = bx[k];
    wave_box(xjj, yjj, bx[0], bx[1], bx[2], bx[3]);
    if (lane == 0)
        for (int k = 0; k < 4; ++k) boxes[16 + wave * 4 + k] = bx[k];
    __syncthreads();
    double ib[4], jb[4];
    for (int t = 0; t < 2; ++t) {
        double *dst = t ? jb : ib;
        const double *src = boxes + 16 * t;
        dst[0] = fmin(fmin(src[0], src[4]), fmin(src[8], src[12]));
        dst[1] = fmax(fmax(src[1], src[5]), fmax(src[9], src[13]));
        dst[2] = fmin(fmin(src[2], src[6]), fmin(src[10], src[14]));
        dst[3] = fmax(fmax(src[3], src[7]), fmax(src[11], src[15]));
    }
    // gap between the boxes per axis (0 when they overlap); every block holds
    // a real particle (nb = ceil(n / 256)), so the boxes are finite
    const double gx = fmax(fmax(ib[0] - jb[1], jb[0] - ib[1]), 0.0);
    const double gy = fmax(fmax(ib[2] - jb[3], jb[2] - ib[3]), 0.0);
    if (!(gx * gx + gy * gy < a.r2max)) return;  // uniform over the block

    uint32_t *h = hist + wave * nbins;
    const float inv_w = a.inv_width;
    const int jstart = d == 0 ? tid + 1 : 0;  // the diagonal tile: j > i only
    // Consecutive j are spatial neighbours (Morton order), so a lane's bin
    // changes rarely: count runs in registers and add a run when it ends.
    // This takes the lanes of a wave off the same LDS word, where coarse bins
    // would otherwise serialise their atomics.
    int32_t run_bin = -1;
    uint32_t run = 0u;
    for (int jj = jstart; jj < 256; ++jj) {
        const double dx = xi - xj[jj];
        const double dy = yi - yj[jj];
        const double d2 = dx * dx + dy * dy;
        if (d2 < a.r2max) {
            // first guess, clamped in FP32 (the product may overflow for a tiny r_max)
            const float g = __fsqrt_rn((float)d2) * inv_w;
            int32_t b = g < (float)nbins ? (int32_t)g : nbins - 1;
            while (b > 0 && d2 < e2[b]) --b;
            while (b + 1 < nbins && d2 >= e2[b + 1]) ++b;
            if (b == run_bin) {
                ++run;
            } else {
                if (run) atomicAdd(&h[run_bin], run);
                run_bin = b;
                run = 1u;
            }
        }
    }
    if (run) atomicAdd(&h[run_bin], run);
    __syncthreads();
    for (int32_t b = tid; b < nbins; b += 256) {
        const uint32_t c = hist[b] + hist[nbins + b] + hist[2 * nbins + b] + hist[3 * nbins + b];
        if (c) atomicAdd(&a.counts[b], (unsigned long long)c);
    }
}

// ---------------------------------------------------------------------------
// Trajectory spectrum.

// s[t] = sum_{p < nuse} series[t * nuse + p]: one wave per row.
__global__ __launch_bounds__(256) void k_row_sums(const double *series, int64_t rows, int64_t nuse, double *s) {
    const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (t >= rows) return;
    double v = 0.0;
    for (int64_t p = lane; p < nuse; p += 64) v += series[(size_t)t * nuse + p];
    v = wave_sum(v);
    if (lane == 0) s[t] = v;
}

// w[m] = e^{-2 pi i m / rows} for m < rows.
__global__ __launch_bounds__(256) void k_twiddles(int64_t rows, double *wre, double *wim) {
    const int64_t m = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (m >= rows) return;
    double sn, cs;
    sincospi((2.0 * (double)m) / (double)rows, &sn, &cs);
    wre[m] = cs;
    wim[m] = -sn;
}

// amp[k] = |sum_t s[t] w[(k t) mod rows]| / nuse for k < rows / 2.
__global__ __launch_bounds__(256) void k_dft(const double *s, const double *wre, const double *wim, int64_t rows,
                                             int64_t nuse, double *amp) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= rows / 2) return;
    double re = 0.0, im = 0.0;
    int64_t m = 0;
    for (int64_t t = 0; t < rows; ++t) {
        re = re + s[t] * wre[m];
        im = im + s[t] * wim[m];
        m += k;
        if (m >= rows) m -= rows;
    }
    amp[k] = sqrt(re * re + im * im) / (double)nuse;
}

}  // namespace

hipError_t launch_kinetic(const double *vx, const double *vy, const double *m, int64_t n, int64_t npad,
                          double *scratch, double *totals, hipStream_t s) {
    const int64_t blocks = npad / 256;
    hipLaunchKernelGGL(k_kinetic_part, dim3((unsigned)blocks), dim3(256), 0, s, vx, vy, m, n,
                       (const double *)nullptr, scratch);
    hipLaunchKernelGGL(k_kinetic_total, dim3(1), dim3(256), 0, s, (const double *)scratch, blocks, totals);
    hipLaunchKernelGGL(k_kinetic_part, dim3((unsigned)blocks), dim3(256), 0, s, vx, vy, m, n,
                       (const double *)totals, scratch);
    hipLaunchKernelGGL(k_kinetic_total, dim3(1), dim3(256), 0, s, (const double *)scratch, blocks, totals + 4);
    return hipGetLastError();
}

size_t pair_hist_lds_bytes(int32_t nbins) {
    return sizeof(double) * ((size_t)(nbins + 2) / 2 * 2 + 512 + 32) + sizeof(uint32_t) * 4 * (size_t)nbins;
}

hipError_t launch_pair_hist(const PairHistLaunch &a, hipStream_t s) {
    const dim3 grid((unsigned)a.nb, (unsigned)(a.nb / 2 + 1));
    hipLaunchKernelGGL(k_pair_hist, grid, dim3(256), pair_hist_lds_bytes(a.nbins), s, a);
    return hipGetLastError();
}

hipError_t launch_spectrum(const double *series, int64_t rows, int64_t nuse, double *s_buf, double *wre,
                           double *wim, double *amp, hipStream_t s) {
    hipLaunchKernelGGL(k_row_sums, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, s, series, rows, nuse, s_buf);
    hipLaunchKernelGGL(k_twiddles, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, s, rows, wre, wim);
    if (rows / 2 > 0)
        hipLaunchKernelGGL(k_dft, dim3((unsigned)((rows / 2 + 255) / 256)), dim3(256), 0, s, (const double *)s_buf,
                           (const double *)wre, (const double *)wim, rows, nuse, amp);
    return hipGetLastError();
}

}  // namespace demon
