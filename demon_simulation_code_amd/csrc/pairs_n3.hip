// pairs_n3.hip -- the single-cloud Newton-3 all-pairs kernel: one workgroup
// per (row S, slice of superblock offsets); the tile code and the work
// decomposition are in pairs_n3.h.
#include "pairs_n3.h"

namespace demon {

namespace {

using namespace pair;
using namespace n3;

template <bool F32, bool DTC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_pairs_n3(N3Row a, int32_t span,
                                                                                   ExpCoef ec) {
    __shared__ RedLds red;
    __shared__ ScrLds scr;
    __shared__ TBoxLds tbox;
    if ((a.flags & kFlagGated) && !a.sc->active) return;
    const int32_t nsb = a.npad / kSuper;
    const int32_t half = nsb / 2;
    // Workgroups go to the 8 XCDs round-robin by linear id, i.e. by
    // blockIdx.x mod 8 (nsb is a multiple of 8).  Giving XCD k the rows
    // [k nsb/8, (k+1) nsb/8) keeps the j superblocks its workgroups visit
    // (rows + 0 .. nsb/2) to 5/8 of the cloud, which its own L2 then holds.
    const int32_t S = (nsb & 7) ? (int32_t)blockIdx.x
                                : (int32_t)(blockIdx.x & 7) * (nsb >> 3) + (int32_t)(blockIdx.x >> 3);
    const int32_t d0 = (int32_t)blockIdx.y * span;
    const int32_t d1 = d0 + span < half + 1 ? d0 + span : half + 1;
    n3_row<F32, DTC>(a, ec, S, d0, d1, (int32_t)blockIdx.y, red, scr, tbox);
}

}  // namespace

hipError_t launch_pairs_n3(const PairLaunch &a, hipStream_t s) {
    const int64_t nsb = a.npad / kSuper;
    if (a.npad % kPad || !a.box || !a.react || !a.rflag || a.span < 1 || a.n > a.npad || (a.fp32pair && !a.qf))
        return hipErrorInvalidValue;
    if ((int64_t)a.slices * a.span < nsb / 2 + 1) return hipErrorInvalidValue;
    int32_t flags = 0;
    if (a.dtcheck) flags |= kFlagDtCheck;
    if (a.gated) flags |= kFlagGated;
    const dim3 grid((unsigned)nsb, (unsigned)a.slices);
    const N3Row row{a.xs, a.ys, a.nc, a.q, a.qf, a.orig, a.box, a.part, a.react, a.rflag, a.sc,
                    (int32_t)a.n, (int32_t)a.npad, flags, a.kappa, a.rc2p};
    auto kern = a.fp32pair ? (a.dtcheck ? k_pairs_n3<true, true> : k_pairs_n3<true, false>)
                           : (a.dtcheck ? k_pairs_n3<false, true> : k_pairs_n3<false, false>);
    hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, row, a.span, exp_coef());
    return hipGetLastError();
}

}  // namespace demon
