// Batched launches of the fused eSRRF block (npx_fused.h): several jobs in one launch.
//
// Used by the parameter sweep (one job per magnification and radius, or several radii of a small magnification
// packed into the wave slots of one block, where they share the gradient lattice; the sensitivities of a job share
// its ring sums) and by the serving loop (frame slices of several requests packed into one launch). Jobs travel by
// value in the kernel arguments; block b finds its job by a scan of the block offsets (at most kMaxJobs,
// wave-uniform) and decomposes its local index into region, phase group and frame chunk. Each job carries its own
// lattice half-width KL: a job whose waves all have ring half-width KL runs the body built for that K, a sweep job
// whose radii differ (vark) runs the body whose waves each take their own K <= KL.
#include <hip/hip_runtime.h>

#include "npx_fused.h"
#include "npx_kernels.h"

namespace npx {

// Lattice half-widths the bodies for mixed ring half-widths are built for: the smallest one >= every K of the job.
static int batched_kl(int kmax) { return kmax <= 2 ? 2 : (kmax <= 4 ? 4 : 6); }

using ckBatch = const __attribute__((address_space(4))) BatchArgs;

template <int MODE>
__global__ __launch_bounds__(kBatchThreads) void esrrf_batched(BatchArgs args) {
    extern __shared__ __align__(16) unsigned char smem[];
    (void)args;
    // the job table is read in place in the kernel-argument segment (scalar loads); a body that takes the job by
    // reference from the by-value parameter makes the compiler copy the 4 KiB block to scratch memory
    ckBatch& b = *(ckBatch*)__builtin_amdgcn_kernarg_segment_ptr();
    const int blk = blockIdx.x;
    int j = 0;
    while (j + 1 < b.njobs && blk >= b.start[j + 1]) ++j;
    j = __builtin_amdgcn_readfirstlane(j);
    const FusedArgs& a = *(const FusedArgs*)(&b.job[j]);
    int local = blk - b.start[j];
    const int nx = b.nregx[j], ny = b.nregy[j];
    const int bx = local % nx;
    local /= nx;
    const int by = local % ny;
    local /= ny;
    const int group = local % a.ngroups;
    const int chunk = local / a.ngroups;
    constexpr int Q = kQTemporal;
    if constexpr (MODE == kModeSweep) {
        if (a.vark) {
            switch (a.KL) {
                case 2: fused_block<2, Q, MODE, true>(a, bx, by, group, chunk, smem); return;
                case 4: fused_block<4, Q, MODE, true>(a, bx, by, group, chunk, smem); return;
                default: fused_block<6, Q, MODE, true>(a, bx, by, group, chunk, smem); return;
            }
        }
    }
    switch (a.KL) {
        case 0: fused_block<0, Q, MODE, false>(a, bx, by, group, chunk, smem); return;
        case 1: fused_block<1, Q, MODE, false>(a, bx, by, group, chunk, smem); return;
        case 2: fused_block<2, Q, MODE, false>(a, bx, by, group, chunk, smem); return;
        case 3: fused_block<3, Q, MODE, false>(a, bx, by, group, chunk, smem); return;
        case 4: fused_block<4, Q, MODE, false>(a, bx, by, group, chunk, smem); return;
        case 5: fused_block<5, Q, MODE, false>(a, bx, by, group, chunk, smem); return;
        default: fused_block<6, Q, MODE, false>(a, bx, by, group, chunk, smem); return;
    }
}

hipError_t launch_batched(const BatchArgs& b, int mode, hipStream_t st) {
    if (b.njobs < 1 || b.njobs > kMaxJobs) return hipErrorInvalidValue;
    if (mode != kModeTemporal && mode != kModeTemporalVar && mode != kModeSweep) return hipErrorInvalidValue;
    const int nw = kBatchThreads / 64;
    size_t lds = 0;
    for (int j = 0; j < b.njobs; ++j) {
        const FusedArgs& a = b.job[j];
        if (a.nwaves != nw || a.ngroups * a.nwaves < a.nphase * a.nrad) return hipErrorInvalidValue;
        if (a.nrad < 1 || a.nrad > kMaxRad || a.ns < 1 || a.ns > kMaxSens) return hipErrorInvalidValue;
        if (a.KL < 0 || a.KL > kMaxK) return hipErrorInvalidValue;
        if (a.vark) {
            int kmax = 0;
            for (int r = 0; r < a.nrad; ++r) kmax = a.K[r] > kmax ? a.K[r] : kmax;
            if (mode != kModeSweep || a.KL != batched_kl(kmax)) return hipErrorInvalidValue;
        } else {
            for (int r = 0; r < a.nrad; ++r)
                if (a.K[r] != a.KL) return hipErrorInvalidValue;
        }
        const size_t lj = fused_lds_bytes(a.KL, kQTemporal, nw, mode);
        lds = lj > lds ? lj : lds;
    }
    const int nblocks = b.start[b.njobs];
    if (nblocks <= 0) return hipSuccess;
    if (lds > static_cast<size_t>(kLdsLimit)) return hipErrorInvalidConfiguration;
    const dim3 grid(nblocks), block(kBatchThreads);
    if (mode == kModeTemporal)
        hipLaunchKernelGGL((esrrf_batched<kModeTemporal>), grid, block, lds, st, b);
    else if (mode == kModeTemporalVar)
        hipLaunchKernelGGL((esrrf_batched<kModeTemporalVar>), grid, block, lds, st, b);
    else
        hipLaunchKernelGGL((esrrf_batched<kModeSweep>), grid, block, lds, st, b);
    return hipGetLastError();
}

int batched_lattice_k(int kmax) { return batched_kl(kmax); }

}  // namespace npx
