// Batched launches of the fused eSRRF block (npx_fused.h): several jobs in one launch.
//
// Used by the parameter sweep (one job per magnification; its radii share the gradient lattice and its
// sensitivities share the ring sums) and by the serving loop (frame slices of several requests packed into
// one launch). Jobs travel by value in the kernel arguments; block b finds its job by a scan of the block
// offsets (at most kMaxJobs, wave-uniform) and decomposes its local index into region, phase group and
// frame chunk. The lattice is built for the largest ring half-width KL of the launch; each wave runs the
// ring sum of its own radius.
#include <hip/hip_runtime.h>

#include "npx_fused.h"
#include "npx_kernels.h"

namespace npx {

template <int KL, int MODE>
__global__ __launch_bounds__(kBatchThreads) void esrrf_batched(BatchArgs b) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int blk = blockIdx.x;
    int j = 0;
    while (j + 1 < b.njobs && blk >= b.start[j + 1]) ++j;
    j = __builtin_amdgcn_readfirstlane(j);
    const FusedArgs& a = b.job[j];
    int local = blk - b.start[j];
    const int nx = b.nregx[j], ny = b.nregy[j];
    const int bx = local % nx;
    local /= nx;
    const int by = local % ny;
    local /= ny;
    const int group = local % a.ngroups;
    const int chunk = local / a.ngroups;
    fused_block<KL, kQTemporal, MODE, true>(a, bx, by, group, chunk, smem);
}

template <int KL>
static hipError_t launch_batched_k(const BatchArgs& b, int mode, dim3 grid, size_t lds, hipStream_t st) {
    const dim3 block(kBatchThreads);
    if (mode == kModeTemporal)
        hipLaunchKernelGGL((esrrf_batched<KL, kModeTemporal>), grid, block, lds, st, b);
    else if (mode == kModeTemporalVar)
        hipLaunchKernelGGL((esrrf_batched<KL, kModeTemporalVar>), grid, block, lds, st, b);
    else if (mode == kModeSweep)
        hipLaunchKernelGGL((esrrf_batched<KL, kModeSweep>), grid, block, lds, st, b);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// Lattice half-widths the batched kernels are built for: the launch uses the smallest one >= every job's K.
static int batched_kl(int kmax) { return kmax <= 2 ? 2 : (kmax <= 4 ? 4 : 6); }

hipError_t launch_batched(const BatchArgs& b, int mode, hipStream_t st) {
    if (b.njobs < 1 || b.njobs > kMaxJobs) return hipErrorInvalidValue;
    int kmax = 0;
    for (int j = 0; j < b.njobs; ++j) {
        const FusedArgs& a = b.job[j];
        if (a.nwaves * 64 != kBatchThreads || a.ngroups * a.nwaves < a.nphase * a.nrad) return hipErrorInvalidValue;
        if (a.nrad < 1 || a.nrad > kMaxRad || a.ns < 1 || a.ns > kMaxSens) return hipErrorInvalidValue;
        for (int r = 0; r < a.nrad; ++r) kmax = a.K[r] > kmax ? a.K[r] : kmax;
    }
    if (kmax > kMaxK) return hipErrorInvalidValue;
    const int KL = batched_kl(kmax);
    for (int j = 0; j < b.njobs; ++j)
        if (b.job[j].KL != KL) return hipErrorInvalidValue;
    const int nblocks = b.start[b.njobs];
    if (nblocks <= 0) return hipSuccess;
    const size_t lds = fused_lds_bytes(KL, kQTemporal, kBatchThreads / 64, mode);
    if (lds > static_cast<size_t>(kLdsLimit)) return hipErrorInvalidConfiguration;
    const dim3 grid(nblocks);
    switch (KL) {
        case 2: return launch_batched_k<2>(b, mode, grid, lds, st);
        case 4: return launch_batched_k<4>(b, mode, grid, lds, st);
        default: return launch_batched_k<6>(b, mode, grid, lds, st);
    }
}

int batched_lattice_k(int kmax) { return batched_kl(kmax); }

}  // namespace npx
