// Kernel-side types shared by the launchers (npx_kernels.hip) and the C-ABI (npx_capi.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace npx {

constexpr int kModeFrames = 0;       // per-frame RGC maps (eSRRF_ST.run contract)
constexpr int kModeTemporal = 1;     // AVG / TAC2 moments streamed over the frames
constexpr int kModeTemporalVar = 2;  // ... plus the sum of squared deviations for VAR
constexpr int kModeSweep = 3;        // AVG for up to kMaxRad radii x kMaxSens sensitivities of one magnification
constexpr int kModeArrays = 4;       // per-frame RGC maps from caller gradients sampled x2M and a magnified image
constexpr int kMaxRad = 4;           // radii of one sweep job (they share the gradient lattice)
constexpr int kMaxSens = 4;          // sensitivities of one sweep job (they share the ring sums)
constexpr int kMaxJobs = 8;          // jobs of one batched launch (passed by value as kernel arguments)
constexpr int kMaxK = 6;             // largest ring half-width int(2 * radius) the kernels are built for
constexpr int kQFrames = 16;         // input rows per lane, per-frame-map kernel of one 16-wave block per region
constexpr int kQTemporal = 8;        // input rows per lane, temporal and sweep kernels
constexpr int kPrefetch = 4;         // input-tile elements per thread prefetched for the next frame
constexpr int kBatchThreads = 1024;  // block size of batched launches (16 wavefronts)
constexpr float kRingScale = 65536.f;  // s: the lattice holds s G/|G|, the indicator is clamp(-s u / |dx|, 0, 1)

// LDS tile geometry of the fused kernel for lattice half-width K and Q rows per lane.
template <int K, int Q>
struct Geometry {
    static constexpr int HG = K / 2 + 3;          // gradient-tile halo (input pixels)
    static constexpr int HI = HG + 1;             // input-tile halo (Roberts reads row/col - 1)
    static constexpr int TIR = Q + 2 * HI, TIC = 64 + 2 * HI;
    static constexpr int TGR = Q + 2 * HG, TGC = 64 + 2 * HG;
    static constexpr int LR = 2 * Q + 2 * K;      // lattice rows per block
    static constexpr int LP = 64 + K;             // lattice float2 per row in one column-parity plane
    static constexpr int NW = 2 * K + 1;
    static constexpr int LB = 128 + 2 * K;        // lattice columns per block
    static constexpr int RS = ((2 * NW + 3 * NW * NW + 1 + (NW + 1) * (NW + 1)) + 15) / 16 * 16;  // ring block floats
    static constexpr int RC = ((2 * NW + 2 * NW * NW) + 3) / 4 * 4;    // LDS floats per wave: dy, dx, (d, t) pairs
    static constexpr int RN = ((1 + (NW + 1) * (NW + 1)) + 3) / 4 * 4;  // ... then wsum + prefix sums
    static constexpr int RW = RC + RN;
};

constexpr int kLdsLimit = 160 * 1024;  // LDS bytes one workgroup may use on gfx950

// Device copies of the plan tables (npx_plan.h).
struct DevTables {
    const float4* lrw;   // lattice row taps [2H+1]
    const int* lrb;
    const float4* lcw;   // lattice column taps [2W+1]
    const int* lcb;
    const float4* crw;   // intensity row taps [HM]
    const int* crb;
    const float4* ccw;   // intensity column taps [WM]
    const int* ccb;
};

// One eSRRF job: a stack (or a frame range of it) with its parameters, tables and outputs. The single-job
// kernels take it by value; a batched launch takes up to kMaxJobs of them by value.
//
// Waves of a block are assigned output phases radius by radius: wave slot g = group * nwaves + wave covers
// radius g / nphase and phase g % nphase, so every radius of a sweep job reads the same gradient lattice.
struct FusedArgs {
    const void* in;   // (N_total, H, W) float32 or uint16
    int u16;
    int N, H, W, M, HM, WM;  // N: one past the last frame this job reads
    int f0;                  // first frame this job reads
    int nphase, nwaves, ngroups;
    int fpc, nchunk;         // frames per chunk; chunk c covers [f0 + c * fpc, min(N, f0 + (c + 1) * fpc))
    int chunk0;              // partial-result slot of chunk 0 (frame slices of one stack share one partial array)
    int direct;              // 1: write the final maps; 0: write per-chunk partials to `part` or the running state
    int n0;                  // temporal: frames already folded into `state` (resumed when > 0)
    int iw;
    int KL;                  // lattice half-width (max of K[])
    int vark;                // batched sweep jobs: 1 when the radii's half-widths K[] differ (each <= KL, KL from
                             // batched_lattice_k); 0: every K[] equals KL
    int nrad;
    int ns;
    // taps of the gradient lattice (depend on H, W, M) and of the intensity Catmull-Rom (H, W, M)
    const float4* lrw;
    const int* lrb;
    const float4* lcw;
    const int* lcb;
    const float4* crw;
    const int* crb;
    const float4* ccw;
    const int* ccb;
    // per radius: half-width, margin, and the reference's tSS = 2 sigma^2 and tSO = 2 sigma + 1 (float32) from which
    // the kernels form each block's ring geometry
    int K[kMaxRad];
    int margin[kMaxRad];
    float tss[kMaxRad];
    float tso[kMaxRad];
    // per sensitivity: value and power mode (0: none, 1: general, 2/3/4: integer)
    float sens[kMaxSens];
    int sens_pow[kMaxSens];
    // outputs
    float* out;   // frames mode: (N, HM, WM)
    float* part;  // temporal: [slot][5][HM*WM]; sweep: [slot][nrad][ns][HM*WM]
    float* state; // temporal, frame slices of a served job: running moments [5][HM*WM] (mean, M2, co-moment,
                  // first value, previous value), read when n0 > 0 and written unless `direct`
    float* avg;
    float* var;
    float* tac2;
    float* sweep_avg[kMaxRad * kMaxSens];  // sweep: AVG map of (radius r, sensitivity s) at [r * kMaxSens + s]
    // arrays mode: column and row gradients (N, 2 HM, 2 WM) and the intensity image (N, HM, WM) of the caller;
    // alat = 1: the gradients hold only the lattice, compact (N, 2H + 1, 2W + 1): rows / columns k M and the last one
    const float* agx;
    const float* agy;
    const float* aimg;
    int alat;
};

// Several jobs in one launch: job j owns blocks [start[j], start[j + 1]), laid out as
// ((chunk * ngroups + group) * nregy + by) * nregx + bx.
struct BatchArgs {
    int njobs;
    int start[kMaxJobs + 1];
    int nregx[kMaxJobs];
    int nregy[kMaxJobs];
    FusedArgs job[kMaxJobs];
};
static_assert(sizeof(BatchArgs) <= 4096, "batched kernel arguments must fit the 4 KiB kernel-argument limit");

// Input rows per lane of the per-frame-map and arrays kernels: the tall tile (kQFrames) when one block of 16 waves
// covers every phase of a region (M = 4: one block per CU either way, and a taller tile spreads the lattice build over
// more rows); otherwise kQTemporal, whose smaller LDS footprint lets more blocks share a CU (M = 2: 4-wave blocks, 4
// blocks per CU instead of 2).
inline int frames_q(int nwaves, int ngroups) { return nwaves == 16 && ngroups == 1 ? kQFrames : kQTemporal; }

// Dynamic LDS bytes of the fused kernels for lattice half-width KL, Q rows per lane, nwaves per block.
size_t fused_lds_bytes(int KL, int Q, int nwaves, int mode);

hipError_t launch_fused(const FusedArgs& a, dim3 grid, dim3 block, int mode, hipStream_t st);
hipError_t launch_batched(const BatchArgs& b, int mode, hipStream_t st);
int batched_lattice_k(int kmax);  // lattice half-width of a batched sweep job of mixed half-widths, widest kmax
hipError_t launch_means_merge(const float* part, int nchunk, int fpc, int N, size_t P, int nout, float* const* outs,
                              hipStream_t st);
hipError_t launch_merge(const float* part, int nchunk, int fpc, int N, size_t P, float* avg, float* var, float* tac2,
                        hipStream_t st);
// ... continuing a running state of n0 frames ([5][P]); to_state: write the merged moments back to the state.
hipError_t launch_merge_state(const float* part, int nchunk, int fpc, int N, size_t P, float* state, int n0,
                              int to_state, float* avg, float* var, float* tac2, hipStream_t st);
hipError_t launch_rgc_arrays(const float* gx, const float* gy, const float* img, int N, int HM, int WM, int M,
                             float fwhm, float tSO, float tSS, int pow_mode, float S, int iw, int margin, float* out,
                             hipStream_t st);
hipError_t launch_moments_stack(const float* stack, int N, size_t P, float* avg, float* var, float* tac2,
                                hipStream_t st);
hipError_t launch_macro_pixel(float* img, int S, int HM, int WM, int M, float* cmap, hipStream_t st);

}  // namespace npx
