// Kernel-side types shared by the launchers (npx_kernels.hip) and the C-ABI (npx_capi.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace npx {

constexpr int kModeFrames = 0;    // per-frame RGC maps (eSRRF_ST.run contract)
constexpr int kModeTemporal = 1;  // AVG / TAC2 moments streamed over the frames
constexpr int kModeTemporalVar = 2;  // ... plus the sum of squared deviations for VAR
constexpr int kQFrames = 16;      // input rows per lane, per-frame-map kernel
constexpr int kQTemporal = 8;     // input rows per lane, temporal kernel (5 moments per pixel in registers)
constexpr int kPrefetch = 4;      // input-tile elements per thread prefetched for the next frame

// LDS tile geometry of the fused kernel for window half-width K and Q rows per lane.
template <int K, int Q>
struct Geometry {
    static constexpr int HG = K / 2 + 3;          // gradient-tile halo (input pixels)
    static constexpr int HI = HG + 1;             // input-tile halo (Roberts reads row/col - 1)
    static constexpr int TIR = Q + 2 * HI, TIC = 64 + 2 * HI;
    static constexpr int TGR = Q + 2 * HG, TGC = 64 + 2 * HG;
    static constexpr int LR = 2 * Q + 2 * K;      // lattice rows per block
    static constexpr int LP = 64 + K;             // lattice float2 per row in one column-parity plane
    static constexpr int NW = 2 * K + 1;
    static constexpr int RS = ((2 * NW + 3 * NW * NW + 1 + (NW + 1) * (NW + 1)) + 15) / 16 * 16;  // ring block floats
};

// Device copies of the plan tables (npx_plan.h).
struct DevTables {
    const float* ring;   // [phase] blocks: dy[j], dx[i], (d, t)[i][j], w[i][j]
    const float* wsum;   // [phase]
    const float4* lrw;   // lattice row taps [2H+1]
    const int* lrb;
    const float4* lcw;   // lattice column taps [2W+1]
    const int* lcb;
    const float4* crw;   // intensity row taps [HM]
    const int* crb;
    const float4* ccw;   // intensity column taps [WM]
    const int* ccb;
};

struct FusedArgs {
    const void* in;
    int u16;
    int N, H, W, M, HM, WM, margin;
    int nphase, nwaves, ngroups;
    int fpc, nchunk;  // frames per chunk (blockIdx.z / ngroups = chunk)
    int pow_mode;
    float S;
    int iw;
    DevTables t;
    float* out;   // frames mode: (N, HM, WM)
    float* part;  // temporal mode, nchunk > 1: [nchunk][5][HM*WM]
    float* avg;
    float* var;
    float* tac2;
};

hipError_t launch_fused(const FusedArgs& a, int K, dim3 grid, dim3 block, int mode, hipStream_t st);
hipError_t launch_merge(const float* part, int nchunk, int fpc, int N, size_t P, float* avg, float* var, float* tac2,
                        hipStream_t st);
hipError_t launch_rgc_arrays(const float* gx, const float* gy, const float* img, int N, int HM, int WM, int M,
                             float fwhm, float tSO, float tSS, int pow_mode, float S, int iw, int margin, float* out,
                             hipStream_t st);
hipError_t launch_moments_stack(const float* stack, int N, size_t P, float* avg, float* var, float* tac2,
                                hipStream_t st);
hipError_t launch_macro_pixel(float* img, int S, int HM, int WM, int M, float* cmap, hipStream_t st);

}  // namespace npx
