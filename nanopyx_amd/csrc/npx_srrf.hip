// SRRF radiality on gfx950 (Radiality, src/nanopyx/core/transform/_le_radiality.pyx:86-126, per-pixel kernel
// src/include/_c_sr_radiality.c:15-51), and its C ABI entry point npx_radiality.
//
// One block = TI (64, one per lane) output columns x TJ output rows of one frame chunk, 4 waves; wave w takes
// the rows w, w + 4, ... For each frame the block stages its input tile (with a one-pixel halo, zero outside
// the image) in LDS, forms the central-difference gradients (Gx, Gy) of the tile in LDS, and evaluates every
// output pixel: 12 ring samples of the Catmull-Rom interpolated gradient, each a 4 x 4 tap sum read from the
// gradient tile. The tap bases and weights of a sample depend on the output column (x) or row (y) only, so they
// come from host tables built with the reference's float32 positions and float64 cubic weights: the column
// table of a lane sits in its VGPRs for the whole block, the row table of the block's rows in LDS (all lanes of
// a wave read the same row entry, a broadcast). Intensity weighting multiplies by the caller's xM image or, when
// none is given (the SRRF workflow), by the xM Catmull-Rom magnification computed here from the input tile.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/npx_esrrf.h"
#include "npx_host.h"
#include "npx_plan.h"

namespace npx {

constexpr int kSrrfThreads = 256;
constexpr int kSrrfTI = 64;
constexpr int kSrrfTJ = 32;
constexpr int kSE = SrrfEntry::kFloats;

struct SrrfArgs {
    const void* in;
    const float* interp;  // (N, HM, WM) intensity image, or nullptr: magnify the input in-kernel
    float* out;
    const float* row;     // [HM][kSE]
    const float* col;     // [WM][kSE]
    const int* tile_r0;   // [row bands]
    const int* tile_c0;   // [column tiles]
    int u16, N, H, W, M, HM, WM, fpc;
    int TH, TW;
    int j_lo, j_hi, i_lo, i_hi;
    int positivity, iw;
    float R;
    float xr[12], yr[12];
};

__host__ __device__ inline int srrf_align4(int n) { return (n + 3) & ~3; }

// LDS layout (floats): input tile (TH + 2) x (TW + 2) | gradient tile TH x TW float2 | row table TJ x kSE
__host__ __device__ inline int srrf_in_floats(int TH, int TW) { return srrf_align4((TH + 2) * (TW + 2)); }
inline size_t srrf_lds_bytes(int TH, int TW) {
    return sizeof(float) * (static_cast<size_t>(srrf_in_floats(TH, TW)) + 2 * TH * TW + kSrrfTJ * kSE);
}

template <bool U16, bool INTERP>
__global__ __launch_bounds__(kSrrfThreads) void srrf_radiality(SrrfArgs a) {
    extern __shared__ float4 smem4[];
    float* const s_in = reinterpret_cast<float*>(smem4);
    float2* const s_g = reinterpret_cast<float2*>(s_in + srrf_in_floats(a.TH, a.TW));
    float* const s_row = reinterpret_cast<float*>(s_g + a.TH * a.TW);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int bx = blockIdx.x, by = blockIdx.y;
    const int i = bx * kSrrfTI + lane;
    const int r0 = a.tile_r0[by], c0 = a.tile_c0[bx];
    const int IW = a.TW + 2, IH = a.TH + 2;
    const int f_begin = blockIdx.z * a.fpc;
    const int f_end = min(a.N, f_begin + a.fpc);

    // row table of the band
    for (int t = threadIdx.x; t < kSrrfTJ * kSE; t += kSrrfThreads) {
        const int j = by * kSrrfTJ + t / kSE;
        s_row[t] = j < a.HM ? a.row[static_cast<size_t>(j) * kSE + t % kSE] : 0.f;
    }
    // column table of this lane's column, in VGPRs
    const float* ce = a.col + static_cast<size_t>(min(i, a.WM - 1)) * kSE;
    int cb[12];
    float wx[12][4], dx[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        cb[k] = __float_as_int(ce[SrrfEntry::kBase + k]);
#pragma unroll
        for (int u = 0; u < 4; ++u) wx[k][u] = ce[SrrfEntry::kW + 4 * k + u];
        dx[k] = ce[SrrfEntry::kD + k];
    }
    const int cbi = __float_as_int(ce[SrrfEntry::kIBase]);
    float wci[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) wci[u] = ce[SrrfEntry::kIW + u];
    const bool col_in = i >= a.i_lo && i < a.i_hi;
    const float invR = 1.0f / a.R;

    const size_t plane = static_cast<size_t>(a.H) * a.W, planeM = static_cast<size_t>(a.HM) * a.WM;
    for (int f = f_begin; f < f_end; ++f) {
        __syncthreads();  // the previous frame is done with the tiles (and the row table is staged)
        // (1) input tile rows r0-1 .. r0+TH, columns c0-1 .. c0+TW; 0 outside the image
        for (int t = threadIdx.x; t < IH * IW; t += kSrrfThreads) {
            const int y = r0 - 1 + t / IW, x = c0 - 1 + t % IW;
            float v = 0.f;
            if (y >= 0 && y < a.H && x >= 0 && x < a.W) {
                const size_t o = static_cast<size_t>(f) * plane + static_cast<size_t>(y) * a.W + x;
                v = U16 ? static_cast<float>(static_cast<const uint16_t*>(a.in)[o])
                        : static_cast<const float*>(a.in)[o];
            }
            s_in[t] = v;
        }
        __syncthreads();
        // (2) gradients (_c_gradients.c:3-25): interior pixels only, 0 on the image border and outside
        for (int t = threadIdx.x; t < a.TH * a.TW; t += kSrrfThreads) {
            const int r = t / a.TW, c = t % a.TW;
            const int y = r0 + r, x = c0 + c;
            float2 g = make_float2(0.f, 0.f);
            if (y >= 1 && y < a.H - 1 && x >= 1 && x < a.W - 1) {
                const float* p = s_in + (r + 1) * IW + (c + 1);
                g.x = -p[-1] + p[1];
                g.y = -p[-IW] + p[IW];
            }
            s_g[t] = g;
        }
        __syncthreads();
        // (3) radiality of this wave's rows
        for (int jj = wv; jj < kSrrfTJ; jj += kSrrfThreads / 64) {
            const int j = by * kSrrfTJ + jj;
            if (j >= a.HM) break;
            float v = 0.f;
            if (col_in && j >= a.j_lo && j < a.j_hi) {
                const float* re = s_row + jj * kSE;
                float div = 0.f;
#pragma unroll
                for (int k = 0; k < 12; ++k) {
                    const int rb = __float_as_int(re[SrrfEntry::kBase + k]);
                    const float4 wy = *reinterpret_cast<const float4*>(re + SrrfEntry::kW + 4 * k);
                    const float dy = re[SrrfEntry::kD + k];
                    const float2* g = s_g + rb * a.TW + cb[k];
                    // _c_interpolate order: columns outer, rows inner
                    float qx = 0.f, qy = 0.f;
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const float2 t0 = g[u], t1 = g[a.TW + u], t2 = g[2 * a.TW + u], t3 = g[3 * a.TW + u];
                        float px = t0.x * wy.x;
                        float py = t0.y * wy.x;
                        px = fmaf(t1.x, wy.y, px);
                        py = fmaf(t1.y, wy.y, py);
                        px = fmaf(t2.x, wy.z, px);
                        py = fmaf(t2.y, wy.z, py);
                        px = fmaf(t3.x, wy.w, px);
                        py = fmaf(t3.y, wy.w, py);
                        qx = fmaf(px, wx[k][u], qx);
                        qy = fmaf(py, wx[k][u], qy);
                    }
                    // Dk = (1 - |G x (xc - x0, yc - y0)| / (|G| R))^2, 0 where |G| = 0 (_c_sr_radiality.c:6-13)
                    const float g2 = fmaf(qx, qx, qy * qy);
                    const float cr = fabsf(fmaf(qy, dx[k], -qx * dy));
                    float dk = fmaf(-cr * __builtin_amdgcn_rsqf(g2), invR, 1.f);
                    dk = g2 > 0.f ? dk * dk : 0.f;
                    const float dot = fmaf(qx, a.xr[k], qy * a.yr[k]);
                    div += dot > 0.f ? -dk : dk;
                }
                div = div / 12.f;
                v = a.positivity ? fmaxf(div, 0.f) : div;
                if (a.iw) {
                    float I;
                    if (INTERP) {
                        I = a.interp[static_cast<size_t>(f) * planeM + static_cast<size_t>(j) * a.WM + i];
                    } else {
                        // xM Catmull-Rom of the input at (j / M, i / M), columns outer, rows inner
                        const int rbi = __float_as_int(re[SrrfEntry::kIBase]);
                        const float4 wr = *reinterpret_cast<const float4*>(re + SrrfEntry::kIW);
                        const float* p = s_in + rbi * IW + cbi;
                        I = 0.f;
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            float s = p[u] * wr.x;
                            s = fmaf(p[IW + u], wr.y, s);
                            s = fmaf(p[2 * IW + u], wr.z, s);
                            s = fmaf(p[3 * IW + u], wr.w, s);
                            I = fmaf(s, wci[u], I);
                        }
                    }
                    v *= I;
                }
            }
            if (i < a.WM) a.out[static_cast<size_t>(f) * planeM + static_cast<size_t>(j) * a.WM + i] = v;
        }
    }
}

// ------------------------------------------------------------------------------------------ host side

namespace {

using npx::host::fail;
using npx::host::hip_fail;

struct SrrfDev {
    void* mem = nullptr;
    const float* row = nullptr;
    const float* col = nullptr;
    const int* tile_r0 = nullptr;
    const int* tile_c0 = nullptr;
};

struct SrrfCached {
    SrrfPlan host;
    std::map<int, SrrfDev> dev;
};

using SKey = std::tuple<int, int, int, uint64_t, int>;
std::mutex s_mu;
std::map<SKey, std::unique_ptr<SrrfCached>> s_plans;

uint64_t dbits(double v) {
    uint64_t b;
    std::memcpy(&b, &v, sizeof b);
    return b;
}

// Cached plan of (H, W, M, ringRadius, border) with its tables on the current device (uploaded once, blocking).
int srrf_plan(int H, int W, int M, double ringRadius, int border, hipStream_t st, const SrrfPlan** hp,
              const SrrfDev** dp) {
    std::lock_guard<std::mutex> lk(s_mu);
    const SKey key{H, W, M, dbits(ringRadius), border};
    auto it = s_plans.find(key);
    if (it == s_plans.end()) {
        auto p = std::make_unique<SrrfCached>();
        std::string err;
        if (!build_srrf_plan(H, W, M, ringRadius, border, kSrrfTJ, kSrrfTI, p->host, &err))
            return fail(NPX_ERR_ARG, err);
        if (srrf_lds_bytes(p->host.TH, p->host.TW) > 160 * 1024)
            return fail(NPX_ERR_UNSUPPORTED, "ring radius too large for the radiality tile");
        it = s_plans.emplace(key, std::move(p)).first;
    }
    SrrfCached* c = it->second.get();
    int d = 0;
    hipError_t e = hipGetDevice(&d);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    auto dv = c->dev.find(d);
    if (dv == c->dev.end()) {
        const SrrfPlan& h = c->host;
        const size_t b_row = h.row.size() * sizeof(float), b_col = h.col.size() * sizeof(float);
        const size_t b_r0 = h.tile_r0.size() * sizeof(int), b_c0 = h.tile_c0.size() * sizeof(int);
        std::vector<uint8_t> blob(b_row + b_col + b_r0 + b_c0);
        std::memcpy(blob.data(), h.row.data(), b_row);
        std::memcpy(blob.data() + b_row, h.col.data(), b_col);
        std::memcpy(blob.data() + b_row + b_col, h.tile_r0.data(), b_r0);
        std::memcpy(blob.data() + b_row + b_col + b_r0, h.tile_c0.data(), b_c0);
        if (npx::host::capturing(st)) return fail(NPX_ERR_ARG, "first use of a radiality plan under capture");
        void* mem = nullptr;
        e = hipMalloc(&mem, blob.size());
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(radiality tables)");
        e = hipMemcpy(mem, blob.data(), blob.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e != hipSuccess) {
            (void)hipFree(mem);
            return hip_fail(e, "upload of radiality tables");
        }
        SrrfDev sd;
        auto* base = static_cast<uint8_t*>(mem);
        sd.mem = mem;
        sd.row = reinterpret_cast<const float*>(base);
        sd.col = reinterpret_cast<const float*>(base + b_row);
        sd.tile_r0 = reinterpret_cast<const int*>(base + b_row + b_col);
        sd.tile_c0 = reinterpret_cast<const int*>(base + b_row + b_col + b_r0);
        dv = c->dev.emplace(d, sd).first;
    }
    *hp = &c->host;
    *dp = &dv->second;
    return NPX_OK;
}

int g_srrf_target_blocks = 2048;

}  // namespace
}  // namespace npx

extern "C" {

int npx_radiality(const void* d_img, int dtype, const float* d_interp, int N, int H, int W, int M, double ringRadius,
                  int border, int positivity, int iw, float* d_out, void* stream) {
    using namespace npx;
    if (!d_img || !d_out) return host::fail(NPX_ERR_ARG, "NULL array pointer");
    if (dtype != NPX_DTYPE_F32 && dtype != NPX_DTYPE_U16) return host::fail(NPX_ERR_ARG, "dtype must be f32 or u16");
    if (N < 1 || H < 1 || W < 1 || M < 1) return host::fail(NPX_ERR_ARG, "bad shape");
    if (border < 0) return host::fail(NPX_ERR_ARG, "border must be >= 0");
    const SrrfPlan* h = nullptr;
    const SrrfDev* dv = nullptr;
    hipStream_t st = static_cast<hipStream_t>(stream);
    int rc = srrf_plan(H, W, M, ringRadius, border, st, &h, &dv);
    if (rc != NPX_OK) return rc;
    SrrfArgs a{};
    a.in = d_img;
    a.interp = d_interp;
    a.out = d_out;
    a.row = dv->row;
    a.col = dv->col;
    a.tile_r0 = dv->tile_r0;
    a.tile_c0 = dv->tile_c0;
    a.u16 = dtype == NPX_DTYPE_U16;
    a.N = N;
    a.H = H;
    a.W = W;
    a.M = M;
    a.HM = h->HM;
    a.WM = h->WM;
    a.TH = h->TH;
    a.TW = h->TW;
    a.j_lo = h->j_lo;
    a.j_hi = h->j_hi;
    a.i_lo = h->i_lo;
    a.i_hi = h->i_hi;
    a.positivity = positivity ? 1 : 0;
    a.iw = iw ? 1 : 0;
    a.R = h->R;
    for (int k = 0; k < 12; ++k) {
        a.xr[k] = h->xr[k];
        a.yr[k] = h->yr[k];
    }
    const int ntx = static_cast<int>(h->tile_c0.size()), nty = static_cast<int>(h->tile_r0.size());
    const int tiles = ntx * nty;
    int chunks = std::max(1, std::min(N, (g_srrf_target_blocks + tiles - 1) / tiles));
    a.fpc = (N + chunks - 1) / chunks;
    chunks = (N + a.fpc - 1) / a.fpc;
    const size_t lds = srrf_lds_bytes(h->TH, h->TW);
    const dim3 grid(ntx, nty, chunks), block(kSrrfThreads);
    if (a.u16) {
        if (d_interp) hipLaunchKernelGGL((srrf_radiality<true, true>), grid, block, lds, st, a);
        else hipLaunchKernelGGL((srrf_radiality<true, false>), grid, block, lds, st, a);
    } else {
        if (d_interp) hipLaunchKernelGGL((srrf_radiality<false, true>), grid, block, lds, st, a);
        else hipLaunchKernelGGL((srrf_radiality<false, false>), grid, block, lds, st, a);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return host::hip_fail(e, "radiality launch");
    return NPX_OK;
}

int npx_radiality_info(int H, int W, int M, double ringRadius, int border, int* info, float* ring) {
    using namespace npx;
    if (!info || !ring) return host::fail(NPX_ERR_ARG, "NULL output");
    SrrfPlan p;
    std::string err;
    if (!build_srrf_plan(H, W, M, ringRadius, border, kSrrfTJ, kSrrfTI, p, &err)) return host::fail(NPX_ERR_ARG, err);
    const int v[12] = {p.HM, p.WM, p.j_lo, p.j_hi, p.i_lo, p.i_hi, p.TH, p.TW, kSrrfTJ, kSrrfTI,
                       static_cast<int>(p.tile_r0.size()), static_cast<int>(p.tile_c0.size())};
    std::memcpy(info, v, sizeof v);
    ring[0] = p.R;
    for (int k = 0; k < 12; ++k) {
        ring[1 + k] = p.xr[k];
        ring[13 + k] = p.yr[k];
    }
    return NPX_OK;
}

int npx_radiality_tables(int H, int W, int M, double ringRadius, int border, float* row, float* col, int* tile_r0,
                         int* tile_c0) {
    using namespace npx;
    if (!row || !col || !tile_r0 || !tile_c0) return host::fail(NPX_ERR_ARG, "NULL output");
    SrrfPlan p;
    std::string err;
    if (!build_srrf_plan(H, W, M, ringRadius, border, kSrrfTJ, kSrrfTI, p, &err)) return host::fail(NPX_ERR_ARG, err);
    std::memcpy(row, p.row.data(), p.row.size() * sizeof(float));
    std::memcpy(col, p.col.data(), p.col.size() * sizeof(float));
    std::memcpy(tile_r0, p.tile_r0.data(), p.tile_r0.size() * sizeof(int));
    std::memcpy(tile_c0, p.tile_c0.data(), p.tile_c0.size() * sizeof(int));
    return NPX_OK;
}

}  // extern "C"
