// SRRF radiality on gfx950 (Radiality, src/nanopyx/core/transform/_le_radiality.pyx:86-126, per-pixel kernel
// src/include/_c_sr_radiality.c:15-51), and its C ABI entry point npx_radiality.
//
// One block = TI (64, one per lane) output columns x TJ output rows of one frame chunk, 8 waves. For each frame the
// block stages its input tile (one-pixel halo, zero outside the image) in LDS and forms the central-difference
// gradients (Gx, Gy) of the tile. A ring sample is a 4 x 4 Catmull-Rom sum of the gradient at
// ((i + 0.5 + xRing) / M, (j + 0.5 + yRing) / M); its x taps and weights depend on the output column i only and its
// y taps on the output row j only, so the sum is taken in two passes per group of 4 samples: a row pass
// interpolates every gradient-tile row at each lane's sample x position (RX, in LDS), then the column pass
// combines 4 RX rows per output pixel and sample. Tap bases and weights come from host tables built with the
// reference's float32 positions and float64 cubic weights: a lane's column entry sits in its VGPRs for the whole
// block, the block's row entries in LDS (one address per wave: a broadcast). The running ring sum of each pixel
// stays in LDS between sample groups. Intensity weighting multiplies by the caller's xM image or, when none is
// given (the SRRF workflow), by the xM Catmull-Rom magnification computed here from the input tile.
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

constexpr int kSrrfThreads = 512;
constexpr int kSrrfWaves = kSrrfThreads / 64;
constexpr int kSrrfTI = 64;
constexpr int kSrrfGroup = 4;  // ring samples per row-pass group (12 = 3 groups)
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
    int TJ, TH, TW;
    int j_lo, j_hi, i_lo, i_hi;
    int positivity, iw;
};

// LDS layout (floats): input tile (TH + 2) x (TW + 2) | gradient tile TH x TW float2 | RX kSrrfGroup x TH x 64
// float2 | row table TJ x kSE | ring sums TJ x 64
__host__ __device__ inline int srrf_in_floats(int TH, int TW) { return ((TH + 2) * (TW + 2) + 3) & ~3; }
__host__ __device__ inline int srrf_g_floats(int TH, int TW) { return (2 * TH * TW + 3) & ~3; }
inline size_t srrf_lds_bytes(int TJ, int TH, int TW) {
    return sizeof(float) * (static_cast<size_t>(srrf_in_floats(TH, TW)) + srrf_g_floats(TH, TW) +
                            2 * kSrrfGroup * TH * 64 + TJ * kSE + TJ * 64);
}

template <bool U16, bool INTERP>
__global__ __launch_bounds__(kSrrfThreads) void srrf_radiality(SrrfArgs a) {
    extern __shared__ float4 smem4[];
    const int TJ = a.TJ, TH = a.TH, TW = a.TW;
    float* const s_in = reinterpret_cast<float*>(smem4);
    float2* const s_g = reinterpret_cast<float2*>(s_in + srrf_in_floats(TH, TW));
    float2* const s_rx = reinterpret_cast<float2*>(reinterpret_cast<float*>(s_g) + srrf_g_floats(TH, TW));
    float* const s_row = reinterpret_cast<float*>(s_rx + kSrrfGroup * TH * 64);
    float* const s_div = s_row + TJ * kSE;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int bx = blockIdx.x, by = blockIdx.y;
    const int i = bx * kSrrfTI + lane;
    const int r0 = a.tile_r0[by], c0 = a.tile_c0[bx];
    const int IW = TW + 2, IH = TH + 2;
    const int f_begin = blockIdx.z * a.fpc;
    const int f_end = min(a.N, f_begin + a.fpc);

    // row table of the band
    for (int t = threadIdx.x; t < TJ * kSE; t += kSrrfThreads) {
        const int j = by * TJ + t / kSE;
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
        for (int t = threadIdx.x; t < TH * TW; t += kSrrfThreads) {
            const int r = t / TW, c = t % TW;
            const int y = r0 + r, x = c0 + c;
            float2 g = make_float2(0.f, 0.f);
            if (y >= 1 && y < a.H - 1 && x >= 1 && x < a.W - 1) {
                const float* p = s_in + (r + 1) * IW + (c + 1);
                g.x = -p[-1] + p[1];
                g.y = -p[-IW] + p[IW];
            }
            s_g[t] = g;
        }
#pragma unroll
        for (int grp = 0; grp < 12 / kSrrfGroup; ++grp) {
            __syncthreads();  // gradient tile ready / the previous group's column pass is done with RX
            // (3) row pass: RX[kk][r][lane] = sum_u wx[k][u] G[r][cb[k] + u] for the group's samples
            for (int r = wv; r < TH; r += kSrrfWaves) {
                const float2* gr = s_g + r * TW;
#pragma unroll
                for (int kk = 0; kk < kSrrfGroup; ++kk) {
                    const int k = grp * kSrrfGroup + kk;
                    const float2* gp = gr + cb[k];
                    const float2 t0 = gp[0], t1 = gp[1], t2 = gp[2], t3 = gp[3];
                    float2 v;
                    v.x = t0.x * wx[k][0];
                    v.y = t0.y * wx[k][0];
                    v.x = fmaf(t1.x, wx[k][1], v.x);
                    v.y = fmaf(t1.y, wx[k][1], v.y);
                    v.x = fmaf(t2.x, wx[k][2], v.x);
                    v.y = fmaf(t2.y, wx[k][2], v.y);
                    v.x = fmaf(t3.x, wx[k][3], v.x);
                    v.y = fmaf(t3.y, wx[k][3], v.y);
                    s_rx[(kk * TH + r) * 64 + lane] = v;
                }
            }
            __syncthreads();
            // (4) column pass: the group's samples of this wave's rows, added to the pixel's running ring sum
            for (int jj = wv; jj < TJ; jj += kSrrfWaves) {
                const int j = by * TJ + jj;
                if (j >= a.HM) break;
                const bool in = col_in && j >= a.j_lo && j < a.j_hi;
                float div = grp == 0 ? 0.f : s_div[jj * 64 + lane];
                if (in) {
                    const float* re = s_row + jj * kSE;
                    // the group's 4 bases, 16 weights and 4 offsets of this row: 6 aligned 16-byte broadcast reads
                    const float4 rb4 = *reinterpret_cast<const float4*>(re + SrrfEntry::kBase + grp * kSrrfGroup);
                    const float4 dy4 = *reinterpret_cast<const float4*>(re + SrrfEntry::kD + grp * kSrrfGroup);
                    const float rbs[4] = {rb4.x, rb4.y, rb4.z, rb4.w}, dys[4] = {dy4.x, dy4.y, dy4.z, dy4.w};
#pragma unroll
                    for (int kk = 0; kk < kSrrfGroup; ++kk) {
                        const int k = grp * kSrrfGroup + kk;
                        const int rb = __float_as_int(rbs[kk]);
                        const float4 wy = *reinterpret_cast<const float4*>(re + SrrfEntry::kW + 4 * k);
                        const float dy = dys[kk];
                        const float2* rp = s_rx + (kk * TH + rb) * 64 + lane;
                        const float2 t0 = rp[0], t1 = rp[64], t2 = rp[128], t3 = rp[192];
                        float qx = t0.x * wy.x;
                        float qy = t0.y * wy.x;
                        qx = fmaf(t1.x, wy.y, qx);
                        qy = fmaf(t1.y, wy.y, qy);
                        qx = fmaf(t2.x, wy.z, qx);
                        qy = fmaf(t2.y, wy.z, qy);
                        qx = fmaf(t3.x, wy.w, qx);
                        qy = fmaf(t3.y, wy.w, qy);
                        // Dk = (1 - |G x (xc - x0, yc - y0)| / (|G| R))^2, 0 where |G| = 0 (_c_sr_radiality.c:6-13);
                        // dx, dy are (xc - x0) / R and (yc - y0) / R, a unit vector, so 1 - |cr| / |G| lies in [0, 1]:
                        // the clamp to [0, 1] (the FMA's clamp bit) changes nothing there, and where |G| = 0
                        // (0 * inf = NaN) it gives the reference's 0. The sample counts -Dk where G points along the
                        // ring vector (x0 - xc, y0 - yc), i.e. where G . (dx, dy) < 0 (the reference tests G . ring
                        // with the unrounded ring coordinates; the two differ only where G . ring ~ 0, where Dk ~ 0):
                        // dk times dk with the dot product's sign copied onto it.
                        const float g2 = fmaf(qx, qx, qy * qy);
                        const float cr = fabsf(fmaf(qy, dx[k], -qx * dy));
                        const float dk = __builtin_amdgcn_fmed3f(fmaf(-cr, __builtin_amdgcn_rsqf(g2), 1.f), 0.f, 1.f);
                        const float dot = fmaf(qx, dx[k], qy * dy);
                        div = fmaf(dk, copysignf(dk, dot), div);
                    }
                }
                if (grp + 1 < 12 / kSrrfGroup) {
                    s_div[jj * 64 + lane] = div;
                    continue;
                }
                // (5) last group: ring mean, positivity, intensity weight, store
                float v = 0.f;
                if (in) {
                    div = div / 12.f;
                    v = a.positivity ? fmaxf(div, 0.f) : div;
                    if (a.iw) {
                        float I;
                        if (INTERP) {
                            I = a.interp[static_cast<size_t>(f) * planeM + static_cast<size_t>(j) * a.WM + i];
                        } else {
                            // xM Catmull-Rom of the input at (j / M, i / M)
                            const float* re = s_row + jj * kSE;
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

// Plan with the tallest block (64, 32 or 16 output rows) whose LDS fits two blocks per CU, else one. Returns
// NPX_OK, NPX_ERR_ARG (bad parameters) or NPX_ERR_UNSUPPORTED (the ring does not fit a tile in LDS).
int build_srrf_plan_fit(int H, int W, int M, double ringRadius, int border, SrrfPlan& p, std::string* err) {
    for (size_t limit : {static_cast<size_t>(80 * 1024), static_cast<size_t>(160 * 1024)}) {
        for (int TJ : {64, 32, 16}) {
            if (!build_srrf_plan(H, W, M, ringRadius, border, TJ, kSrrfTI, p, err)) return NPX_ERR_ARG;
            if (srrf_lds_bytes(p.TJ, p.TH, p.TW) <= limit) return NPX_OK;
        }
    }
    if (err) *err = "ring radius too large for the radiality tile (LDS)";
    return NPX_ERR_UNSUPPORTED;
}

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
        const int rc = build_srrf_plan_fit(H, W, M, ringRadius, border, p->host, &err);
        if (rc != NPX_OK) return fail(rc, err);
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

namespace host {
void release_srrf() {
    std::lock_guard<std::mutex> lk(s_mu);
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) cur = -1;
    for (auto& kv : s_plans) {
        for (auto& d : kv.second->dev)
            if (hipSetDevice(d.first) == hipSuccess) (void)hipFree(d.second.mem);
        kv.second->dev.clear();
    }
    if (cur >= 0) (void)hipSetDevice(cur);
}
}  // namespace host
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
    a.TJ = h->TJ;
    a.TH = h->TH;
    a.TW = h->TW;
    a.j_lo = h->j_lo;
    a.j_hi = h->j_hi;
    a.i_lo = h->i_lo;
    a.i_hi = h->i_hi;
    a.positivity = positivity ? 1 : 0;
    a.iw = iw ? 1 : 0;
    const int ntx = static_cast<int>(h->tile_c0.size()), nty = static_cast<int>(h->tile_r0.size());
    const int tiles = ntx * nty;
    int chunks = std::max(1, std::min(N, (g_srrf_target_blocks + tiles - 1) / tiles));
    a.fpc = (N + chunks - 1) / chunks;
    chunks = (N + a.fpc - 1) / a.fpc;
    const size_t lds = srrf_lds_bytes(h->TJ, h->TH, h->TW);
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
    const int rc = build_srrf_plan_fit(H, W, M, ringRadius, border, p, &err);
    if (rc != NPX_OK) return host::fail(rc, err);
    const int v[12] = {p.HM, p.WM, p.j_lo, p.j_hi, p.i_lo, p.i_hi, p.TH, p.TW, p.TJ, kSrrfTI,
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
    const int rc = build_srrf_plan_fit(H, W, M, ringRadius, border, p, &err);
    if (rc != NPX_OK) return host::fail(rc, err);
    std::memcpy(row, p.row.data(), p.row.size() * sizeof(float));
    std::memcpy(col, p.col.data(), p.col.size() * sizeof(float));
    std::memcpy(tile_r0, p.tile_r0.data(), p.tile_r0.size() * sizeof(int));
    std::memcpy(tile_c0, p.tile_c0.data(), p.tile_c0.size() * sizeof(int));
    return NPX_OK;
}

}  // extern "C"
