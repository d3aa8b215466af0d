// Catmull-Rom shift and magnify of image stacks on gfx950 (ShiftAndMagnify / CRShiftAndMagnify,
// src/nanopyx/core/transform/_le_interpolation_catmull_rom.pyx:143-170; per pixel
// src/include/_c_interpolation_catmull_rom.c:21-52) and its C ABI entry point. The four taps and weights of an output
// row depend on the row only, and those of a column on the column only, so the host tabulates them (float32
// coordinates, float64 cubic weights, 0 for taps outside the image and for coordinates outside it); the kernel sums
// the 16 weighted taps, columns outer and rows inner as the reference does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/npx_esrrf.h"
#include "npx_host.h"
#include "npx_plan.h"

namespace npx {

__global__ __launch_bounds__(256) void shift_magnify(const void* in, int u16, int N, int H, int W, int HM, int WM,
                                                     const int* rb, const float4* rw, const int* cb,
                                                     const float4* cw, float* out) {
    const size_t plane = static_cast<size_t>(HM) * WM, n = plane * N, iplane = static_cast<size_t>(H) * W;
    for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < n;
         t += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const size_t f = t / plane, p = t % plane;
        const int i = static_cast<int>(p / WM), j = static_cast<int>(p % WM);
        const int r0 = rb[i], c0 = cb[j];
        const float4 wr4 = rw[i], wc4 = cw[j];
        const float wr[4] = {wr4.x, wr4.y, wr4.z, wr4.w}, wc[4] = {wc4.x, wc4.y, wc4.z, wc4.w};
        float q = 0.f;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int cn = c0 + u;
            if (wc[u] == 0.f || cn < 0 || cn >= W) continue;
            float s = 0.f;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int rn = r0 + v;
                if (wr[v] == 0.f || rn < 0 || rn >= H) continue;
                const size_t o = f * iplane + static_cast<size_t>(rn) * W + cn;
                const float x = u16 ? static_cast<float>(static_cast<const uint16_t*>(in)[o])
                                    : static_cast<const float*>(in)[o];
                s = fmaf(x, wr[v], s);
            }
            q = fmaf(s, wc[u], q);
        }
        out[t] = q;
    }
}

namespace {

using npx::host::fail;
using npx::host::hip_fail;

struct SmTables {
    void* mem = nullptr;
    const int *rb = nullptr, *cb = nullptr;
    const float4 *rw = nullptr, *cw = nullptr;
};
std::mutex sm_mu;
std::map<std::tuple<int, int, int, uint32_t, uint32_t, uint32_t, uint32_t>, SmTables> sm_tables;

uint32_t fbits(float v) {
    uint32_t b;
    std::memcpy(&b, &v, sizeof b);
    return b;
}

int sm_tables_for(int H, int W, float sr, float sc, float mr, float mc, int HM, int WM, hipStream_t st,
                  const SmTables** out) {
    std::lock_guard<std::mutex> lk(sm_mu);
    int d = 0;
    hipError_t e = hipGetDevice(&d);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    const auto key = std::make_tuple(d, H, W, fbits(sr), fbits(sc), fbits(mr), fbits(mc));
    auto it = sm_tables.find(key);
    if (it == sm_tables.end()) {
        if (npx::host::capturing(st)) return fail(NPX_ERR_ARG, "first use of a shift-magnify plan under capture");
        std::vector<int> br(HM), bc(WM);
        std::vector<float4> wr(HM), wc(WM);
        for (int i = 0; i < HM; ++i) {
            const CrTap t = cr_taps_shift(i, mr, sr, H);
            br[i] = t.base;
            wr[i] = make_float4(t.w[0], t.w[1], t.w[2], t.w[3]);
        }
        for (int j = 0; j < WM; ++j) {
            const CrTap t = cr_taps_shift(j, mc, sc, W);
            bc[j] = t.base;
            wc[j] = make_float4(t.w[0], t.w[1], t.w[2], t.w[3]);
        }
        const size_t a = HM * sizeof(float4), b = WM * sizeof(float4), c = HM * sizeof(int), dd = WM * sizeof(int);
        std::vector<uint8_t> blob(a + b + c + dd);
        std::memcpy(blob.data(), wr.data(), a);
        std::memcpy(blob.data() + a, wc.data(), b);
        std::memcpy(blob.data() + a + b, br.data(), c);
        std::memcpy(blob.data() + a + b + c, bc.data(), dd);
        void* mem = nullptr;
        e = hipMalloc(&mem, blob.size());
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(shift-magnify tables)");
        e = hipMemcpy(mem, blob.data(), blob.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e != hipSuccess) {
            (void)hipFree(mem);
            return hip_fail(e, "upload of shift-magnify tables");
        }
        auto* m = static_cast<uint8_t*>(mem);
        SmTables tb;
        tb.mem = mem;
        tb.rw = reinterpret_cast<const float4*>(m);
        tb.cw = reinterpret_cast<const float4*>(m + a);
        tb.rb = reinterpret_cast<const int*>(m + a + b);
        tb.cb = reinterpret_cast<const int*>(m + a + b + c);
        it = sm_tables.emplace(key, tb).first;
    }
    *out = &it->second;
    return NPX_OK;
}

}  // namespace

namespace host {
void release_interp() {
    std::lock_guard<std::mutex> lk(sm_mu);
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) cur = -1;
    for (auto& kv : sm_tables)
        if (hipSetDevice(std::get<0>(kv.first)) == hipSuccess) (void)hipFree(kv.second.mem);
    sm_tables.clear();
    if (cur >= 0) (void)hipSetDevice(cur);
}
}  // namespace host
}  // namespace npx

extern "C" {

int npx_shift_magnify_shape(int H, int W, float mag_row, float mag_col, int* HM, int* WM) {
    if (!HM || !WM) return npx::host::fail(NPX_ERR_ARG, "NULL output");
    // <int>(rows * magnification_row) in float32 (_le_interpolation_catmull_rom.pyx:156-157)
    *HM = static_cast<int>(static_cast<float>(H) * mag_row);
    *WM = static_cast<int>(static_cast<float>(W) * mag_col);
    return NPX_OK;
}

int npx_shift_magnify(const void* d_img, int dtype, int N, int H, int W, float shift_row, float shift_col,
                      float mag_row, float mag_col, float* d_out, void* stream) {
    using namespace npx;
    if (!d_img || !d_out) return host::fail(NPX_ERR_ARG, "NULL array pointer");
    if (dtype != NPX_DTYPE_F32 && dtype != NPX_DTYPE_U16) return host::fail(NPX_ERR_ARG, "dtype must be f32 or u16");
    if (N < 1 || H < 1 || W < 1 || !(mag_row > 0) || !(mag_col > 0)) return host::fail(NPX_ERR_ARG, "bad shape");
    int HM = 0, WM = 0;
    npx_shift_magnify_shape(H, W, mag_row, mag_col, &HM, &WM);
    if (HM < 1 || WM < 1) return host::fail(NPX_ERR_ARG, "empty output");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const SmTables* tb = nullptr;
    int rc = sm_tables_for(H, W, shift_row, shift_col, mag_row, mag_col, HM, WM, st, &tb);
    if (rc != NPX_OK) return rc;
    const size_t n = static_cast<size_t>(N) * HM * WM;
    const int grid = static_cast<int>(std::min<size_t>((n + 255) / 256, size_t(65536)));
    hipLaunchKernelGGL(shift_magnify, dim3(grid), dim3(256), 0, st, d_img, dtype == NPX_DTYPE_U16 ? 1 : 0, N, H, W,
                       HM, WM, tb->rb, tb->rw, tb->cb, tb->cw, d_out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return host::hip_fail(e, "shift-magnify launch");
    return NPX_OK;
}

}  // extern "C"
