// eSRRF 3D on gfx950 (eSRRF3D, src/nanopyx/core/transform/_le_esrrf3d.pyx:73-175; per-voxel kernel
// src/include/_c_sr_radial_gradient_convergence.c:92-204; gradients src/include/_c_gradients.c:75-116;
// interpolation src/nanopyx/core/transform/_interpolation.pyx:54-103), and its C ABI entry points.
//
// Per chunk of frames: (1) grad3d: x and y central differences / 2 and the z backward difference / 2 of every
// voxel (neighbours clamped); (2) magnify_xy: Catmull-Rom xMxy of the intensity and the three gradients of every
// slice (identity at Mxy = 1), stored interleaved as float4 (Gx, Gy, Gz, I) per magnified pixel of each input slice;
// the reference's z "linear" step repeats each slice Mz times (its slice position is an integer), so no z copy is
// made: magnified slice zM reads input slice zM / Mz; (3) rgc3d: one thread per output voxel walks the chunk's
// frames, sums the window samples (one float4 gather each), applies the sensitivity power and the intensity weight,
// and folds the value into the running mean (or Welford M2 and mean) kept in registers, reading and writing the
// caller's state once per chunk.
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

struct E3Args {
    const void* in;       // (F, S, R, C) float32 or uint16
    float4* mag;          // (chunk, S, RM, CM) float4: Gx, Gy, Gz, intensity
    float* grad;          // (chunk, 3, S, R, C)
    float* out;           // (SM, RM, CM): running mean (average) or M2 (std)
    float* mean;          // (SM, RM, CM) running mean of the std mode
    const int* crb_r;     // [RM] first tap row,    crw_r [RM] 4 weights
    const float4* crw_r;
    const int* crb_c;     // [CM]
    const float4* crw_c;
    int u16, F, S, R, C, Mxy, Mz, SM, RM, CM;
    int f0, nf;           // first frame of the chunk, frames in the chunk
    int margin, margin_z;
    int j0, j1, k0, k1;   // window: x/y offsets [j0, j1], z offsets [k0, k1]
    float tSO, tSO_z, tSS, tSS_z, S_pow;
    int pow_on, mode, iw, last;  // last: this chunk ends the stack (std: take sqrt(M2 / F))
};

__device__ inline float e3_load(const E3Args& a, size_t o) {
    return a.u16 ? static_cast<float>(static_cast<const uint16_t*>(a.in)[o]) : static_cast<const float*>(a.in)[o];
}

// (1) gradients of every voxel of the chunk's frames
__global__ __launch_bounds__(256) void e3_grad3d(E3Args a) {
    const size_t vol = static_cast<size_t>(a.S) * a.R * a.C;
    const size_t n = vol * a.nf;
    for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < n;
         t += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const size_t fl = t / vol, v = t % vol;
        const int x = static_cast<int>(v % a.C), y = static_cast<int>((v / a.C) % a.R),
                  z = static_cast<int>(v / (static_cast<size_t>(a.R) * a.C));
        const size_t base = (a.f0 + fl) * vol, plane = static_cast<size_t>(a.R) * a.C;
        const int zm = z <= 0 ? 0 : z - 1;
        const int yp = y >= a.R - 1 ? a.R - 1 : y + 1, ym = y <= 0 ? 0 : y - 1;
        const int xp = x >= a.C - 1 ? a.C - 1 : x + 1, xm = x <= 0 ? 0 : x - 1;
        const size_t row = base + z * plane + static_cast<size_t>(y) * a.C;
        const float c = e3_load(a, row + x);
        const float gx = (e3_load(a, row + xp) - e3_load(a, row + xm)) / 2;
        const float gy = (e3_load(a, base + z * plane + static_cast<size_t>(yp) * a.C + x) -
                          e3_load(a, base + z * plane + static_cast<size_t>(ym) * a.C + x)) / 2;
        const float gz = (c - e3_load(a, base + zm * plane + static_cast<size_t>(y) * a.C + x)) / 2;
        float* g = a.grad + fl * 3 * vol;
        g[v] = gx;
        g[vol + v] = gy;
        g[2 * vol + v] = gz;
    }
}

// (2) Catmull-Rom xMxy of (Gx, Gy, Gz, intensity) of every slice: columns outer, rows inner, as _c_interpolate.
// Blocks of 64 x 4 output pixels of one (frame, slice) plane: x covers CM, y covers RM, z the planes (strided).
constexpr int kE3MagRows = 4;
__global__ __launch_bounds__(64 * kE3MagRows) void e3_magnify_xy(E3Args a) {
    const size_t vol = static_cast<size_t>(a.S) * a.R * a.C, plane = static_cast<size_t>(a.R) * a.C;
    const size_t planeM = static_cast<size_t>(a.RM) * a.CM;
    const int cM = blockIdx.x * 64 + threadIdx.x, rM = blockIdx.y * kE3MagRows + threadIdx.y;
    if (cM >= a.CM || rM >= a.RM) return;
    for (int zs = blockIdx.z; zs < a.S * a.nf; zs += gridDim.z) {
        const int fl = zs / a.S, s = zs - fl * a.S;
        const float* g = a.grad + static_cast<size_t>(fl) * 3 * vol + s * plane;
        const size_t ibase = (static_cast<size_t>(a.f0) + fl) * vol + s * plane;
        float4 r;
        if (a.Mxy == 1) {
            const size_t o = static_cast<size_t>(rM) * a.C + cM;
            r = make_float4(g[o], g[vol + o], g[2 * vol + o], e3_load(a, ibase + o));
        } else {
            const int rb = a.crb_r[rM], cb = a.crb_c[cM];
            const float4 wr = a.crw_r[rM], wc = a.crw_c[cM];
            const float wrs[4] = {wr.x, wr.y, wr.z, wr.w}, wcs[4] = {wc.x, wc.y, wc.z, wc.w};
            float q0 = 0.f, q1 = 0.f, q2 = 0.f, q3 = 0.f;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int cn = cb + u;
                if (wcs[u] == 0.f || cn < 0 || cn >= a.C) continue;
                float p0 = 0.f, p1 = 0.f, p2 = 0.f, p3 = 0.f;
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    const int rn = rb + w;
                    if (wrs[w] == 0.f || rn < 0 || rn >= a.R) continue;
                    const size_t o = static_cast<size_t>(rn) * a.C + cn;
                    p0 = fmaf(g[o], wrs[w], p0);
                    p1 = fmaf(g[vol + o], wrs[w], p1);
                    p2 = fmaf(g[2 * vol + o], wrs[w], p2);
                    p3 = fmaf(e3_load(a, ibase + o), wrs[w], p3);
                }
                q0 = fmaf(p0, wcs[u], q0);
                q1 = fmaf(p1, wcs[u], q1);
                q2 = fmaf(p2, wcs[u], q2);
                q3 = fmaf(p3, wcs[u], q3);
            }
            r = make_float4(q0, q1, q2, q3);
        }
        // the convergence uses only G's direction (the sign of G . d and |G x d| / |G|): store G / |G| (0 for G = 0)
        const float n2 = fmaf(r.x, r.x, fmaf(r.y, r.y, r.z * r.z));
        const float inv = n2 > 0.f ? rsqrtf(n2) : 0.f;
        r.x *= inv;
        r.y *= inv;
        r.z *= inv;
        a.mag[static_cast<size_t>(zs) * planeM + static_cast<size_t>(rM) * a.CM + cM] = r;
    }
}

// (3) RGC of every output voxel over the chunk's frames, folded into the running statistics. The window geometry
// (validity, offsets, distance, weight and the weight sum) does not depend on the frame, so it is formed once per
// batch of kE3Batch frames and the batch's ring sums are kept in registers.
constexpr int kE3Batch = 16;
constexpr float kE3Scale = 65536.f;  // s of the indicator clamp(-s g . d, 0, 1) (a power of two)

// Voxels of a block: 64 consecutive columns (the lanes) x kE3Waves rows of one phase (stride M), one row per wave, so
// the waves' window samples come from largely the same magnified rows and meet in L1. A thread takes the ZB output
// slices of one source slice (ZB = Mz when Mz is 2 or 4, else 1): with Mz a power of two their centres zc = sM / Mz
// and sample positions vz = zc + k are exact, so every sub-slice sees the same offsets (dz = k), distances, weights
// and source records and they differ only in which z offsets fall inside the volume. The thread gathers each record
// once and adds its term to the sums of the sub-slices it is valid for: the sums are those of separate threads, bit
// for bit.
constexpr int kE3Waves = 4;
__host__ __device__ inline int e3_zb(int Mz) { return (Mz == 2 || Mz == 4) ? Mz : 1; }

// Window sums of the ZB sub-slices of one output column and row for NB frames of a batch (rg[r * NBM + q], NBM =
// kE3Batch / ZB frames at most) and their weight sums. The NB gathers of a window sample are issued together, before
// any of them is used.
template <int ZB, int NB>
__device__ __forceinline__ void e3_window(const E3Args& a, const float4* m, size_t volM, size_t planeM, float xc,
                                          float yc, const float (&zc)[ZB], float ylim, float xlim, float zlim,
                                          float (&rg)[kE3Batch], float (&wsum)[ZB]) {
    // no contraction of a * b + c into FMAs beyond the explicit fmaf calls: every batch size (a separate inlined
    // copy of this body) then rounds alike, so results do not depend on how the frames are batched or chunked
#pragma clang fp contract(off)
    constexpr int NBM = kE3Batch / ZB;
    static_assert(NB <= NBM, "batch larger than the sums kept per sub-slice");
    for (int j = a.j0; j <= a.j1; ++j) {
        const float vy = yc + j;
        if (!(0.f < vy && vy <= ylim)) continue;
        const float dy = vy - yc;
        const size_t oy = static_cast<size_t>(static_cast<int>(a.Mxy * vy)) * a.CM;
        for (int i = a.j0; i <= a.j1; ++i) {
            const float vx = xc + i;
            if (!(0.f < vx && vx <= xlim)) continue;
            const float dx = vx - xc;
            const float dxy2 = dx * dx + dy * dy;
            if (!(sqrtf(dxy2) <= a.tSO)) continue;
            const size_t oyx = oy + static_cast<int>(a.Mxy * vx);
            const float wxy = __expf(-dxy2 / a.tSS);
            for (int k = a.k0; k <= a.k1; ++k) {
                bool ok[ZB];
                int first = -1;
#pragma unroll
                for (int r = 0; r < ZB; ++r) {
                    const float vz = zc[r] + k;
                    ok[r] = 0.f < vz && vz <= zlim;
                    if (ok[r] && first < 0) first = r;
                }
                if (first < 0) continue;
                float zf = zc[0];
#pragma unroll
                for (int r = 1; r < ZB; ++r) zf = first == r ? zc[r] : zf;
                const float vz = zf + k;
                const float dz = vz - zf;  // the same for every sub-slice
                const float d2 = dxy2 + dz * dz;
                if (d2 == 0.f || !(dz <= a.tSO_z)) continue;
                const float4* g = m + static_cast<size_t>(static_cast<int>(vz * a.Mz) / a.Mz) * planeM +
                                  oyx;  // source slice of the repeated volume (the same for every sub-slice)
                float3 G[NB];  // g = G / |G| of the batch's frames
#pragma unroll
                for (int q = 0; q < NB; ++q) {
                    const float4 rr = g[q * volM];
                    G[q] = make_float3(rr.x, rr.y, rr.z);
                }
                const float dist = sqrtf(d2);
                float w = dist * (wxy * __expf(-dz * dz / a.tSS_z));
                w = (w * w) * (w * w);
#pragma unroll
                for (int r = 0; r < ZB; ++r) wsum[r] += ok[r] ? w : 0.f;
                const float wd = w * __builtin_amdgcn_rcpf(dist);  // Dk w = w - |g x d| * w / |d|
                // [g . d < 0] as clamp(-s g . d, 0, 1) on the clamp bit of the last FMA: exact for |g . d| >= 1/s;
                // below that the term it scales is under w (g . d)^2 / |d|^2
                const float sdx = -kE3Scale * dx, sdy = -kE3Scale * dy, sdz = -kE3Scale * dz;
#pragma unroll
                for (int q = 0; q < NB; ++q) {
                    const float3 gq = G[q];
                    const float neg =
                        __builtin_amdgcn_fmed3f(fmaf(gq.x, sdx, fmaf(gq.y, sdy, gq.z * sdz)), 0.f, 1.f);
                    const float c1 = fmaf(gq.y, dz, -gq.z * dy), c2 = fmaf(gq.z, dx, -gq.x * dz),
                                c3 = fmaf(gq.x, dy, -gq.y * dx);
                    const float cr2 = fmaf(c1, c1, fmaf(c2, c2, c3 * c3));
                    const float term = fmaf(-__builtin_amdgcn_sqrtf(cr2), wd, w);
#pragma unroll
                    for (int r = 0; r < ZB; ++r) rg[r * NBM + q] = fmaf(ok[r] ? neg : 0.f, term, rg[r * NBM + q]);
                }
            }
        }
    }
}

template <int ZB>
__global__ __launch_bounds__(64 * kE3Waves) void e3_rgc3d(E3Args a) {
#pragma clang fp contract(off)
    constexpr int NBM = kE3Batch / ZB;
    const size_t volM = static_cast<size_t>(a.S) * a.RM * a.CM, planeM = static_cast<size_t>(a.RM) * a.CM;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int ncx = (a.CM + 63) / 64, nG = (a.R + kE3Waves - 1) / kE3Waves, nz = (a.SM + ZB - 1) / ZB;
    const size_t ntiles = static_cast<size_t>(ncx) * a.Mxy * nG * nz;
    for (size_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        size_t rest = tile;
        const int cx = static_cast<int>(rest % ncx);
        rest /= ncx;
        const int p = static_cast<int>(rest % a.Mxy);
        rest /= a.Mxy;
        const int g = static_cast<int>(rest % nG);
        const int sz = static_cast<int>(rest / nG);
        const int cM = cx * 64 + lane, rM = p + a.Mxy * (kE3Waves * g + wv), sM0 = sz * ZB;
        if (cM >= a.CM || rM >= a.RM) continue;
        const bool in_xy = rM >= a.margin && rM < a.RM - a.margin && cM >= a.margin && cM < a.CM - a.margin;
        bool in[ZB];
        float acc[ZB], mu[ZB], zc[ZB];
        bool any = false;
#pragma unroll
        for (int r = 0; r < ZB; ++r) {
            const int sM = sM0 + r;
            const size_t t = static_cast<size_t>(sM) * planeM + static_cast<size_t>(rM) * a.CM + cM;
            in[r] = sM < a.SM && in_xy && sM >= a.margin_z && sM < a.SM - a.margin_z;
            any = any || in[r];
            acc[r] = (a.f0 == 0 || sM >= a.SM) ? 0.f : a.out[t];
            mu[r] = (a.mode == 1 && a.f0 > 0 && sM < a.SM) ? a.mean[t] : 0.f;
            zc[r] = static_cast<float>(sM) / a.Mz;  // window coordinates exactly as the reference forms them
        }
        if (any) {
            const float xc = static_cast<float>(cM) / a.Mxy, yc = static_cast<float>(rM) / a.Mxy;
            const float ylim = static_cast<float>(a.RM / a.Mxy - 1), xlim = static_cast<float>(a.CM / a.Mxy - 1);
            const float zlim = static_cast<float>(a.SM / a.Mz - 1);
            // intensity of the source slice, the same for every sub-slice
            const size_t self = static_cast<size_t>(sM0 / a.Mz) * planeM + static_cast<size_t>(rM) * a.CM + cM;
            // batches of equal size (at most NBM frames)
            const int nbat = (a.nf + NBM - 1) / NBM, bsz = (a.nf + nbat - 1) / nbat;
            for (int fb = 0; fb < a.nf; fb += bsz) {
                const int nb = min(bsz, a.nf - fb);
                const float4* m = a.mag + fb * volM;
                float rg[kE3Batch];
#pragma unroll
                for (int q = 0; q < kE3Batch; ++q) rg[q] = 0.f;
                float wsum[ZB];
#pragma unroll
                for (int r = 0; r < ZB; ++r) wsum[r] = 0.f;
                switch (nb) {
#define NPX_E3_CASE(n) \
    case n:            \
        if constexpr (n <= NBM) e3_window<ZB, (n <= NBM ? n : 1)>(a, m, volM, planeM, xc, yc, zc, ylim, xlim, zlim, rg, wsum); \
        break;
                    NPX_E3_CASE(1) NPX_E3_CASE(2) NPX_E3_CASE(3) NPX_E3_CASE(4)
                    NPX_E3_CASE(5) NPX_E3_CASE(6) NPX_E3_CASE(7) NPX_E3_CASE(8)
                    NPX_E3_CASE(9) NPX_E3_CASE(10) NPX_E3_CASE(11) NPX_E3_CASE(12)
                    NPX_E3_CASE(13) NPX_E3_CASE(14) NPX_E3_CASE(15) NPX_E3_CASE(16)
#undef NPX_E3_CASE
                    default: break;
                }
#pragma unroll
                for (int r = 0; r < ZB; ++r) {
                    if (!in[r]) continue;
#pragma unroll
                    for (int q = 0; q < NBM; ++q) {
                        if (q >= nb) break;
                        float rgc = rg[r * NBM + q] / wsum[r];
                        if (rgc >= 0.f && a.pow_on) rgc = powf(rgc, a.S_pow);
                        else if (rgc < 0.f) rgc = 0.f;
                        if (a.iw) rgc *= m[q * volM + self].w;
                        const float fn = static_cast<float>(a.f0 + fb + q + 1);
                        if (a.mode == 0) {
                            acc[r] = acc[r] + (rgc - acc[r]) / fn;
                        } else {
                            const float delta = rgc - mu[r];
                            mu[r] += delta / fn;
                            acc[r] += delta * (rgc - mu[r]);
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < ZB; ++r) {
            const int sM = sM0 + r;
            if (sM >= a.SM) continue;
            const size_t t = static_cast<size_t>(sM) * planeM + static_cast<size_t>(rM) * a.CM + cM;
            float v = acc[r];
            if (a.mode == 1 && !a.last) a.mean[t] = mu[r];
            if (a.mode == 1 && a.last) v = sqrtf(v / static_cast<float>(a.F));
            a.out[t] = v;
        }
    }
}

// ------------------------------------------------------------------------------------------ host side

namespace {

using npx::host::fail;
using npx::host::hip_fail;

struct E3Dev {
    void* mem = nullptr;
    const int *crb_r = nullptr, *crb_c = nullptr;
    const float4 *crw_r = nullptr, *crw_c = nullptr;
};

std::mutex e3_mu;
std::map<std::tuple<int, int, int, int>, E3Dev> e3_tables;  // (device, R, C, Mxy)

int e3_tables_for(int R, int C, int Mxy, hipStream_t st, const E3Dev** out) {
    std::lock_guard<std::mutex> lk(e3_mu);
    int d = 0;
    hipError_t e = hipGetDevice(&d);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    const auto key = std::make_tuple(d, R, C, Mxy);
    auto it = e3_tables.find(key);
    if (it == e3_tables.end()) {
        if (npx::host::capturing(st)) return fail(NPX_ERR_ARG, "first use of an eSRRF 3D plan under capture");
        const int RM = R * Mxy, CM = C * Mxy;
        std::vector<int> br(RM), bc(CM);
        std::vector<float4> wr(RM), wc(CM);
        const float mag = static_cast<float>(Mxy);
        for (int r = 0; r < RM; ++r) {
            const CrTap t = cr_taps(r, mag, R);
            br[r] = t.base;
            wr[r] = make_float4(t.w[0], t.w[1], t.w[2], t.w[3]);
        }
        for (int c = 0; c < CM; ++c) {
            const CrTap t = cr_taps(c, mag, C);
            bc[c] = t.base;
            wc[c] = make_float4(t.w[0], t.w[1], t.w[2], t.w[3]);
        }
        const size_t bwr = RM * sizeof(float4), bwc = CM * sizeof(float4), bbr = RM * sizeof(int);
        const size_t bbc = CM * sizeof(int);
        std::vector<uint8_t> blob(bwr + bwc + bbr + bbc);
        std::memcpy(blob.data(), wr.data(), bwr);
        std::memcpy(blob.data() + bwr, wc.data(), bwc);
        std::memcpy(blob.data() + bwr + bwc, br.data(), bbr);
        std::memcpy(blob.data() + bwr + bwc + bbr, bc.data(), bbc);
        void* mem = nullptr;
        e = hipMalloc(&mem, blob.size());
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(eSRRF 3D tables)");
        e = hipMemcpy(mem, blob.data(), blob.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e != hipSuccess) {
            (void)hipFree(mem);
            return hip_fail(e, "upload of eSRRF 3D tables");
        }
        auto* b = static_cast<uint8_t*>(mem);
        E3Dev t;
        t.mem = mem;
        t.crw_r = reinterpret_cast<const float4*>(b);
        t.crw_c = reinterpret_cast<const float4*>(b + bwr);
        t.crb_r = reinterpret_cast<const int*>(b + bwr + bwc);
        t.crb_c = reinterpret_cast<const int*>(b + bwr + bwc + bbr);
        it = e3_tables.emplace(key, t).first;
    }
    *out = &it->second;
    return NPX_OK;
}

size_t g_e3_chunk_bytes = size_t(1) << 30;  // workspace per chunk of frames (magnified volumes + gradients)

int e3_chunk(int F, int S, int R, int C, int Mxy) {
    const size_t per = static_cast<size_t>(S) * R * Mxy * C * Mxy * sizeof(float4) +
                       static_cast<size_t>(3) * S * R * C * sizeof(float);
    return static_cast<int>(std::max<size_t>(1, std::min<size_t>(static_cast<size_t>(F), g_e3_chunk_bytes / per)));
}

size_t e3_tiles(const E3Args& a) {
    const int Zb = e3_zb(a.Mz);
    return static_cast<size_t>((a.CM + 63) / 64) * a.Mxy * ((a.R + kE3Waves - 1) / kE3Waves) * ((a.SM + Zb - 1) / Zb);
}

int e3_grid(size_t n) { return static_cast<int>(std::min<size_t>((n + 255) / 256, size_t(65536))); }

}  // namespace

namespace host {
void release_esrrf3d() {
    std::lock_guard<std::mutex> lk(e3_mu);
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) cur = -1;
    for (auto& kv : e3_tables)
        if (hipSetDevice(std::get<0>(kv.first)) == hipSuccess) (void)hipFree(kv.second.mem);
    e3_tables.clear();
    if (cur >= 0) (void)hipSetDevice(cur);
}
}  // namespace host
}  // namespace npx

extern "C" {

size_t npx_esrrf3d_chunk_bytes(size_t bytes) {
    const size_t prev = npx::g_e3_chunk_bytes;
    if (bytes) npx::g_e3_chunk_bytes = bytes;
    return prev;
}

size_t npx_esrrf3d_workspace(int F, int S, int R, int C, int Mxy, int Mz, int mode) {
    if (F < 1 || S < 1 || R < 1 || C < 1 || Mxy < 1 || Mz < 1) return 0;
    const int ch = npx::e3_chunk(F, S, R, C, Mxy);
    const size_t mag = static_cast<size_t>(ch) * S * R * Mxy * C * Mxy * sizeof(float4);
    const size_t grad = static_cast<size_t>(ch) * 3 * S * R * C * sizeof(float);
    const size_t mean = mode == 1 ? static_cast<size_t>(S) * Mz * R * Mxy * C * Mxy * sizeof(float) : 0;
    return ((mag + 255) & ~size_t(255)) + ((grad + 255) & ~size_t(255)) + mean;
}

int npx_esrrf3d(const void* d_img, int dtype, int F, int S, int R, int C, int Mxy, int Mz, double radius,
                double radius_z, double sensitivity, int mode, int iw, float* d_out, void* d_work, size_t work_bytes,
                void* stream) {
    using namespace npx;
    if (!d_img || !d_out || !d_work) return host::fail(NPX_ERR_ARG, "NULL array pointer");
    if (dtype != NPX_DTYPE_F32 && dtype != NPX_DTYPE_U16) return host::fail(NPX_ERR_ARG, "dtype must be f32 or u16");
    if (F < 1 || S < 1 || R < 1 || C < 1 || Mxy < 1 || Mz < 1) return host::fail(NPX_ERR_ARG, "bad shape");
    if (mode != 0 && mode != 1) return host::fail(NPX_ERR_ARG, "mode must be average (0) or std (1)");
    if (!(radius >= 0) || !(radius_z >= 0)) return host::fail(NPX_ERR_ARG, "radii must be >= 0");
    const size_t need = npx_esrrf3d_workspace(F, S, R, C, Mxy, Mz, mode);
    if (work_bytes < need) return host::fail(NPX_ERR_WORKSPACE, "workspace smaller than npx_esrrf3d_workspace()");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const E3Dev* tb = nullptr;
    int rc = e3_tables_for(R, C, Mxy, st, &tb);
    if (rc != NPX_OK) return rc;
    // parameters as src/nanopyx/core/transform/_le_esrrf3d.pyx:80-92 derives them
    const float sigma = static_cast<float>(radius / 2.355);
    const float sigma_z = static_cast<float>(radius_z / 2.355);
    const float fwhm = static_cast<float>(radius), fwhm_z = static_cast<float>(radius_z);
    E3Args a{};
    a.in = d_img;
    a.u16 = dtype == NPX_DTYPE_U16;
    a.F = F;
    a.S = S;
    a.R = R;
    a.C = C;
    a.Mxy = Mxy;
    a.Mz = Mz;
    a.SM = S * Mz;
    a.RM = R * Mxy;
    a.CM = C * Mxy;
    a.margin = static_cast<int>(radius * Mxy);
    a.margin_z = static_cast<int>(radius_z * Mz);
    a.tSS = 2 * sigma * sigma;
    a.tSO = 2 * sigma + 1;
    a.tSS_z = 2 * sigma_z * sigma_z;
    a.tSO_z = 2 * sigma_z + 1;
    a.j0 = -static_cast<int>(fwhm);
    a.j1 = static_cast<int>(fwhm + 1);
    a.k0 = -static_cast<int>(fwhm_z);
    a.k1 = static_cast<int>(fwhm_z + 1);
    a.S_pow = static_cast<float>(sensitivity);
    a.pow_on = a.S_pow > 1.0f;
    a.mode = mode;
    a.iw = iw ? 1 : 0;
    a.out = d_out;
    a.crb_r = tb->crb_r;
    a.crw_r = tb->crw_r;
    a.crb_c = tb->crb_c;
    a.crw_c = tb->crw_c;
    const int ch = e3_chunk(F, S, R, C, Mxy);
    const size_t mag = static_cast<size_t>(ch) * S * a.RM * a.CM * sizeof(float4);
    const size_t grad = static_cast<size_t>(ch) * 3 * S * R * C * sizeof(float);
    auto* w = static_cast<uint8_t*>(d_work);
    a.mag = reinterpret_cast<float4*>(w);
    a.grad = reinterpret_cast<float*>(w + ((mag + 255) & ~size_t(255)));
    a.mean = mode == 1 ? reinterpret_cast<float*>(w + ((mag + 255) & ~size_t(255)) + ((grad + 255) & ~size_t(255)))
                       : nullptr;
    for (int f0 = 0; f0 < F; f0 += ch) {
        a.f0 = f0;
        a.nf = std::min(ch, F - f0);
        a.last = f0 + a.nf >= F;
        const size_t ng = static_cast<size_t>(a.nf) * S * R * C;
        hipLaunchKernelGGL(e3_grad3d, dim3(e3_grid(ng)), dim3(256), 0, st, a);
        hipLaunchKernelGGL(e3_magnify_xy,
                           dim3((a.CM + 63) / 64, (a.RM + kE3MagRows - 1) / kE3MagRows, std::min(S * a.nf, 65535)),
                           dim3(64, kE3MagRows), 0, st, a);
        const dim3 grid(static_cast<int>(std::min<size_t>(e3_tiles(a), size_t(1) << 20))), block(64 * kE3Waves);
        switch (e3_zb(Mz)) {
            case 2: hipLaunchKernelGGL(e3_rgc3d<2>, grid, block, 0, st, a); break;
            case 4: hipLaunchKernelGGL(e3_rgc3d<4>, grid, block, 0, st, a); break;
            default: hipLaunchKernelGGL(e3_rgc3d<1>, grid, block, 0, st, a); break;
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return host::hip_fail(e, "eSRRF 3D launch");
    }
    return NPX_OK;
}

}  // extern "C"
