// Fourier (zero-padded spectrum) interpolation of image stacks on gfx950 with hipFFT, as eSRRF's use_fht branch
// calls it (fht_space_interpolation, src/nanopyx/core/transform/interpolation.py:58-292), and the Roberts cross
// gradients of a stack (GradientRobertsCross, src/nanopyx/core/transform/_le_roberts_cross_gradients.pyx:67-83;
// per pixel src/include/_c_gradients.c:53-72), with their C ABI entry points.
//
// One interpolation = for each tile of the frame (the whole frame when neither side exceeds the tile size) and each
// batch of frames: the tile mirror-padded to a power-of-two square in float64 (fht_pad), a batched double-precision
// real-to-complex 2D FFT, then the reference's irfft2 of the M-times larger half spectrum whose row quadrants sit at
// its top and bottom. Only the first Ww/2+1 columns of that spectrum are non-zero, so the inverse runs as: those
// columns, stored column-major with the quadrants placed (fht_embed_cols), through a batched complex inverse FFT along
// the rows; a transpose into the enlarged half spectrum, zero beyond those columns, with the imaginary parts of the
// DC and Nyquist columns set to 0 as irfft ignores them (fht_to_half); a batched complex-to-real inverse FFT along
// the columns. Then per-frame min/max of the float32 result (fht_minmax), and one pass that rescales to the tile's
// input range, crops the padding and writes (or, for tiles, adds the Hann-weighted tile into the output and its
// window into the weight; fht_finish); tiled outputs are then divided by the summed window (fht_normalise).
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/npx_esrrf.h"
#include "npx_host.h"

namespace npx {

// ------------------------------------------------------------------------------------------ Roberts cross

__global__ __launch_bounds__(256) void roberts_cross(const void* in, int u16, int N, int H, int W, float* gc,
                                                     float* gr) {
    const size_t n = static_cast<size_t>(N) * H * W, plane = static_cast<size_t>(H) * W;
    for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < n;
         t += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const size_t f = t / plane, p = t % plane;
        const int r1 = static_cast<int>(p / W), c1 = static_cast<int>(p % W);
        const int r0 = r1 > 0 ? r1 - 1 : 0, c0 = c1 > 0 ? c1 - 1 : 0;
        auto ld = [&](int r, int c) {
            const size_t o = f * plane + static_cast<size_t>(r) * W + c;
            return u16 ? static_cast<float>(static_cast<const uint16_t*>(in)[o]) : static_cast<const float*>(in)[o];
        };
        const float a01 = ld(r0, c1), a10 = ld(r1, c0), a00 = ld(r0, c0), a11 = ld(r1, c1);
        // _c_gradients.c:53-72, the same order of additions
        gc[t] = a01 - a10 + a11 - a00;
        gr[t] = -a01 + a10 + a11 - a00;
    }
}

// ------------------------------------------------------------------------------------------ Fourier interpolation

struct FhtTile {
    const float* in;     // (N, h, w) stack
    float* out;          // (N, h*M, w*M)
    float* weight;       // (N, h*M, w*M) summed window (tiled), else unused
    double* pad;         // (B, Hw, Ww)
    double2* spec;       // (B, Hw, Ww/2+1)
    double2* cols;       // (B, Ww/2+1, HwM): the non-zero columns of the enlarged spectrum, column-major
    double2* half;       // (B, HwM, WwM/2+1): enlarged half spectrum after the row pass
    double* real;        // (B, HwM, WwM): the inverse
    float* mm;           // (B, 4): input min, input max, output min, output max (as ordered ints for atomics)
    const float* win_r;  // [th*M] window rows (tiled)
    const float* win_c;  // [tw*M]
    int N, h, w, M, f0, B;
    int r0, c0, th, tw;  // tile origin and size in input pixels
    int mirror, tiled;
    int Hw, Ww;          // padded working size
    int yroi, xroi;      // crop offsets of the tile inside the enlarged grid (output pixels)
    int tiles_h, tiles_w, mid_h, mid_w, off_h, off_w;  // mirror tiling: copies per axis, middle copy, crop offsets
};

__device__ inline int fmod2(int v) { return ((v % 2) + 2) % 2; }

// ordered-int encoding of floats for atomic min/max
__device__ inline int f2o(float v) {
    const int i = __float_as_int(v);
    return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ inline float o2f(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7fffffff); }

__global__ __launch_bounds__(256) void fht_init_mm(FhtTile a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < a.B) {
        int* m = reinterpret_cast<int*>(a.mm) + 4 * t;
        m[0] = f2o(INFINITY);
        m[1] = f2o(-INFINITY);
        m[2] = f2o(INFINITY);
        m[3] = f2o(-INFINITY);
    }
}

// mirror padding (interpolation.py:22-55) into float64, and the tile's input min / max
__global__ __launch_bounds__(256) void fht_pad(FhtTile a) {
    const size_t plane = static_cast<size_t>(a.Hw) * a.Ww, n = plane * a.B;
    const int fb = blockIdx.y;  // one frame per grid row keeps the min/max reduction per frame
    float lo = INFINITY, hi = -INFINITY;
    for (size_t p = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; p < plane;
         p += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const int y = static_cast<int>(p / a.Ww), x = static_cast<int>(p % a.Ww);
        int sy = y, sx = x;
        if (a.mirror) {
            const int Y = y + a.off_h, X = x + a.off_w;  // position in the tiles_h x tiles_w mirror grid
            const int ty = Y / a.th, tx = X / a.tw, qy = Y % a.th, qx = X % a.tw;
            sy = fmod2(ty - a.mid_h) == 0 ? qy : a.th - 1 - qy;
            sx = fmod2(tx - a.mid_w) == 0 ? qx : a.tw - 1 - qx;
        }
        const float v = a.in[static_cast<size_t>(a.f0 + fb) * a.h * a.w + static_cast<size_t>(a.r0 + sy) * a.w +
                             (a.c0 + sx)];
        a.pad[fb * plane + p] = static_cast<double>(v);
        lo = fminf(lo, v);
        hi = fmaxf(hi, v);
    }
    (void)n;
    // block reduction then one atomic per block
    __shared__ float s_lo[256], s_hi[256];
    s_lo[threadIdx.x] = lo;
    s_hi[threadIdx.x] = hi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            s_lo[threadIdx.x] = fminf(s_lo[threadIdx.x], s_lo[threadIdx.x + s]);
            s_hi[threadIdx.x] = fmaxf(s_hi[threadIdx.x], s_hi[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int* m = reinterpret_cast<int*>(a.mm) + 4 * fb;
        atomicMin(m + 0, f2o(s_lo[0]));
        atomicMax(m + 1, f2o(s_hi[0]));
    }
}

// the Ww/2+1 non-zero columns of the enlarged spectrum, column-major: rows [0, h2) from the top quadrant, rows
// [HwM - (Hw - h2), HwM) from the bottom one, zero between (interpolation.py:86-89)
__global__ __launch_bounds__(256) void fht_embed_cols(FhtTile a) {
    const int HM = a.Hw * a.M, hc = a.Ww / 2 + 1, h2 = (a.Hw + 1) / 2;
    const size_t plane = static_cast<size_t>(hc) * HM, n = plane * a.B;
    for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < n;
         t += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const size_t fb = t / plane, p = t % plane;
        const int c = static_cast<int>(p / HM), R = static_cast<int>(p % HM);
        int r = -1;
        if (R < h2) r = R;
        else if (R >= HM - (a.Hw - h2)) r = a.Hw - (HM - R);
        a.cols[t] = r >= 0 ? a.spec[fb * static_cast<size_t>(a.Hw) * hc + static_cast<size_t>(r) * hc + c]
                           : make_double2(0.0, 0.0);
    }
}

// transposed into the enlarged half spectrum (HwM x WwM/2+1): columns beyond Ww/2 zero, imaginary parts of the DC
// and (even WwM) Nyquist columns zero
__global__ __launch_bounds__(256) void fht_to_half(FhtTile a) {
    const int HM = a.Hw * a.M, WM = a.Ww * a.M, hc = a.Ww / 2 + 1, wh = WM / 2 + 1;
    const size_t plane = static_cast<size_t>(HM) * wh, n = plane * a.B;
    for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < n;
         t += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const size_t fb = t / plane, p = t % plane;
        const int y = static_cast<int>(p / wh), c = static_cast<int>(p % wh);
        double2 v = make_double2(0.0, 0.0);
        if (c < hc) v = a.cols[fb * static_cast<size_t>(hc) * HM + static_cast<size_t>(c) * HM + y];
        if (c == 0 || (WM % 2 == 0 && c == WM / 2)) v.y = 0.0;
        a.half[t] = v;
    }
}

// float32 result of the inverse (real part / (HwM WwM)) and its per-frame min / max
__global__ __launch_bounds__(256) void fht_minmax(FhtTile a) {
    const int HM = a.Hw * a.M, WM = a.Ww * a.M;
    const size_t plane = static_cast<size_t>(HM) * WM;
    const int fb = blockIdx.y;
    const double norm = 1.0 / (static_cast<double>(HM) * WM);
    float lo = INFINITY, hi = -INFINITY;
    for (size_t p = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; p < plane;
         p += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const float v = static_cast<float>(a.real[fb * plane + p] * norm);
        lo = fminf(lo, v);
        hi = fmaxf(hi, v);
    }
    __shared__ float s_lo[256], s_hi[256];
    s_lo[threadIdx.x] = lo;
    s_hi[threadIdx.x] = hi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            s_lo[threadIdx.x] = fminf(s_lo[threadIdx.x], s_lo[threadIdx.x + s]);
            s_hi[threadIdx.x] = fmaxf(s_hi[threadIdx.x], s_hi[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int* m = reinterpret_cast<int*>(a.mm) + 4 * fb;
        atomicMin(m + 2, f2o(s_lo[0]));
        atomicMax(m + 3, f2o(s_hi[0]));
    }
}

// rescale to the input range (float32 arithmetic, interpolation.py:90-97), crop, write or blend
__global__ __launch_bounds__(256) void fht_finish(FhtTile a) {
    const int HM = a.Hw * a.M, WM = a.Ww * a.M;
    const int oh = a.th * a.M, ow = a.tw * a.M;
    const size_t plane = static_cast<size_t>(HM) * WM, tplane = static_cast<size_t>(oh) * ow, n = tplane * a.B;
    const double norm = 1.0 / (static_cast<double>(HM) * WM);
    const size_t oplane = static_cast<size_t>(a.h) * a.M * a.w * a.M;
    for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < n;
         t += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const size_t fb = t / tplane, p = t % tplane;
        const int y = static_cast<int>(p / ow), x = static_cast<int>(p % ow);
        const int* m = reinterpret_cast<const int*>(a.mm) + 4 * fb;
        const float lo = o2f(m[0]), hi = o2f(m[1]), omin = o2f(m[2]), omax = o2f(m[3]);
        float v = static_cast<float>(a.real[fb * plane + static_cast<size_t>(y + a.yroi) * WM + (x + a.xroi)] * norm);
        if (omax != omin) {
            const float scale = static_cast<float>((static_cast<double>(hi) - lo) / (static_cast<double>(omax) - omin));
            v = v - omin;
            v = v * scale;
            v = v + lo;
        }
        const size_t o = static_cast<size_t>(a.f0 + fb) * oplane +
                         static_cast<size_t>(a.r0 * a.M + y) * (a.w * a.M) + (a.c0 * a.M + x);
        if (a.tiled) {
            const float wv = a.win_r[y] * a.win_c[x];
            a.out[o] += v * wv;
            a.weight[o] += wv;
        } else {
            a.out[o] = v;
        }
    }
}

__global__ __launch_bounds__(256) void fht_normalise(float* out, const float* weight, size_t n) {
    for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < n;
         t += static_cast<size_t>(gridDim.x) * blockDim.x)
        if (weight[t] > 0.f) out[t] /= weight[t];
}

__global__ __launch_bounds__(256) void fill_zero(float* p, size_t n) {
    for (size_t t = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; t < n;
         t += static_cast<size_t>(gridDim.x) * blockDim.x)
        p[t] = 0.f;
}

namespace {

using npx::host::fail;
using npx::host::hip_fail;

int grid_of(size_t n) { return static_cast<int>(std::min<size_t>((n + 255) / 256, size_t(65536))); }

// Padded working size of a tile and its mirror-grid parameters (interpolation.py:22-55).
struct Work {
    int Hw, Ww, tiles_h, tiles_w, mid_h, mid_w, off_h, off_w;
};
Work work_of(int th, int tw, bool mirror) {
    Work k{};
    if (!mirror) {
        k.Hw = th;
        k.Ww = tw;
        return k;
    }
    auto p2 = [](int v) { return v > 1 ? 1 << static_cast<int>(std::ceil(std::log2(static_cast<double>(v)))) : 1; };
    const int n = std::max(p2(th), p2(tw));
    int sw = static_cast<int>(std::ceil(static_cast<double>(n) / tw));
    if (sw % 2 == 0) ++sw;
    int sh = static_cast<int>(std::ceil(static_cast<double>(n) / th));
    if (sh % 2 == 0) ++sh;
    k.Hw = k.Ww = n;
    k.tiles_h = sh;
    k.tiles_w = sw;
    k.mid_h = (sh - 1) / 2;
    k.mid_w = (sw - 1) / 2;
    k.off_h = (sh * th - n) / 2;
    k.off_w = (sw * tw - n) / 2;
    return k;
}

std::vector<int> starts_of(int dim, int tile, int overlap) {
    const int step = std::max(1, tile - 2 * overlap);
    std::vector<int> s;
    for (int v = 0; v < dim; v += step) s.push_back(v);
    if (!s.empty() && s.back() + tile < dim) s.push_back(std::max(0, dim - tile));
    return s;
}

struct Plans {
    hipfftHandle fwd = 0, cols = 0, rows = 0;
};
std::mutex f_mu;
std::map<std::tuple<int, int, int, int, int>, Plans> f_plans;  // (device, Hw, Ww, M, batch)

int plans_for(int Hw, int Ww, int M, int B, Plans* out) {
    std::lock_guard<std::mutex> lk(f_mu);
    int d = 0;
    hipError_t e = hipGetDevice(&d);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    const auto key = std::make_tuple(d, Hw, Ww, M, B);
    auto it = f_plans.find(key);
    if (it == f_plans.end()) {
        Plans p;
        int nf[2] = {Hw, Ww}, nr[1] = {Hw * M}, nc[1] = {Ww * M};
        const int hc = Ww / 2 + 1;
        if (hipfftPlanMany(&p.fwd, 2, nf, nullptr, 1, 0, nullptr, 1, 0, HIPFFT_D2Z, B) != HIPFFT_SUCCESS)
            return fail(NPX_ERR_HIP, "hipfftPlanMany (forward)");
        // inverse along the rows of every non-zero column (contiguous columns of length HwM)
        if (hipfftPlanMany(&p.cols, 1, nr, nullptr, 1, 0, nullptr, 1, 0, HIPFFT_Z2Z, B * hc) != HIPFFT_SUCCESS) {
            hipfftDestroy(p.fwd);
            return fail(NPX_ERR_HIP, "hipfftPlanMany (columns)");
        }
        // complex-to-real inverse along the columns of every enlarged row
        if (hipfftPlanMany(&p.rows, 1, nc, nullptr, 1, 0, nullptr, 1, 0, HIPFFT_Z2D, B * Hw * M) != HIPFFT_SUCCESS) {
            hipfftDestroy(p.fwd);
            hipfftDestroy(p.cols);
            return fail(NPX_ERR_HIP, "hipfftPlanMany (rows)");
        }
        it = f_plans.emplace(key, p).first;
    }
    *out = it->second;
    return NPX_OK;
}

constexpr size_t kFhtBatchBytes = size_t(1) << 31;  // the enlarged complex grids of one batch of frames

// Batch size and per-part bytes of the workspace for tiles of at most th x tw input pixels.
struct Layout {
    int B;
    size_t pad, spec, cols, half, real, mm, win, weight, total;
};
Layout layout_of(int N, int h, int w, int M, int th, int tw, bool mirror, bool tiled) {
    const Work k = work_of(th, tw, mirror);
    const size_t HM = static_cast<size_t>(k.Hw) * M, WM = static_cast<size_t>(k.Ww) * M;
    const size_t per = HM * (WM / 2 + 1) * sizeof(double2) + HM * WM * sizeof(double);
    Layout L{};
    L.B = static_cast<int>(std::max<size_t>(
        1, std::min<size_t>({static_cast<size_t>(N), kFhtBatchBytes / per, size_t(4096)})));
    auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
    L.pad = al(static_cast<size_t>(L.B) * k.Hw * k.Ww * sizeof(double));
    L.spec = al(static_cast<size_t>(L.B) * k.Hw * (k.Ww / 2 + 1) * sizeof(double2));
    L.cols = al(static_cast<size_t>(L.B) * (k.Ww / 2 + 1) * HM * sizeof(double2));
    L.half = al(static_cast<size_t>(L.B) * HM * (WM / 2 + 1) * sizeof(double2));
    L.real = al(static_cast<size_t>(L.B) * HM * WM * sizeof(double));
    L.mm = al(static_cast<size_t>(L.B) * 4 * sizeof(float));
    L.win = tiled ? al(static_cast<size_t>(th + tw) * M * sizeof(float)) : 0;
    L.weight = tiled ? al(static_cast<size_t>(N) * h * M * w * M * sizeof(float)) : 0;
    L.total = L.pad + L.spec + L.cols + L.half + L.real + L.mm + L.win + L.weight;
    return L;
}

bool tiled_of(int h, int w, int tile) { return h > tile || w > tile; }

}  // namespace

namespace host {
void release_fht() {
    std::lock_guard<std::mutex> lk(f_mu);
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) cur = -1;
    for (auto& kv : f_plans) {
        if (hipSetDevice(std::get<0>(kv.first)) != hipSuccess) continue;
        hipfftDestroy(kv.second.fwd);
        hipfftDestroy(kv.second.cols);
        hipfftDestroy(kv.second.rows);
    }
    f_plans.clear();
    if (cur >= 0) (void)hipSetDevice(cur);
}
}  // namespace host
}  // namespace npx

extern "C" {

int npx_roberts(const void* d_img, int dtype, int N, int H, int W, float* d_gc, float* d_gr, void* stream) {
    using namespace npx;
    if (!d_img || !d_gc || !d_gr) return host::fail(NPX_ERR_ARG, "NULL array pointer");
    if (dtype != NPX_DTYPE_F32 && dtype != NPX_DTYPE_U16) return host::fail(NPX_ERR_ARG, "dtype must be f32 or u16");
    if (N < 1 || H < 1 || W < 1) return host::fail(NPX_ERR_ARG, "bad shape");
    const size_t n = static_cast<size_t>(N) * H * W;
    hipLaunchKernelGGL(roberts_cross, dim3(grid_of(n)), dim3(256), 0, static_cast<hipStream_t>(stream), d_img,
                       dtype == NPX_DTYPE_U16 ? 1 : 0, N, H, W, d_gc, d_gr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return host::hip_fail(e, "roberts launch");
    return NPX_OK;
}

size_t npx_fht_workspace(int N, int h, int w, int M, int mirror, int tile_size, int overlap) {
    using namespace npx;
    if (N < 1 || h < 1 || w < 1 || M < 1) return 0;
    const int tile = tile_size > 0 ? tile_size : 256;
    const bool tiled = tiled_of(h, w, tile);
    (void)overlap;
    const int th = tiled ? std::min(tile, h) : h, tw = tiled ? std::min(tile, w) : w;
    // tiles are at most tile x tile; the padded size of a smaller tile is never larger
    return layout_of(N, h, w, M, th, tw, tiled ? true : mirror != 0, tiled).total;
}

int npx_fht_interpolate(const float* d_in, int N, int h, int w, int M, int mirror, int tile_size, int overlap,
                        float* d_out, void* d_work, size_t work_bytes, void* stream) {
    using namespace npx;
    if (!d_in || !d_out || !d_work) return host::fail(NPX_ERR_ARG, "NULL array pointer");
    if (N < 1 || h < 1 || w < 1 || M < 1) return host::fail(NPX_ERR_ARG, "bad shape");
    const int tile = tile_size > 0 ? tile_size : 256;
    const bool tiled = tiled_of(h, w, tile);
    const size_t need = npx_fht_workspace(N, h, w, M, mirror, tile_size, overlap);
    if (work_bytes < need) return host::fail(NPX_ERR_WORKSPACE, "workspace smaller than npx_fht_workspace()");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int th_max = tiled ? std::min(tile, h) : h, tw_max = tiled ? std::min(tile, w) : w;
    const Layout L = layout_of(N, h, w, M, th_max, tw_max, tiled ? true : mirror != 0, tiled);
    auto* wb = static_cast<uint8_t*>(d_work);
    FhtTile a{};
    a.in = d_in;
    a.out = d_out;
    a.N = N;
    a.h = h;
    a.w = w;
    a.M = M;
    a.pad = reinterpret_cast<double*>(wb);
    a.spec = reinterpret_cast<double2*>(wb + L.pad);
    size_t off = L.pad + L.spec;
    a.cols = reinterpret_cast<double2*>(wb + off);
    off += L.cols;
    a.half = reinterpret_cast<double2*>(wb + off);
    off += L.half;
    a.real = reinterpret_cast<double*>(wb + off);
    off += L.real;
    a.mm = reinterpret_cast<float*>(wb + off);
    off += L.mm;
    float* win = tiled ? reinterpret_cast<float*>(wb + off) : nullptr;
    off += L.win;
    a.weight = tiled ? reinterpret_cast<float*>(wb + off) : nullptr;
    a.tiled = tiled ? 1 : 0;
    a.mirror = tiled ? 1 : (mirror ? 1 : 0);
    const size_t nout = static_cast<size_t>(N) * h * M * w * M;
    if (tiled) {
        hipLaunchKernelGGL(fill_zero, dim3(grid_of(nout)), dim3(256), 0, st, d_out, nout);
        hipLaunchKernelGGL(fill_zero, dim3(grid_of(nout)), dim3(256), 0, st, a.weight, nout);
    }
    const std::vector<int> rs = tiled ? starts_of(h, tile, overlap) : std::vector<int>{0};
    const std::vector<int> cs = tiled ? starts_of(w, tile, overlap) : std::vector<int>{0};
    std::vector<float> hwin;
    for (int r0 : rs) {
        const int r1 = tiled ? std::min(r0 + tile, h) : h;
        for (int c0 : cs) {
            const int c1 = tiled ? std::min(c0 + tile, w) : w;
            a.r0 = r0;
            a.c0 = c0;
            a.th = r1 - r0;
            a.tw = c1 - c0;
            const Work k = work_of(a.th, a.tw, a.mirror != 0);
            a.Hw = k.Hw;
            a.Ww = k.Ww;
            a.tiles_h = k.tiles_h;
            a.tiles_w = k.tiles_w;
            a.mid_h = k.mid_h;
            a.mid_w = k.mid_w;
            a.off_h = k.off_h;
            a.off_w = k.off_w;
            a.yroi = a.mirror ? (k.Hw * M - M * a.th) / 2 : 0;  // interpolation.py:104-107
            a.xroi = a.mirror ? (k.Ww * M - M * a.tw) / 2 : 0;
            if (tiled) {
                // np.hanning(n + 2)[1:-1] in float32, rows then columns (interpolation.py:170-173); NumPy forms
                // the window as 0.5 + 0.5 cos(pi k / (L - 1)) over k = 1 - L, 3 - L, ..., L - 1 with L = n + 2
                const int oh = a.th * M, ow = a.tw * M;
                hwin.assign(oh + ow, 0.f);
                auto hann = [](int i, int n) {
                    const double k = 2.0 * (i + 1) + 1.0 - (n + 2);
                    return static_cast<float>(0.5 + 0.5 * std::cos(M_PI * k / (n + 1)));
                };
                for (int i = 0; i < oh; ++i) hwin[i] = hann(i, oh);
                for (int i = 0; i < ow; ++i) hwin[oh + i] = hann(i, ow);
                hipError_t e = hipMemcpyAsync(win, hwin.data(), hwin.size() * sizeof(float), hipMemcpyHostToDevice,
                                              st);
                if (e == hipSuccess) e = hipStreamSynchronize(st);  // hwin is reused by the next tile
                if (e != hipSuccess) return host::hip_fail(e, "window upload");
                a.win_r = win;
                a.win_c = win + oh;
            }
            for (int f0 = 0; f0 < N; f0 += L.B) {
                a.f0 = f0;
                a.B = std::min(L.B, N - f0);
                Plans pl;
                int rc = plans_for(a.Hw, a.Ww, M, a.B, &pl);
                if (rc != NPX_OK) return rc;
                const size_t plane = static_cast<size_t>(a.Hw) * a.Ww;
                const size_t bigp = static_cast<size_t>(a.Hw) * M * a.Ww * M;
                hipLaunchKernelGGL(fht_init_mm, dim3((a.B + 255) / 256), dim3(256), 0, st, a);
                hipLaunchKernelGGL(fht_pad, dim3(std::min<size_t>((plane + 255) / 256, 1024), a.B), dim3(256), 0, st,
                                   a);
                if (hipfftSetStream(pl.fwd, st) != HIPFFT_SUCCESS ||
                    hipfftExecD2Z(pl.fwd, reinterpret_cast<hipfftDoubleReal*>(a.pad),
                                  reinterpret_cast<hipfftDoubleComplex*>(a.spec)) != HIPFFT_SUCCESS)
                    return host::fail(NPX_ERR_HIP, "hipfftExecD2Z");
                const size_t ncol = static_cast<size_t>(a.Ww / 2 + 1) * a.Hw * M * a.B;
                const size_t nhalf = static_cast<size_t>(a.Hw) * M * (a.Ww * M / 2 + 1) * a.B;
                hipLaunchKernelGGL(fht_embed_cols, dim3(grid_of(ncol)), dim3(256), 0, st, a);
                if (hipfftSetStream(pl.cols, st) != HIPFFT_SUCCESS ||
                    hipfftExecZ2Z(pl.cols, reinterpret_cast<hipfftDoubleComplex*>(a.cols),
                                  reinterpret_cast<hipfftDoubleComplex*>(a.cols), HIPFFT_BACKWARD) != HIPFFT_SUCCESS)
                    return host::fail(NPX_ERR_HIP, "hipfftExecZ2Z (columns)");
                hipLaunchKernelGGL(fht_to_half, dim3(grid_of(nhalf)), dim3(256), 0, st, a);
                if (hipfftSetStream(pl.rows, st) != HIPFFT_SUCCESS ||
                    hipfftExecZ2D(pl.rows, reinterpret_cast<hipfftDoubleComplex*>(a.half),
                                  reinterpret_cast<hipfftDoubleReal*>(a.real)) != HIPFFT_SUCCESS)
                    return host::fail(NPX_ERR_HIP, "hipfftExecZ2D (rows)");
                hipLaunchKernelGGL(fht_minmax, dim3(std::min<size_t>((bigp + 255) / 256, 1024), a.B), dim3(256), 0,
                                   st, a);
                const size_t nt = static_cast<size_t>(a.th) * M * a.tw * M * a.B;
                hipLaunchKernelGGL(fht_finish, dim3(grid_of(nt)), dim3(256), 0, st, a);
                hipError_t e = hipGetLastError();
                if (e != hipSuccess) return host::hip_fail(e, "Fourier interpolation launch");
            }
        }
    }
    if (tiled) hipLaunchKernelGGL(fht_normalise, dim3(grid_of(nout)), dim3(256), 0, st, d_out, a.weight, nout);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return host::hip_fail(e, "Fourier interpolation launch");
    return NPX_OK;
}

}  // extern "C"
