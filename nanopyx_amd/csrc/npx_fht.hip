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
// the rows (one column of the spectrum each); then along the columns, for a power-of-two Ww >= 2 and M >= 2, one
// kernel (fht_rows) evaluates each output row as M polyphase inverse DFTs of length Ww, in pairs as complex FFTs in
// LDS, and writes the float32 result with its per-frame min / max; for other shapes a transpose into the enlarged
// half spectrum, zero beyond those columns, with the imaginary parts of the DC and Nyquist columns set to 0 as irfft
// ignores them (fht_to_half), a batched complex-to-real inverse FFT and a min / max pass (fht_minmax). Then one pass
// rescales to the tile's input range, crops the padding and writes (or, for tiles, adds the Hann-weighted tile into
// the output and its window into the weight; fht_finish); tiled outputs are then divided by the summed window
// (fht_normalise).
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
    float* realf;        // (B, HwM, WwM) float32 of the inverse / (HwM WwM), in place of `real` (fht_rows), or null
    float* mm;           // (B, 4): input min, input max, output min, output max (as ordered ints for atomics)
    const float* win_r;  // [th*M] window rows (tiled)
    const float* win_c;  // [tw*M]
    int N, h, w, M, f0, B;
    int r0, c0, th, tw;  // tile origin and size in input pixels
    int mirror, tiled;
    int Hw, Ww;          // padded working size
    int yroi, xroi;      // crop offsets of the tile inside the enlarged grid (output pixels)
    int tiles_h, tiles_w, mid_h, mid_w, off_h, off_w;  // mirror tiling: copies per axis, middle copy, crop offsets
    int lat;             // >= 2: write only output rows / columns that are multiples of lat and the last ones, compact
};

// Output position g of an axis of n positions is written: all of them, or (lattice output) the multiples of lat and
// the last position, which are what a radial-gradient-convergence pass reads from a x2M gradient (its sample
// positions are k M, clamped to the last row / column).
__device__ inline bool on_lattice(int g, int lat, int n) { return lat < 2 || g % lat == 0 || g == n - 1; }

// Kept positions of a span [g0, g0 + len) of an axis of n positions: all len of them (lat < 2), or the multiples of
// lat in it and the last position of the axis when it falls inside; kept_index gives the t-th one's offset in the span.
__host__ __device__ inline int kept_count(int g0, int len, int n, int lat) {
    if (lat < 2) return len;
    const int k0 = (g0 + lat - 1) / lat, nk = (g0 + len - 1) / lat - k0 + 1;
    return nk + (n - 1 >= g0 && n - 1 < g0 + len && (n - 1) % lat != 0 ? 1 : 0);
}
__device__ inline int kept_index(int g0, int len, int n, int lat, int t) {
    const int k0 = (g0 + lat - 1) / lat, nk = (g0 + len - 1) / lat - k0 + 1;
    return (t < nk ? (k0 + t) * lat : n - 1) - g0;
}

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

// mirror padding (interpolation.py:22-55) into float64, and the tile's input min / max. A block covers kPadRows rows
// of one frame (grid: row groups x frames): each thread finds its columns' sources once and reuses them per row.
constexpr int kPadRows = 8;
__global__ __launch_bounds__(256) void fht_pad(FhtTile a) {
    const int fb = blockIdx.y, y0 = blockIdx.x * kPadRows, y1 = min(y0 + kPadRows, a.Hw);
    const float* in = a.in + static_cast<size_t>(a.f0 + fb) * a.h * a.w;
    double* out = a.pad + static_cast<size_t>(fb) * a.Hw * a.Ww;
    float lo = INFINITY, hi = -INFINITY;
    for (int x = threadIdx.x; x < a.Ww; x += blockDim.x) {
        int sx = x;
        if (a.mirror) {
            const int X = x + a.off_w, tx = X / a.tw, qx = X - tx * a.tw;  // position in the mirror grid
            sx = fmod2(tx - a.mid_w) == 0 ? qx : a.tw - 1 - qx;
        }
        for (int y = y0; y < y1; ++y) {
            int sy = y;
            if (a.mirror) {
                const int Y = y + a.off_h, ty = Y / a.th, qy = Y - ty * a.th;
                sy = fmod2(ty - a.mid_h) == 0 ? qy : a.th - 1 - qy;
            }
            const float v = in[static_cast<size_t>(a.r0 + sy) * a.w + (a.c0 + sx)];
            out[static_cast<size_t>(y) * a.Ww + x] = static_cast<double>(v);
            lo = fminf(lo, v);
            hi = fmaxf(hi, v);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, o));
        hi = fmaxf(hi, __shfl_xor(hi, o));
    }
    __shared__ float s_lo[4], s_hi[4];
    if (threadIdx.x % 64 == 0) {
        s_lo[threadIdx.x / 64] = lo;
        s_hi[threadIdx.x / 64] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < static_cast<int>(blockDim.x / 64); ++k) {
            lo = fminf(lo, s_lo[k]);
            hi = fmaxf(hi, s_hi[k]);
        }
        int* m = reinterpret_cast<int*>(a.mm) + 4 * fb;
        atomicMin(m + 0, f2o(lo));
        atomicMax(m + 1, f2o(hi));
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
// and (even WwM) Nyquist columns zero. A block covers kHalfRows rows of one frame (grid: frames x row groups): the
// column-major source is read kHalfRows contiguous entries per column into LDS, kHalfCols columns at a time, and
// each row is written as contiguous spans; past the non-zero columns only zeros are written.
constexpr int kHalfRows = 16, kHalfCols = 64;
__global__ __launch_bounds__(256) void fht_to_half(FhtTile a) {
    const int HM = a.Hw * a.M, WM = a.Ww * a.M, hc = a.Ww / 2 + 1, wh = WM / 2 + 1;
    const int ngrp = (HM + kHalfRows - 1) / kHalfRows;
    const int fb = blockIdx.x / ngrp, y0 = (blockIdx.x % ngrp) * kHalfRows, nr = min(kHalfRows, HM - y0);
    const double2* src = a.cols + static_cast<size_t>(fb) * hc * HM + y0;
    double2* dst = a.half + (static_cast<size_t>(fb) * HM + y0) * wh;
    __shared__ double2 s[kHalfCols][kHalfRows + 1];
    for (int c0 = 0; c0 < wh; c0 += kHalfCols) {
        const bool filled = c0 < hc;  // block-uniform
        if (filled) {
            __syncthreads();  // the previous span's reads of s are done
            for (int t = threadIdx.x; t < kHalfCols * kHalfRows; t += blockDim.x) {
                const int c = t / kHalfRows, r = t % kHalfRows;
                s[c][r] = (c0 + c < hc && r < nr) ? src[static_cast<size_t>(c0 + c) * HM + r] : make_double2(0.0, 0.0);
            }
            __syncthreads();
        }
        for (int t = threadIdx.x; t < kHalfRows * kHalfCols; t += blockDim.x) {
            const int r = t / kHalfCols, c = t % kHalfCols, C = c0 + c;
            if (r < nr && C < wh) {
                double2 v = filled ? s[c][r] : make_double2(0.0, 0.0);
                if (C == 0 || (WM % 2 == 0 && C == WM / 2)) v.y = 0.0;
                dst[static_cast<size_t>(r) * wh + C] = v;
            }
        }
    }
}

// Row inverse of the enlarged half spectrum without forming it (power-of-two Ww >= 2, M >= 2). Row y holds X[k] for
// k <= K = Ww/2 (the column pass's output) and zeros up to WwM/2, so irfft's unnormalised value at n = M m + r is
//   x[M m + r] = Re X[0] + 2 sum_{k=1..K} Re(X[k] e^{2 pi i k (M m + r) / (Ww M)}),
// a length-Ww inverse DFT over m of G_r: G_r[0] = Re X[0], G_r[k] = Y_r[k] and G_r[Ww-k] = conj(Y_r[k]) for 0 < k < K,
// G_r[K] = 2 Re Y_r[K], with Y_r[k] = X[k] e^{2 pi i k r / (Ww M)}. Phases r = 2p and 2p+1 share one complex FFT
// (G_2p + i G_2p+1, real outputs), so a row is ceil(M/2) complex FFTs of length Ww: mixed radix-4/2 Stockham in LDS
// (each stage reads its butterflies into registers, then writes them back in place). The result is written as the
// float32 of x / (HwM WwM) with its per-frame min / max, replacing the half spectrum, the complex-to-real FFT and a
// separate min / max pass (the float64 grid is never written). A block covers kRowsPerBlock rows of one frame, kRowPair
// rows at a time; twiddles e^{2 pi i t / Ww} and e^{2 pi i j / (Ww M)} (split as j = 64 h + l) sit in LDS.
constexpr int kRowsPerBlock = 8;
constexpr int kRowLdsPoints = 3072;  // double2 of FFT data per block (48 KiB); tables and X come on top
constexpr int kRowLdsLimit = 64 * 1024;
constexpr int kRowThreads = 256;
constexpr int kRowLdsLimit256 = 160 * 1024;  // LDS one workgroup may use (fht_rows256)

__device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

struct RowShape {
    int rb;        // rows per pass (1 or 2)
    int npt;       // FFT entries per pass: rb * ceil(M/2) FFTs of Ww points, fft_stride(Ww) apart
    int nhi;       // entries of the coarse phase-twiddle table
    size_t bytes;  // dynamic LDS
};

// LDS entries of one FFT of Ww points: entry e sits at e + e / 16 (a 16-byte skew every 16 entries spreads the
// strided stage writes over the banks), plus one entry that staggers consecutive FFTs.
__host__ __device__ inline int fft_stride(int Ww) { return Ww + (Ww >> 4) + 1; }
__device__ inline int sk(int e) { return e + (e >> 4); }

__host__ __device__ inline RowShape row_shape(int Ww, int M, int rb) {
    RowShape r{};
    const int K = Ww / 2, P = (M + 1) / 2;
    r.rb = rb;
    r.npt = rb * P * fft_stride(Ww);
    r.nhi = (K * (M - 1)) / 64 + 1;
    r.bytes = sizeof(double2) * (static_cast<size_t>(r.npt) + Ww + 64 + r.nhi + static_cast<size_t>(rb) * (K + 1));
    return r;
}

__global__ __launch_bounds__(kRowThreads) void fht_rows(FhtTile a, int rb) {
    extern __shared__ double2 lds[];
    const int HM = a.Hw * a.M, Ww = a.Ww, M = a.M, N = Ww * M, K = Ww / 2, hc = K + 1, P = (M + 1) / 2;
    const RowShape sh = row_shape(Ww, M, rb);
    const int FS = fft_stride(Ww);
    double2* F = lds;                  // [rb][P][fft_stride(Ww)], skewed entries
    double2* tw = F + sh.npt;          // the Stockham stages' twiddles, stage after stage (Ww - 1 of Ww entries)
    double2* tlo = tw + Ww;            // e^{2 pi i l / N}, l < 64
    double2* thi = tlo + 64;           // e^{2 pi i 64 h / N}, h < nhi
    double2* X = thi + sh.nhi;         // [rb][K + 1]
    const int ngrp = (HM + kRowsPerBlock - 1) / kRowsPerBlock;
    const int fb = blockIdx.x / ngrp, ybeg = (blockIdx.x % ngrp) * kRowsPerBlock;
    const int tid = threadIdx.x;
    for (int t = tid; t < Ww + 64 + sh.nhi; t += kRowThreads) {
        double s = 0.0, c = 1.0;
        if (t < Ww - 1) {  // stage tables: stage (Ns, R) holds e^{2 pi i r k / (Ns R)} at (r - 1) Ns + k
            int base = 0;
            for (int ns = 1; ns < Ww;) {
                const int R = ns * 4 <= Ww ? 4 : 2, cnt = (R - 1) * ns;
                if (t < base + cnt) {
                    const int l = t - base, r = l / ns + 1, k = l - (r - 1) * ns;
                    sincospi(2.0 * (r * k * (Ww / (ns * R))) / Ww, &s, &c);
                    break;
                }
                base += cnt;
                ns *= R;
            }
        } else if (t >= Ww && t < Ww + 64) {
            sincospi(2.0 * (t - Ww) / N, &s, &c);
        } else if (t >= Ww + 64) {
            sincospi(2.0 * 64.0 * (t - Ww - 64) / N, &s, &c);
        }
        tw[t] = make_double2(c, s);
    }
    const int lgW = __builtin_ctz(Ww);
    const double norm = 1.0 / (static_cast<double>(HM) * N);
    const double2* src = a.cols + static_cast<size_t>(fb) * hc * HM;
    float lo = INFINITY, hi = -INFINITY;
    const int nf = rb * P;
    for (int y0 = ybeg; y0 < ybeg + kRowsPerBlock; y0 += rb) {
        __syncthreads();  // tables written; the previous pass's reads of F and X are done
        for (int t = tid; t < rb * hc; t += kRowThreads) {
            const int rr = t % rb, c = t / rb;
            X[rr * hc + c] = y0 + rr < HM ? src[static_cast<size_t>(c) * HM + y0 + rr] : make_double2(0.0, 0.0);
        }
        __syncthreads();
        // Z = G_2p + i G_2p+1 for every (row, pair): one thread per kk < K writes Z[kk] and Z[Ww - kk], the first nf
        // threads of a second step write Z[K]
        const int lgK = lgW - 1;
        for (int t = tid; t < nf << lgK; t += kRowThreads) {
            const int kk = t & (K - 1), f = t >> lgK, rr = f / P, p = f - rr * P, r1 = 2 * p;
            const double2* x = X + rr * hc;
            double2* z = F + f * FS;
            const bool two = r1 + 1 < M;
            if (kk == 0) {
                z[0] = make_double2(x[0].x, two ? x[0].x : 0.0);
            } else {
                const int j1 = kk * r1, j2 = j1 + kk;
                const double2 y1 = cmul(x[kk], cmul(thi[j1 >> 6], tlo[j1 & 63]));
                const double2 y2 = two ? cmul(x[kk], cmul(thi[j2 >> 6], tlo[j2 & 63])) : make_double2(0.0, 0.0);
                z[sk(kk)] = make_double2(y1.x - y2.y, y1.y + y2.x);         // y1 + i y2
                z[sk(Ww - kk)] = make_double2(y1.x + y2.y, -y1.y + y2.x);   // conj(y1) + i conj(y2)
            }
        }
        for (int f = tid; f < nf; f += kRowThreads) {
            const int rr = f / P, p = f - rr * P, r1 = 2 * p, j1 = K * r1, j2 = j1 + K;
            const double2 xk = X[rr * hc + K];
            const double2 y1 = cmul(xk, cmul(thi[j1 >> 6], tlo[j1 & 63]));
            const double2 y2 = r1 + 1 < M ? cmul(xk, cmul(thi[j2 >> 6], tlo[j2 & 63])) : make_double2(0.0, 0.0);
            F[f * FS + sk(K)] = make_double2(2.0 * y1.x, 2.0 * y2.x);
        }
        // Stockham stages: Ns points already combined; radix 4 while it fits, then one radix-2 stage
        int Ns = 1, lgNs = 0;
        const double2* tws = tw;  // this stage's table: twiddle of (r, k) at (r - 1) Ns + k
        while (Ns < Ww) {
            __syncthreads();
            if (Ns * 4 <= Ww) {
                const int lgnb = lgW - 2, nb = 1 << lgnb, total = nf << lgnb;
                double2 v[3][4];
#pragma unroll
                for (int u = 0; u < 3; ++u) {
                    const int b = tid + u * kRowThreads;
                    if (b < total) {
                        const int f = b >> lgnb, j = b & (nb - 1), k = j & (Ns - 1);
                        const double2* in = F + f * FS;
                        const double2 v0 = in[sk(j)], v1 = cmul(in[sk(j + nb)], tws[k]),
                                      v2 = cmul(in[sk(j + 2 * nb)], tws[Ns + k]),
                                      v3 = cmul(in[sk(j + 3 * nb)], tws[2 * Ns + k]);
                        const double2 a0 = make_double2(v0.x + v2.x, v0.y + v2.y);
                        const double2 a1 = make_double2(v0.x - v2.x, v0.y - v2.y);
                        const double2 b0 = make_double2(v1.x + v3.x, v1.y + v3.y);
                        const double2 b1 = make_double2(-(v1.y - v3.y), v1.x - v3.x);  // i (v1 - v3)
                        v[u][0] = make_double2(a0.x + b0.x, a0.y + b0.y);
                        v[u][1] = make_double2(a1.x + b1.x, a1.y + b1.y);
                        v[u][2] = make_double2(a0.x - b0.x, a0.y - b0.y);
                        v[u][3] = make_double2(a1.x - b1.x, a1.y - b1.y);
                    }
                }
                __syncthreads();
#pragma unroll
                for (int u = 0; u < 3; ++u) {
                    const int b = tid + u * kRowThreads;
                    if (b < total) {
                        const int f = b >> lgnb, j = b & (nb - 1), k = j & (Ns - 1);
                        double2* out = F + f * FS;
                        const int o = ((j >> lgNs) << (lgNs + 2)) + k;
#pragma unroll
                        for (int r = 0; r < 4; ++r) out[sk(o + r * Ns)] = v[u][r];
                    }
                }
                tws += 3 * Ns;
                Ns *= 4;
                lgNs += 2;
            } else {
                const int lgnb = lgW - 1, nb = 1 << lgnb, total = nf << lgnb;
                double2 v[6][2];
#pragma unroll
                for (int u = 0; u < 6; ++u) {
                    const int b = tid + u * kRowThreads;
                    if (b < total) {
                        const int f = b >> lgnb, j = b & (nb - 1), k = j & (Ns - 1);
                        const double2* in = F + f * FS;
                        const double2 v0 = in[sk(j)], v1 = cmul(in[sk(j + nb)], tws[k]);
                        v[u][0] = make_double2(v0.x + v1.x, v0.y + v1.y);
                        v[u][1] = make_double2(v0.x - v1.x, v0.y - v1.y);
                    }
                }
                __syncthreads();
#pragma unroll
                for (int u = 0; u < 6; ++u) {
                    const int b = tid + u * kRowThreads;
                    if (b < total) {
                        const int f = b >> lgnb, j = b & (nb - 1), k = j & (Ns - 1);
                        double2* out = F + f * FS;
                        const int o = ((j >> lgNs) << (lgNs + 1)) + k;
                        out[sk(o)] = v[u][0];
                        out[sk(o + Ns)] = v[u][1];
                    }
                }
                tws += Ns;
                Ns *= 2;
                lgNs += 1;
            }
        }
        __syncthreads();
        // pair p at m holds phases 2p (real part) and 2p + 1 (imaginary part): outputs n = M m + 2p and n + 1.
        // m = t / P through a float reciprocal: t < Ww P <= 2^12, so the error stays far below 1 / (2P).
        const float invP = 1.f / P;
        // only positions inside the crop that fht_finish reads are stored; min / max covers the whole grid
        const int oh = a.th * M, ow = a.tw * M, gh = a.h * M, gw = a.w * M;
        for (int rr = 0; rr < rb; ++rr) {
            const int y = y0 + rr;
            if (y >= HM) break;
            const int yo = y - a.yroi;
            const bool row_kept = yo >= 0 && yo < oh && on_lattice(a.r0 * M + yo, a.lat, gh);
            float* dst = a.realf + (static_cast<size_t>(fb) * HM + y) * N;
            const double2* fr = F + rr * P * FS;
            const bool store_all = row_kept && a.lat < 2;
            for (int t = tid; t < Ww * P; t += kRowThreads) {
                const int m = static_cast<int>((t + 0.5f) * invP), p = t - m * P, n = M * m + 2 * p;
                const double2 z = fr[p * FS + sk(m)];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    if (h == 1 && 2 * p + 1 >= M) break;
                    const float v = static_cast<float>((h ? z.y : z.x) * norm);
                    lo = fminf(lo, v);
                    hi = fmaxf(hi, v);
                    const int xo = n + h - a.xroi;
                    if (store_all && xo >= 0 && xo < ow) dst[n + h] = v;
                }
            }
            if (row_kept && a.lat >= 2) {  // lattice output: only its columns inside the crop, same values as above
                const int nkc = kept_count(a.c0 * M, ow, gw, a.lat);
                for (int t = tid; t < nkc; t += kRowThreads) {
                    const int n = kept_index(a.c0 * M, ow, gw, a.lat, t) + a.xroi;
                    const int m = n / M, r = n - m * M;
                    const double2 z = fr[(r >> 1) * FS + sk(m)];
                    dst[n] = static_cast<float>((r & 1 ? z.y : z.x) * norm);
                }
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, o));
        hi = fmaxf(hi, __shfl_xor(hi, o));
    }
    __shared__ float s_lo[kRowThreads / 64], s_hi[kRowThreads / 64];
    if (tid % 64 == 0) {
        s_lo[tid / 64] = lo;
        s_hi[tid / 64] = hi;
    }
    __syncthreads();
    if (tid == 0) {
        for (int k = 1; k < kRowThreads / 64; ++k) {
            lo = fminf(lo, s_lo[k]);
            hi = fmaxf(hi, s_hi[k]);
        }
        int* m = reinterpret_cast<int*>(a.mm) + 4 * fb;
        atomicMin(m + 2, f2o(lo));
        atomicMax(m + 3, f2o(hi));
    }
}

// Row inverse for Ww = 256 (tiles of 256 input pixels and whole frames of 129-256): the same values as fht_rows, each
// length-256 complex FFT as a four-step FFT in registers (256 = 16 x 16) instead of radix-4 stages through LDS.
// For n = 16 n1 + n2 and m = m1 + 16 m2, e^{2 pi i n m / 256} = e^{2 pi i n1 m1 / 16} e^{2 pi i n2 m1 / 256}
// e^{2 pi i n2 m2 / 16}: lane n2 of a 16-lane group takes the 16 points n1 of its column, a 16-point inverse DFT in
// registers, the twiddle e^{2 pi i n2 m1 / 256}; one transpose through LDS inside the wave (no block barrier), and
// lane m1 takes a 16-point inverse DFT over n2, giving outputs m = m1 + 16 m2, stored straight from registers. A wave
// covers 4 rows (one per 16-lane group) and their ceil(M/2) phase pairs one after another; the column-pass values a
// lane needs (X[q] for its 16 points, the mirror points q = 256 - k as conjugates) stay in registers for the row.
// The phase twiddle of point q = 16 n +- j is split as e^{i t 16 n r} (wave-uniform, from LDS) times
// e^{+-i t j r} (per lane), t = 2 pi / (256 M).
constexpr int kR256Threads = 256;
constexpr int kR256Rows = 16;            // rows per block: 4 waves x 4 groups of 16 lanes
constexpr int kR256Ex = 16 * 17;         // exchange entries of one FFT: 16 rows of 16, row stride 17 (bank spread)

__host__ __device__ inline size_t r256_lds_bytes(int M) {
    const int P = (M + 1) / 2;
    return sizeof(double2) * (static_cast<size_t>(kR256Rows) * kR256Ex + 256 + static_cast<size_t>(P) * 50);
}

__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double2 cmulc(double2 a, double2 b) {  // a conj(b)
    return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// inverse DFT of 4 points in place: y[k] = sum_n x[n] i^{n k}
__device__ inline void idft4(double2& x0, double2& x1, double2& x2, double2& x3) {
    const double2 a0 = cadd(x0, x2), a1 = csub(x0, x2), b0 = cadd(x1, x3);
    const double2 b1 = make_double2(-(x1.y - x3.y), x1.x - x3.x);  // i (x1 - x3)
    x0 = cadd(a0, b0);
    x1 = cadd(a1, b1);
    x2 = csub(a0, b0);
    x3 = csub(a1, b1);
}

// inverse DFT of 16 points in place, n = na + 4 nb and k = k0 + 4 k1; output k is left at index 4 (k % 4) + k / 4
__device__ inline void idft16(double2 (&a)[16]) {
    constexpr double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173, h = 0.70710678118654752440;
#pragma unroll
    for (int na = 0; na < 4; ++na) idft4(a[na], a[na + 4], a[na + 8], a[na + 12]);
    // twiddles e^{2 pi i na k0 / 16} of element na + 4 k0
    a[5] = cmul(a[5], make_double2(c1, s1));
    a[9] = cmul(a[9], make_double2(h, h));
    a[13] = cmul(a[13], make_double2(s1, c1));
    a[6] = cmul(a[6], make_double2(h, h));
    a[10] = make_double2(-a[10].y, a[10].x);
    a[14] = cmul(a[14], make_double2(-h, h));
    a[7] = cmul(a[7], make_double2(s1, c1));
    a[11] = cmul(a[11], make_double2(-h, h));
    a[15] = cmul(a[15], make_double2(-c1, -s1));
#pragma unroll
    for (int k0 = 0; k0 < 4; ++k0) idft4(a[4 * k0], a[4 * k0 + 1], a[4 * k0 + 2], a[4 * k0 + 3]);
}

__global__ __launch_bounds__(kR256Threads) void fht_rows256(FhtTile a) {
    extern __shared__ double2 lds[];
    const int HM = a.Hw * a.M, M = a.M, N = 256 * M, P = (M + 1) / 2;
    constexpr int K = 128, hc = K + 1;
    double2* ex = lds;                          // [wave][group][kR256Ex]
    double2* tw = ex + kR256Rows * kR256Ex;     // e^{2 pi i t / 256}, t < 256
    double2* dpe = tw + 256;                    // [P][9] e^{i t 16 n (2p)}, n <= 8
    double2* dpo = dpe + 9 * P;                 // [P][9] e^{i t 16 n (2p + 1)}
    double2* cpe = dpo + 9 * P;                 // [P][16] e^{i t j (2p)}
    double2* cpo = cpe + 16 * P;                // [P][16] e^{i t j (2p + 1)}
    const int tid = threadIdx.x;
    for (int t = tid; t < 256 + 50 * P; t += kR256Threads) {
        long long k;  // exponent of e^{2 pi i k / N} (or / 256 for the first table), reduced exactly
        int den = N;
        if (t < 256) {
            k = t;
            den = 256;
        } else if (t < 256 + 18 * P) {
            const int u = t - 256, odd = u >= 9 * P, v = u - odd * 9 * P, p = v / 9, n = v - p * 9;
            k = 16LL * n * (2 * p + odd);
        } else {
            const int u = t - 256 - 18 * P, odd = u >= 16 * P, v = u - odd * 16 * P, p = v / 16, j = v - p * 16;
            k = static_cast<long long>(j) * (2 * p + odd);
        }
        double sn, cs;
        sincospi(2.0 * static_cast<double>(k % den) / den, &sn, &cs);
        tw[t] = make_double2(cs, sn);
    }
    const int ngrp = (HM + kR256Rows - 1) / kR256Rows;
    const int fb = blockIdx.x / ngrp;
    const int wave = tid >> 6, lane = tid & 63, g = lane >> 4, j = lane & 15;
    const int y = (blockIdx.x % ngrp) * kR256Rows + wave * 4 + g;
    const bool row_ok = y < HM;
    // this lane's column-pass values: point 16 i + j is X[16 i + j] for i < 8, the mirror X[256 - 16 i - j] for i >= 8
    double2 x[16];
    const double2* src = a.cols + static_cast<size_t>(fb) * hc * HM + (row_ok ? y : 0);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int q = i < 8 ? 16 * i + j : 16 * (16 - i) - j;
        x[i] = row_ok ? src[static_cast<size_t>(q) * HM] : make_double2(0.0, 0.0);
    }
    __syncthreads();  // tables written
    const double norm = 1.0 / (static_cast<double>(HM) * N);
    const int yo = y - a.yroi;
    const int oh = a.th * M, ow = a.tw * M, gh = a.h * M, gw = a.w * M;
    const bool row_kept = row_ok && yo >= 0 && yo < oh && on_lattice(a.r0 * M + yo, a.lat, gh);
    float* dst = a.realf + (static_cast<size_t>(fb) * HM + (row_ok ? y : 0)) * N;
    double2* xb = ex + (wave * 4 + g) * kR256Ex;
    const bool lat_div = a.lat >= 2 && M % a.lat == 0;
    float lo = INFINITY, hi = -INFINITY;
    for (int p = 0; p < P; ++p) {
        const bool two = 2 * p + 1 < M;
        const double2 ce = cpe[16 * p + j], co = cpo[16 * p + j];
        double2 v[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int n = i < 8 ? i : 16 - i;
            // phase twiddles of point q: e^{i t q 2p} and e^{i t q (2p + 1)}, q = 16 n + j (i < 8) or 16 n - j
            const double2 te = i < 8 ? cmul(dpe[9 * p + n], ce) : cmulc(dpe[9 * p + n], ce);
            const double2 to = i < 8 ? cmul(dpo[9 * p + n], co) : cmulc(dpo[9 * p + n], co);
            const double2 y1 = cmul(x[i], te);
            const double2 y2 = two ? cmul(x[i], to) : make_double2(0.0, 0.0);
            double2 z = i < 8 ? make_double2(y1.x - y2.y, y1.y + y2.x)     // y1 + i y2
                              : make_double2(y1.x + y2.y, -y1.y + y2.x);   // conj(y1) + i conj(y2)
            if (i == 0 && j == 0) z = make_double2(x[0].x, two ? x[0].x : 0.0);  // DC: real part only
            if (i == 8 && j == 0) z = make_double2(2.0 * y1.x, 2.0 * y2.x);      // k = K
            v[i] = z;
        }
        idft16(v);
        // A[m1] (at 4 (m1 % 4) + m1 / 4) times e^{2 pi i j m1 / 256}, into row m1 of the group's exchange block
#pragma unroll
        for (int m1 = 0; m1 < 16; ++m1) {
            const double2 w = v[4 * (m1 & 3) + (m1 >> 2)];
            xb[17 * m1 + j] = m1 ? cmul(w, tw[j * m1]) : w;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
#pragma unroll
        for (int n2 = 0; n2 < 16; ++n2) v[n2] = xb[17 * j + n2];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // these reads are done before the next pass writes
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        idft16(v);
        if (!row_ok) continue;
        // columns kept (lattice output): when lat divides M, column M m + 2p + h of the grid is a multiple of lat for
        // every m or for none, so the test is per (p, h); otherwise per column
        const int g0 = a.c0 * M - a.xroi;  // global column of grid position 0
        bool keep_ph[2] = {true, true};
        if (lat_div) {
#pragma unroll
            for (int hh = 0; hh < 2; ++hh) keep_ph[hh] = (((g0 + 2 * p + hh) % a.lat) + a.lat) % a.lat == 0;
        }
        // output m = j + 16 m2 (at 4 (m2 % 4) + m2 / 4): phases 2p (real part) and 2p + 1 (imaginary part)
#pragma unroll
        for (int m2 = 0; m2 < 16; ++m2) {
            const double2 z = v[4 * (m2 & 3) + (m2 >> 2)];
            const int n = M * (j + 16 * m2) + 2 * p;
#pragma unroll
            for (int hh = 0; hh < 2; ++hh) {
                if (hh == 1 && !two) break;
                const float val = static_cast<float>((hh ? z.y : z.x) * norm);
                lo = fminf(lo, val);
                hi = fmaxf(hi, val);
                const int xo = n + hh - a.xroi;
                if (!row_kept || xo < 0 || xo >= ow) continue;
                const bool col = a.lat < 2 || (lat_div ? keep_ph[hh] || g0 + n + hh == gw - 1
                                                       : on_lattice(g0 + n + hh, a.lat, gw));
                if (col) dst[n + hh] = val;
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, o));
        hi = fmaxf(hi, __shfl_xor(hi, o));
    }
    __shared__ float s_lo[kR256Threads / 64], s_hi[kR256Threads / 64];
    if (lane == 0) {
        s_lo[wave] = lo;
        s_hi[wave] = hi;
    }
    __syncthreads();
    if (tid == 0) {
        for (int k = 1; k < kR256Threads / 64; ++k) {
            lo = fminf(lo, s_lo[k]);
            hi = fmaxf(hi, s_hi[k]);
        }
        int* m = reinterpret_cast<int*>(a.mm) + 4 * fb;
        atomicMin(m + 2, f2o(lo));
        atomicMax(m + 3, f2o(hi));
    }
}

__device__ inline void mm_update(double d, double norm, float& lo, float& hi) {
    const float v = static_cast<float>(d * norm);
    lo = fminf(lo, v);
    hi = fmaxf(hi, v);
}

// float32 result of the inverse (real part / (HwM WwM)) and its per-frame min / max (grid: row = frame). Even
// planes are read as 16-byte pairs, four loads in flight per thread; min and max are exact, so the reduction order
// does not change them.
__global__ __launch_bounds__(256) void fht_minmax(FhtTile a) {
    const int HM = a.Hw * a.M, WM = a.Ww * a.M;
    const size_t plane = static_cast<size_t>(HM) * WM;
    const int fb = blockIdx.y;
    const double norm = 1.0 / (static_cast<double>(HM) * WM);
    const double* src = a.real + fb * plane;
    float lo = INFINITY, hi = -INFINITY;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    size_t p = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (plane % 2 == 0) {
        const double2* s2 = reinterpret_cast<const double2*>(src);
        const size_t n2 = plane / 2;
        for (; p + 3 * stride < n2; p += 4 * stride) {
            double2 v[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) v[k] = s2[p + k * stride];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                mm_update(v[k].x, norm, lo, hi);
                mm_update(v[k].y, norm, lo, hi);
            }
        }
        for (; p < n2; p += stride) {
            const double2 v = s2[p];
            mm_update(v.x, norm, lo, hi);
            mm_update(v.y, norm, lo, hi);
        }
    } else {
        for (; p < plane; p += stride) mm_update(src[p], norm, lo, hi);
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, o));
        hi = fmaxf(hi, __shfl_xor(hi, o));
    }
    __shared__ float s_lo[4], s_hi[4];
    const int wv = threadIdx.x / 64;
    if (threadIdx.x % 64 == 0) {
        s_lo[wv] = lo;
        s_hi[wv] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < static_cast<int>(blockDim.x / 64); ++k) {
            lo = fminf(lo, s_lo[k]);
            hi = fmaxf(hi, s_hi[k]);
        }
        int* m = reinterpret_cast<int*>(a.mm) + 4 * fb;
        atomicMin(m + 2, f2o(lo));
        atomicMax(m + 3, f2o(hi));
    }
}

// rescale to the input range in float32 as three rounded steps (interpolation.py:90-97), crop, write or blend. A
// block covers one output row of one frame (grid: frames x tile rows); four loads in flight per thread.
__global__ __launch_bounds__(256) void fht_finish(FhtTile a) {
#pragma clang fp contract(off)
    const int HM = a.Hw * a.M, WM = a.Ww * a.M;
    const int oh = a.th * a.M, ow = a.tw * a.M;
    // lattice output: blocks cover only the kept rows (multiples of lat inside the tile's rows, and the last row)
    const int nry = kept_count(a.r0 * a.M, oh, a.h * a.M, a.lat);
    const int fb = blockIdx.x / nry, ty = blockIdx.x % nry;
    const int y = a.lat < 2 ? ty : kept_index(a.r0 * a.M, oh, a.h * a.M, a.lat, ty);
    const int gw = a.w * a.M;
    const double norm = 1.0 / (static_cast<double>(HM) * WM);
    const int* m = reinterpret_cast<const int*>(a.mm) + 4 * fb;
    const float lo = o2f(m[0]), hi = o2f(m[1]), omin = o2f(m[2]), omax = o2f(m[3]);
    const bool rescale = omax != omin;
    const float scale = rescale ? static_cast<float>((static_cast<double>(hi) - lo) / (static_cast<double>(omax) - omin))
                                : 1.f;
    const size_t row = (static_cast<size_t>(fb) * HM + y + a.yroi) * WM + a.xroi;
    const double* src = a.real + row;
    const float* srcf = a.realf ? a.realf + row : nullptr;  // float32 grid written by fht_rows
    const float wr = a.tiled ? a.win_r[y] : 0.f;
    if (a.lat >= 2) {
        // lattice output, compact (N, LH, LW): kept row gy goes to gy / lat (the last row, when not a multiple, to
        // LH - 1); the tile's kept columns likewise
        const int gh = a.h * a.M, LH = kept_count(0, gh, gh, a.lat), LW = kept_count(0, gw, gw, a.lat);
        const int gy = a.r0 * a.M + y, ci = gy % a.lat == 0 ? gy / a.lat : LH - 1;
        const size_t orow = (static_cast<size_t>(a.f0 + fb) * LH + ci) * LW;
        const int nkc = kept_count(a.c0 * a.M, ow, gw, a.lat);
        for (int t = threadIdx.x; t < nkc; t += blockDim.x) {
            const int x = kept_index(a.c0 * a.M, ow, gw, a.lat, t), gx = a.c0 * a.M + x;
            const size_t o = orow + (gx % a.lat == 0 ? gx / a.lat : LW - 1);
            float v = srcf ? srcf[x] : static_cast<float>(src[x] * norm);
            if (rescale) {
                v = v - omin;
                v = v * scale;
                v = v + lo;
            }
            if (a.tiled) {
                const float wv = wr * a.win_c[x];
                a.out[o] += v * wv;
                a.weight[o] += wv;
            } else {
                a.out[o] = v;
            }
        }
        return;
    }
    const size_t wo = static_cast<size_t>(a.w) * a.M;
    const size_t o0 = static_cast<size_t>(a.f0 + fb) * a.h * a.M * wo + (static_cast<size_t>(a.r0) * a.M + y) * wo +
                      static_cast<size_t>(a.c0) * a.M;
    float* out = a.out + o0;
    constexpr int U = 4;
    for (int x0 = threadIdx.x; x0 < ow; x0 += U * blockDim.x) {
        float d[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const int x = x0 + k * static_cast<int>(blockDim.x);
            d[k] = x >= ow ? 0.f : (srcf ? srcf[x] : static_cast<float>(src[x] * norm));
        }
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const int x = x0 + k * static_cast<int>(blockDim.x);
            if (x >= ow) break;
            float v = d[k];
            if (rescale) {
                v = v - omin;
                v = v * scale;
                v = v + lo;
            }
            if (a.tiled) {
                const float wv = wr * a.win_c[x];
                out[x] += v * wv;
                a.weight[o0 + x] += wv;
            } else {
                out[x] = v;
            }
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

// Rows per pass of fht_rows for this width and magnification (0: the shape takes the complex-to-real FFT path):
// a power-of-two Ww >= 2, M >= 2, and the pass's FFT points and tables within the kernel's register and LDS budget.
int rows_pass(int Ww, int M) {
    if (M < 2 || Ww < 2 || (Ww & (Ww - 1)) != 0) return 0;
    for (int rb = 2; rb >= 1; --rb) {
        const RowShape r = row_shape(Ww, M, rb);
        if (r.npt <= kRowLdsPoints && r.bytes <= static_cast<size_t>(kRowLdsLimit)) return rb;
    }
    return 0;
}

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

int npx_fht_interpolate_lattice(const float* d_in, int N, int h, int w, int M, int mirror, int tile_size,
                                int overlap, int stride, float* d_out, void* d_work, size_t work_bytes, void* stream) {
    using namespace npx;
    if (stride < 0) return host::fail(NPX_ERR_ARG, "lattice stride must be >= 0");
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
    a.lat = stride;
    const size_t nout = stride >= 2 ? static_cast<size_t>(N) * kept_count(0, h * M, h * M, stride) *
                                          kept_count(0, w * M, w * M, stride)
                                    : static_cast<size_t>(N) * h * M * w * M;  // compact lattice or every point
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
                const size_t bigp = static_cast<size_t>(a.Hw) * M * a.Ww * M;
                hipLaunchKernelGGL(fht_init_mm, dim3((a.B + 255) / 256), dim3(256), 0, st, a);
                hipLaunchKernelGGL(fht_pad, dim3((a.Hw + kPadRows - 1) / kPadRows, a.B), dim3(256), 0, st, a);
                if (hipfftSetStream(pl.fwd, st) != HIPFFT_SUCCESS ||
                    hipfftExecD2Z(pl.fwd, reinterpret_cast<hipfftDoubleReal*>(a.pad),
                                  reinterpret_cast<hipfftDoubleComplex*>(a.spec)) != HIPFFT_SUCCESS)
                    return host::fail(NPX_ERR_HIP, "hipfftExecD2Z");
                const size_t ncol = static_cast<size_t>(a.Ww / 2 + 1) * a.Hw * M * a.B;
                hipLaunchKernelGGL(fht_embed_cols, dim3(grid_of(ncol)), dim3(256), 0, st, a);
                if (hipfftSetStream(pl.cols, st) != HIPFFT_SUCCESS ||
                    hipfftExecZ2Z(pl.cols, reinterpret_cast<hipfftDoubleComplex*>(a.cols),
                                  reinterpret_cast<hipfftDoubleComplex*>(a.cols), HIPFFT_BACKWARD) != HIPFFT_SUCCESS)
                    return host::fail(NPX_ERR_HIP, "hipfftExecZ2Z (columns)");
                const int rb = rows_pass(a.Ww, M);
                if (rb > 0 || (a.Ww == 256 && M >= 2 && r256_lds_bytes(M) <= size_t(kRowLdsLimit256))) {
                    a.realf = reinterpret_cast<float*>(a.real);
                    if (a.Ww == 256 && r256_lds_bytes(M) <= size_t(kRowLdsLimit256)) {
                        const int ngrp = (a.Hw * M + kR256Rows - 1) / kR256Rows;
                        hipLaunchKernelGGL(fht_rows256, dim3(ngrp * a.B), dim3(kR256Threads), r256_lds_bytes(M), st, a);
                    } else {
                        const int ngrp = (a.Hw * M + kRowsPerBlock - 1) / kRowsPerBlock;
                        hipLaunchKernelGGL(fht_rows, dim3(ngrp * a.B), dim3(kRowThreads), row_shape(a.Ww, M, rb).bytes,
                                           st, a, rb);
                    }
                    if (const int ny = kept_count(a.r0 * M, a.th * M, h * M, a.lat); ny > 0)
                    hipLaunchKernelGGL(fht_finish, dim3(ny * a.B), dim3(256), 0, st, a);
                    hipError_t e = hipGetLastError();
                    if (e != hipSuccess) return host::hip_fail(e, "Fourier interpolation launch");
                    continue;
                }
                a.realf = nullptr;
                const int ngrp = (a.Hw * M + kHalfRows - 1) / kHalfRows;
                hipLaunchKernelGGL(fht_to_half, dim3(ngrp * a.B), dim3(256), 0, st, a);
                if (hipfftSetStream(pl.rows, st) != HIPFFT_SUCCESS ||
                    hipfftExecZ2D(pl.rows, reinterpret_cast<hipfftDoubleComplex*>(a.half),
                                  reinterpret_cast<hipfftDoubleReal*>(a.real)) != HIPFFT_SUCCESS)
                    return host::fail(NPX_ERR_HIP, "hipfftExecZ2D (rows)");
                // about 8192 blocks over the batch, each sweeping 1024 pairs per step
                const size_t mmx = std::min<size_t>((bigp / 2 + 1023) / 1024, std::max(1, 8192 / a.B));
                hipLaunchKernelGGL(fht_minmax, dim3(static_cast<unsigned>(std::max<size_t>(1, mmx)), a.B), dim3(256),
                                   0, st, a);
                if (const int ny = kept_count(a.r0 * M, a.th * M, h * M, a.lat); ny > 0)
                    hipLaunchKernelGGL(fht_finish, dim3(ny * a.B), dim3(256), 0, st, a);
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

int npx_fht_interpolate(const float* d_in, int N, int h, int w, int M, int mirror, int tile_size, int overlap,
                        float* d_out, void* d_work, size_t work_bytes, void* stream) {
    return npx_fht_interpolate_lattice(d_in, N, h, w, M, mirror, tile_size, overlap, 0, d_out, d_work, work_bytes,
                                       stream);
}

}  // extern "C"
