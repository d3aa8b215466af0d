// gfx950 kernels of the eSRRF hot path.
//
// esrrf_fused: one launch runs Roberts-cross gradients, the Catmull-Rom x2M gradient lattice, the RGC
// ring sum, the Catmull-Rom xM intensity weight and (temporal mode) the AVG/VAR/TAC2 moments for a
// stack of frames. A block owns a 64-column x Q-row region of the input image and up to 16 of the M*M
// output phases (one wavefront per phase, one lane per input column, Q rows per lane); every lane of a
// wavefront shares the phase, so the ring geometry (A, B, w per window offset) is wave-uniform and is
// read through scalar loads. Per frame the block stages the input tile + halo, the Roberts gradients and
// the normalised gradient lattice in LDS; the lattice is stored in two column-parity planes so that the
// 64 lanes of a read touch 64 consecutive float2. Ring sums accumulate in registers.
//
// Reference semantics: src/include/_c_sr_radial_gradient_convergence.c:17-90 (ring sum),
// src/include/_c_interpolation_catmull_rom.c:6-52 (taps), src/include/_c_gradients.c:53-72 (Roberts),
// src/nanopyx/core/transform/_le_esrrf.pyx:172-195 (composition),
// src/nanopyx/core/transform/sr_temporal_correlations.py:46-75,205-221 (temporal moments).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "npx_fused.h"
#include "npx_kernels.h"

namespace npx {

size_t fused_lds_bytes(int KL, int Q, int nwaves, int mode) {
#define NPX_LDS_CASE(k)                                                                      \
    case k:                                                                                  \
        return Q == kQFrames ? lds_bytes_kq<k, kQFrames>(nwaves, mode) : lds_bytes_kq<k, kQTemporal>(nwaves, mode);
    switch (KL) {
        NPX_LDS_CASE(0)
        NPX_LDS_CASE(1)
        NPX_LDS_CASE(2)
        NPX_LDS_CASE(3)
        NPX_LDS_CASE(4)
        NPX_LDS_CASE(5)
        NPX_LDS_CASE(6)
        default: return 0;
    }
#undef NPX_LDS_CASE
}

// One job: grid (regions x, regions y, phase groups x frame chunks).
template <int K, int Q, int MODE>
__global__ __launch_bounds__(1024) void esrrf_fused(FusedArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    fused_block<K, Q, MODE, false>(a, blockIdx.x, blockIdx.y, blockIdx.z % a.ngroups, blockIdx.z / a.ngroups, smem);
}

// Sweep chunks: the mean over all frames is the frame-count-weighted mean of the chunk means. part holds
// [chunk][nout][P]; output o of pixel k is outs[o][k].
struct MeanOuts {
    float* o[kMaxRad * kMaxSens];
};

__global__ __launch_bounds__(256) void means_merge(const float* __restrict__ part, int nchunk, int fpc, int N, size_t P,
                                                   int nout, MeanOuts outs) {
    const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= P) return;
    for (int o = 0; o < nout; ++o) {
        float mu = 0.f, n = 0.f;
        for (int ch = 0; ch < nchunk; ++ch) {
            const int nb_i = min(N, (ch + 1) * fpc) - ch * fpc;
            if (nb_i <= 0) break;
            const float nb = static_cast<float>(nb_i);
            const float mb = part[(static_cast<size_t>(ch) * nout + o) * P + k];
            n += nb;
            mu += (mb - mu) * (nb / n);
        }
        outs.o[o][k] = mu;
    }
}

// Merges per-chunk moments in frame order (Chan for the mean/M2; the lag-1 co-moment is re-centred on the
// merged mean and gains the cross term of the two frames that meet at the chunk boundary). With a running state
// of n0 > 0 frames, the chunks continue it; the result goes to the maps or (to_state) back to the state.
__global__ __launch_bounds__(256) void moments_merge(const float* __restrict__ part, int nchunk, int fpc, int N,
                                                     size_t P, float* avg, float* var, float* tac2, float* state,
                                                     int n0, int to_state) {
    const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= P) return;
    float n = 0.f, mu = 0.f, m2 = 0.f, c = 0.f, first = 0.f, last = 0.f;
    if (state && n0 > 0) {
        n = static_cast<float>(n0);
        mu = state[k];
        m2 = state[P + k];
        c = state[2 * P + k];
        first = state[3 * P + k];
        last = state[4 * P + k];
    }
    for (int ch = 0; ch < nchunk; ++ch) {
        const int f0 = ch * fpc;
        const int nb_i = min(N, f0 + fpc) - f0;
        if (nb_i <= 0) break;
        const float nb = static_cast<float>(nb_i);
        const float* b = part + static_cast<size_t>(ch) * 5 * P + k;
        const float mub = b[0], m2b = b[P], cb = b[2 * P], fb = b[3 * P], lb = b[4 * P];
        if (n == 0.f) {
            n = nb;
            mu = mub;
            m2 = m2b;
            c = cb;
            first = fb;
            last = lb;
            continue;
        }
        const float nn = n + nb;
        const float delta = mub - mu;
        const float mun = mu + delta * (nb / nn);
        // re-centre both co-moments on mun: C(m) = C - e*((last - mu) + (first - mu)) + (n - 1) e^2, e = mu - m
        const float ea = mu - mun, eb = mub - mun;
        const float ca = c - ea * ((last - mu) + (first - mu)) + (n - 1.0f) * ea * ea;
        const float cbm = cb - eb * ((lb - mub) + (fb - mub)) + (nb - 1.0f) * eb * eb;
        c = ca + cbm + (last - mun) * (fb - mun);
        m2 = m2 + m2b + delta * delta * (n * nb / nn);
        mu = mun;
        n = nn;
        last = lb;
    }
    if (to_state) {
        state[k] = mu;
        state[P + k] = m2;
        state[2 * P + k] = c;
        state[3 * P + k] = first;
        state[4 * P + k] = last;
        return;
    }
    if (avg) avg[k] = mu;
    if (var) var[k] = m2 / n;
    if (tac2) tac2[k] = c / (n - 1.0f);
}

// ------------------------------------------------------------------------------------------ RGC on caller arrays

// RadialGradientConvergence on arbitrary (N, 2HM, 2WM) gradient arrays and an (N, HM, WM) intensity,
// read exactly where _c_calculate_rgc reads them (:17-90), one thread per output pixel, float32
// throughout (the distance weight uses expf where the reference uses double exp).
__global__ __launch_bounds__(256) void rgc_arrays(const float* __restrict__ gxi, const float* __restrict__ gyi,
                                                  const float* __restrict__ img, int N, int HM, int WM, int M,
                                                  float fwhm, float tSO, float tSS, int pow_mode, float S,
                                                  int iw, int margin, float* __restrict__ out) {
    const size_t P = static_cast<size_t>(HM) * WM;
    const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= P * N) return;
    const int f = static_cast<int>(k / P);
    const size_t pix = k - static_cast<size_t>(f) * P;
    const int rM = static_cast<int>(pix / WM);
    const int cM = static_cast<int>(pix - static_cast<size_t>(rM) * WM);
    if (rM < margin || rM >= HM - margin || cM < margin || cM >= WM - margin) {
        out[k] = 0.f;
        return;
    }
    const float G2 = 2.0f;
    const float xc = static_cast<float>(2 * cM + 1) / static_cast<float>(2 * M);
    const float yc = static_cast<float>(2 * rM + 1) / static_cast<float>(2 * M);
    const int start = -static_cast<int>(G2 * fwhm);
    const int end = static_cast<int>(G2 * fwhm + 1.0f);
    const int gw = static_cast<int>(WM * G2);
    const float* gx = gxi + static_cast<size_t>(f) * 4 * P;
    const float* gy = gyi + static_cast<size_t>(f) * 4 * P;
    const float inv_tSS = 1.0f / tSS;
    float rgc = 0.f, dws = 0.f;
    for (int j = start; j < end; ++j) {
        const float vy = static_cast<float>(static_cast<int>(G2 * yc) + j) / G2;
        if (!(0 < vy && vy <= HM - 1)) continue;
        for (int i = start; i < end; ++i) {
            const float vx = static_cast<float>(static_cast<int>(G2 * xc) + i) / G2;
            if (!(0 < vx && vx <= WM - 1)) continue;
            const float dx = vx - xc, dy = vy - yc;
            const float d = sqrtf(dx * dx + dy * dy);
            if (!(d != 0.f && d <= tSO)) continue;
            float pyf = vy * M * G2, pxf = vx * M * G2;
            if (pyf >= HM * G2) pyf = HM * G2 - 1;
            if (pxf >= WM * G2) pxf = WM * G2 - 1;
            if (pyf < 0) pyf = 0;
            if (pxf < 0) pxf = 0;
            const size_t o = static_cast<size_t>(static_cast<int>(pyf)) * gw + static_cast<int>(pxf);
            const float Gx = gx[o], Gy = gy[o];
            const float e = d * __expf(-(d * d) * inv_tSS);
            const float e2 = e * e;
            const float dw = e2 * e2;
            dws += dw;
            // G.d < 0 implies G != 0, so |G| > 0 below (the reference's NaN branch is never taken on this path)
            if (Gx * dx + Gy * dy < 0) {
                const float Dk = 1.0f - fabsf(Gy * dx - Gx * dy) * __builtin_amdgcn_rsqf(Gx * Gx + Gy * Gy) *
                                            __builtin_amdgcn_rcpf(d);
                rgc += Dk * dw;
            }
        }
    }
    rgc = dws > 0 ? rgc / dws : 0.f;
    if (rgc < 0) rgc = 0.f;
    rgc = apply_pow(rgc, pow_mode, S);
    if (iw) rgc = rgc * img[k];
    out[k] = rgc;
}

// ------------------------------------------------------------------------------------------ temporal / post

// AVG / VAR / TAC2 of a device (N, P) stack with the streaming moments above (one thread per pixel).
__global__ __launch_bounds__(256) void moments_stack(const float* __restrict__ stack, int N, size_t P,
                                                     float* avg, float* var, float* tac2) {
    const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= P) return;
    float mu = 0.f, m2 = 0.f, c = 0.f, first = 0.f, prev = 0.f;
    for (int t = 0; t < N; ++t)
        moments_push(stack[static_cast<size_t>(t) * P + k], t, 1.0f / static_cast<float>(t + 1), mu, m2, c, first,
                     prev);
    if (avg) avg[k] = mu;
    if (var) var[k] = m2 / static_cast<float>(N);
    if (tac2) tac2[k] = c / static_cast<float>(N - 1);
}

// macro_pixel_corrector (src/nanopyx/core/transform/mpcorrector.pyx:7-73): per slice and phase (y, x),
// the mean over the rows x cols macro-pixels of img / (rows * cols).
__global__ __launch_bounds__(256) void macro_pixel_sums(const float* __restrict__ img, int HM, int WM, int M,
                                                        float* __restrict__ cmap) {
    const int ph = blockIdx.x, s = blockIdx.y;
    const int y = ph / M, x = ph % M;
    const int rows = HM / M, cols = WM / M;
    const float scale = static_cast<float>(rows * cols);
    const float* im = img + static_cast<size_t>(s) * HM * WM;
    float acc = 0.f;
    for (int k = threadIdx.x; k < rows * cols; k += blockDim.x) {
        const int ry = k / cols, rx = k - ry * cols;
        acc += im[static_cast<size_t>(ry * M + y) * WM + rx * M + x] / scale;
    }
    __shared__ float red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) cmap[s * M * M + ph] = red[0];
}

__global__ __launch_bounds__(256) void macro_pixel_divide(float* __restrict__ img, int S, int HM, int WM, int M,
                                                          const float* __restrict__ cmap) {
    const size_t P = static_cast<size_t>(HM) * WM;
    const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= P * S) return;
    const int s = static_cast<int>(k / P);
    const size_t pix = k - static_cast<size_t>(s) * P;
    const int yM = static_cast<int>(pix / WM), xM = static_cast<int>(pix % WM);
    const float c = cmap[s * M * M + (yM % M) * M + (xM % M)];
    if (c != 0.f) img[k] = img[k] / c;
}

// ------------------------------------------------------------------------------------------ launchers

template <int K>
static hipError_t launch_fused_k(const FusedArgs& a, dim3 grid, dim3 block, int mode, hipStream_t st) {
    const int Q = (mode == kModeFrames || mode == kModeArrays) ? frames_q(a.nwaves, a.ngroups) : kQTemporal;
    const size_t dyn = fused_lds_bytes(K, Q, a.nwaves, mode);
    if (dyn > static_cast<size_t>(kLdsLimit)) return hipErrorInvalidConfiguration;
    if (mode == kModeFrames && Q == kQFrames)
        hipLaunchKernelGGL((esrrf_fused<K, kQFrames, kModeFrames>), grid, block, dyn, st, a);
    else if (mode == kModeFrames)
        hipLaunchKernelGGL((esrrf_fused<K, kQTemporal, kModeFrames>), grid, block, dyn, st, a);
    else if (mode == kModeArrays && Q == kQFrames)
        hipLaunchKernelGGL((esrrf_fused<K, kQFrames, kModeArrays>), grid, block, dyn, st, a);
    else if (mode == kModeArrays)
        hipLaunchKernelGGL((esrrf_fused<K, kQTemporal, kModeArrays>), grid, block, dyn, st, a);
    else if (mode == kModeTemporal)
        hipLaunchKernelGGL((esrrf_fused<K, kQTemporal, kModeTemporal>), grid, block, dyn, st, a);
    else if (mode == kModeTemporalVar)
        hipLaunchKernelGGL((esrrf_fused<K, kQTemporal, kModeTemporalVar>), grid, block, dyn, st, a);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_fused(const FusedArgs& a, dim3 grid, dim3 block, int mode, hipStream_t st) {
    if (a.nrad != 1 || a.K[0] != a.KL) return hipErrorInvalidValue;
    switch (a.KL) {
        case 0: return launch_fused_k<0>(a, grid, block, mode, st);
        case 1: return launch_fused_k<1>(a, grid, block, mode, st);
        case 2: return launch_fused_k<2>(a, grid, block, mode, st);
        case 3: return launch_fused_k<3>(a, grid, block, mode, st);
        case 4: return launch_fused_k<4>(a, grid, block, mode, st);
        case 5: return launch_fused_k<5>(a, grid, block, mode, st);
        case 6: return launch_fused_k<6>(a, grid, block, mode, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_means_merge(const float* part, int nchunk, int fpc, int N, size_t P, int nout, float* const* outs,
                              hipStream_t st) {
    if (nout < 1 || nout > kMaxRad * kMaxSens) return hipErrorInvalidValue;
    MeanOuts o{};
    for (int i = 0; i < nout; ++i) o.o[i] = outs[i];
    const unsigned blocks = static_cast<unsigned>((P + 255) / 256);
    hipLaunchKernelGGL(means_merge, dim3(blocks), dim3(256), 0, st, part, nchunk, fpc, N, P, nout, o);
    return hipGetLastError();
}

hipError_t launch_merge(const float* part, int nchunk, int fpc, int N, size_t P, float* avg, float* var, float* tac2,
                        hipStream_t st) {
    return launch_merge_state(part, nchunk, fpc, N, P, nullptr, 0, 0, avg, var, tac2, st);
}

hipError_t launch_merge_state(const float* part, int nchunk, int fpc, int N, size_t P, float* state, int n0,
                              int to_state, float* avg, float* var, float* tac2, hipStream_t st) {
    const unsigned blocks = static_cast<unsigned>((P + 255) / 256);
    hipLaunchKernelGGL(moments_merge, dim3(blocks), dim3(256), 0, st, part, nchunk, fpc, N, P, avg, var, tac2, state,
                       n0, to_state);
    return hipGetLastError();
}

hipError_t launch_rgc_arrays(const float* gx, const float* gy, const float* img, int N, int HM, int WM, int M,
                             float fwhm, float tSO, float tSS, int pow_mode, float S, int iw, int margin, float* out,
                             hipStream_t st) {
    const size_t total = static_cast<size_t>(N) * HM * WM;
    const unsigned blocks = static_cast<unsigned>((total + 255) / 256);
    hipLaunchKernelGGL(rgc_arrays, dim3(blocks), dim3(256), 0, st, gx, gy, img, N, HM, WM, M, fwhm, tSO, tSS, pow_mode,
                       S, iw, margin, out);
    return hipGetLastError();
}

hipError_t launch_moments_stack(const float* stack, int N, size_t P, float* avg, float* var, float* tac2,
                                hipStream_t st) {
    const unsigned blocks = static_cast<unsigned>((P + 255) / 256);
    hipLaunchKernelGGL(moments_stack, dim3(blocks), dim3(256), 0, st, stack, N, P, avg, var, tac2);
    return hipGetLastError();
}

hipError_t launch_macro_pixel(float* img, int S, int HM, int WM, int M, float* cmap, hipStream_t st) {
    hipLaunchKernelGGL(macro_pixel_sums, dim3(M * M, S), dim3(256), 0, st, img, HM, WM, M, cmap);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const size_t total = static_cast<size_t>(S) * HM * WM;
    hipLaunchKernelGGL(macro_pixel_divide, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, st, img, S,
                       HM, WM, M, cmap);
    return hipGetLastError();
}

}  // namespace npx
