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

#include "npx_kernels.h"

namespace npx {

// ------------------------------------------------------------------------------------------ helpers

__device__ __forceinline__ float apply_pow(float v, int pow_mode, float S) {
    // reference: if RGC >= 0 && S > 1: RGC = pow(RGC, S) (:83-87); RGC here is never negative. Integer
    // sensitivities are exact products; others use 2^(S log2 v) on the hardware log2/exp2 (a few ulp).
    if (pow_mode == 2) return v * v;
    if (pow_mode == 3) return v * v * v;
    if (pow_mode == 4) {
        const float s = v * v;
        return s * s;
    }
    if (pow_mode == 1) return v > 0.f ? __builtin_amdgcn_exp2f(S * __builtin_amdgcn_logf(v)) : 0.f;
    return v;
}

// One step of the streaming moments. State: mu (mean), m2 (sum of squared deviations), c (lag-1
// co-moment sum_t (x_t - mu)(x_{t+1} - mu)), first and prev values; n = values already absorbed.
template <bool WITH_M2 = true>
__device__ __forceinline__ void moments_push(float v, int n, float inv_np1, float& mu, float& m2, float& c,
                                             float& first, float& prev) {
    if (n == 0) {
        mu = v;
        if (WITH_M2) m2 = 0.f;
        c = 0.f;
        first = v;
        prev = v;
        return;
    }
    const float delta = v - mu;
    const float mun = mu + delta * inv_np1;
    const float dmu = mun - mu;
    c = c + dmu * ((prev - mu) + (first - mu)) + static_cast<float>(n - 1) * dmu * dmu + (prev - mun) * (v - mun);
    if (WITH_M2) m2 = m2 + delta * (v - mun);
    mu = mun;
    prev = v;
}

// ------------------------------------------------------------------------------------------ fused kernel

__device__ __forceinline__ float load_px(const FusedArgs& a, int f, int gy, int gx) {
    if (gy < 0 || gy >= a.H || gx < 0 || gx >= a.W) return 0.f;
    const size_t o = (static_cast<size_t>(f) * a.H + gy) * a.W + gx;
    return a.u16 ? static_cast<float>(static_cast<const uint16_t*>(a.in)[o]) : static_cast<const float*>(a.in)[o];
}

template <int K, int Q, int MODE>
__global__ __launch_bounds__(1024) void esrrf_fused(FusedArgs a) {
    using G = Geometry<K, Q>;
    constexpr int NW = G::NW, RS = G::RS;
    constexpr int HG = G::HG, HI = G::HI;
    constexpr int TIR = G::TIR, TIC = G::TIC;
    constexpr int TGR = G::TGR, TGC = G::TGC;
    constexpr int LR = G::LR, LP = G::LP, LPL = LR * LP;
    constexpr int LB = 128 + 2 * K;  // lattice columns per block
    constexpr int NR = 2 * Q + 2 * K - 1;
    constexpr int RDX = NW, RDT = 2 * NW, RSUM = 2 * NW + 3 * NW * NW, RPRE = RSUM + 1;

    __shared__ float s_in[TIR * TIC];
    __shared__ float s_rx[LR * TGC];   // gradients filtered along rows (one row per lattice row)
    __shared__ float s_ry[LR * TGC];
    __shared__ float2 s_lat[2 * LPL];
    // the Roberts gradients are dead once the row pass has read them, before the lattice is written: they
    // share the lattice's storage
    static_assert(2 * TGR * TGC <= 4 * LPL, "gradient tile must fit in the lattice storage");
    float* const s_gc = reinterpret_cast<float*>(s_lat);
    float* const s_gr = s_gc + TGR * TGC;
    // frame-invariant tables of this block, staged once: lattice row / column taps (tile-relative bases,
    // zero weights where the ring test excludes the lattice line), intensity row taps per (wave, q)
    __shared__ float4 s_lrw[LR];
    __shared__ int s_lrb[LR];
    __shared__ float4 s_lcw[LB];
    __shared__ int s_lcb[LB];
    __shared__ float4 s_crw[16 * Q];
    __shared__ int s_crb[16 * Q];
    // dynamic: nwaves ring blocks of RS floats, then (temporal mode) the first frame's value of every pixel
    extern __shared__ float s_ring[];

    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const int x0 = blockIdx.x * 64;
    const int y0 = blockIdx.y * Q;
    const int group = blockIdx.z % a.ngroups;
    const int chunk = blockIdx.z / a.ngroups;
    const int f_begin = chunk * a.fpc;
    const int f_end = min(a.N, f_begin + a.fpc);
    const int ph = __builtin_amdgcn_readfirstlane(group * a.nwaves + wave);
    const bool active = ph < a.nphase;
    const int M = a.M;
    const int py = active ? ph / M : 0;
    const int px = active ? ph % M : 0;
    const int ey = (2 * py + 1) / M;
    const int ex = (2 * px + 1) / M;
    const int xi = x0 + lane;
    const int cM = M * xi + px;
    const int a_lo = 2 * y0 - K;
    const int b_lo = 2 * x0 - K;

    // ---- stage the frame-invariant tables
    for (int k = tid; k < a.nwaves * RS; k += nthr) {
        const int w = k / RS, phw = group * a.nwaves + w;
        s_ring[k] = phw < a.nphase ? a.t.ring[static_cast<size_t>(phw) * RS + (k - w * RS)] : 0.f;
    }
    for (int la = tid; la < LR; la += nthr) {
        const int aa = a_lo + la;
        float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
        int b = 0;
        if (aa >= 1 && aa <= 2 * a.HM - 2) {  // rows a >= 2H read the clamped last row (:45-52)
            const int ai = min(aa, 2 * a.H);
            w = a.t.lrw[ai];
            b = a.t.lrb[ai] - (y0 - HG);
        }
        s_lrw[la] = w;
        s_lrb[la] = b;
    }
    for (int lb = tid; lb < LB; lb += nthr) {
        const int bb = b_lo + lb;
        float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
        int b = 0;
        if (bb >= 1 && bb <= 2 * a.WM - 2) {
            const int bi = min(bb, 2 * a.W);
            w = a.t.lcw[bi];
            b = a.t.lcb[bi] - (x0 - HG);
        }
        s_lcw[lb] = w;
        s_lcb[lb] = b;
    }
    for (int k = tid; k < a.nwaves * Q; k += nthr) {
        const int w = k / Q, q = k - w * Q, phw = group * a.nwaves + w;
        const int rM = min(M * (y0 + q) + (phw < a.nphase ? phw / M : 0), a.HM - 1);
        s_crw[k] = a.t.crw[rM];
        s_crb[k] = a.t.crb[rM] - (y0 - HI);
    }
    const int cMs = min(cM, a.WM - 1);
    const float4 wcol = a.t.ccw[cMs];
    const int ccol = a.t.ccb[cMs] - (x0 - HI);

    constexpr bool WITH_M2 = MODE == kModeTemporalVar;
    float mu[Q], m2[WITH_M2 ? Q : 1], cc[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) mu[q] = cc[q] = 0.f;
#pragma unroll
    for (int q = 0; q < (WITH_M2 ? Q : 1); ++q) m2[q] = 0.f;
    // first and previous frame's value of every pixel of the wave: [wave][q][lane] and [nwaves + wave][q][lane]
    float* s_first = s_ring + a.nwaves * RS + (wave * Q) * 64 + lane;
    float* s_prev = s_first + a.nwaves * Q * 64;

    // input-tile elements this thread stages; the first kPrefetch are fetched one frame ahead
    float pf[kPrefetch];
#pragma unroll
    for (int k = 0; k < kPrefetch; ++k) {
        const int idx = tid + k * nthr;
        pf[k] = 0.f;
        if (idx < TIR * TIC && f_begin < f_end) {
            const int r = idx / TIC, c = idx - r * TIC;
            pf[k] = load_px(a, f_begin, y0 - HI + r, x0 - HI + c);
        }
    }
    const float* ringL = s_ring + wave * RS;

    for (int f = f_begin; f < f_end; ++f) {
        // (1) input tile + halo of frame f (zero outside the image; uint16 converted as a plain cast)
#pragma unroll
        for (int k = 0; k < kPrefetch; ++k) {
            const int idx = tid + k * nthr;
            if (idx < TIR * TIC) s_in[idx] = pf[k];
        }
        for (int idx = tid + kPrefetch * nthr; idx < TIR * TIC; idx += nthr) {
            const int r = idx / TIC, c = idx - r * TIC;
            s_in[idx] = load_px(a, f, y0 - HI + r, x0 - HI + c);
        }
        __syncthreads();
        // prefetch frame f + 1 while this frame is processed (consumed at the top of the next iteration)
        if (f + 1 < f_end) {
#pragma unroll
            for (int k = 0; k < kPrefetch; ++k) {
                const int idx = tid + k * nthr;
                if (idx < TIR * TIC) {
                    const int r = idx / TIC, c = idx - r * TIC;
                    pf[k] = load_px(a, f + 1, y0 - HI + r, x0 - HI + c);
                }
            }
        }
        // (2) Roberts cross (_c_gradients.c:53-72); gradients outside the image are 0 (CR skips them)
        for (int idx = tid; idx < TGR * TGC; idx += nthr) {
            const int r = idx / TGC, c = idx - r * TGC;
            const int gy = y0 - HG + r, gx = x0 - HG + c;
            float gcv = 0.f, grv = 0.f;
            if (gy >= 0 && gy < a.H && gx >= 0 && gx < a.W) {
                const int r1 = gy - (y0 - HI), c1 = gx - (x0 - HI);
                const int r0 = (gy > 0 ? gy - 1 : 0) - (y0 - HI);
                const int c0 = (gx > 0 ? gx - 1 : 0) - (x0 - HI);
                const float a_r0c1 = s_in[r0 * TIC + c1];
                const float a_r1c0 = s_in[r1 * TIC + c0];
                const float a_r0c0 = s_in[r0 * TIC + c0];
                const float a_r1c1 = s_in[r1 * TIC + c1];
                gcv = a_r0c1 - a_r1c0 + a_r1c1 - a_r0c0;
                grv = -a_r0c1 + a_r1c0 + a_r1c1 - a_r0c0;
            }
            s_gc[idx] = gcv;
            s_gr[idx] = grv;
        }
        __syncthreads();
        // (3a) Catmull-Rom along rows for every lattice row: the x2M gradient image sampled at row a/2
        for (int idx = tid; idx < LR * TGC; idx += nthr) {
            const int la = idx / TGC, c = idx - la * TGC;
            const float4 wr = s_lrw[la];
            const int rb = s_lrb[la];
            const float wrv[4] = {wr.x, wr.y, wr.z, wr.w};
            float sx = 0.f, sy = 0.f;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int tr = min(max(rb + u, 0), TGR - 1);
                sx += s_gc[tr * TGC + c] * wrv[u];
                sy += s_gr[tr * TGC + c] * wrv[u];
            }
            s_rx[idx] = sx;
            s_ry[idx] = sy;
        }
        __syncthreads();
        // (3b) ... then along columns, normalised: lattice point (a, b) holds G/|G| at (a/2, b/2); stored in two
        // column-parity planes so the 64 lanes of one read touch 64 consecutive float2
        for (int idx = tid; idx < 2 * LPL; idx += nthr) {
            const int plane = idx / LPL;
            const int rem = idx - plane * LPL;
            const int la = rem / LP;
            const int pos = rem - la * LP;
            const int lb = 2 * pos + plane;
            const float4 wc = s_lcw[lb];
            const int cb = s_lcb[lb];
            const float wcv[4] = {wc.x, wc.y, wc.z, wc.w};
            float gx = 0.f, gy = 0.f;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int tc = min(max(cb + v, 0), TGC - 1);
                gx += s_rx[la * TGC + tc] * wcv[v];
                gy += s_ry[la * TGC + tc] * wcv[v];
            }
            const float n2 = gx * gx + gy * gy;
            float2 g = make_float2(0.f, 0.f);
            if (n2 > 0.f) {
                const float inv = rsqrtf(n2);
                g = make_float2(gx * inv, gy * inv);
            }
            s_lat[idx] = g;
        }
        __syncthreads();

        // (4) ring sums. Sample (j, i) of pixel q reads lattice row 2q + ey + j + K, column 2*lane + ex + i + K.
        // With u = g.(dx, dy) and v = g x (dx, dy) (g = G/|G|, d = |(dx, dy)|, t = w/d), the reference's
        // contribution [G.d < 0] * w * (1 - |G x d| / (|G| d)) is t * med3(0, -u, d - |v|). g.x*dx and g.y*dx are
        // formed once per lattice point of the column and shared by the samples that read it. The ring
        // constants come from LDS (the same address for all lanes: one broadcast read).
        if (active) {
            float acc[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) acc[q] = 0.f;
            // ring constants through scalar loads (wave-uniform, SGPRs): keeps the VGPRs for pixels and lattice rows
            using cfloat = const __attribute__((address_space(4))) float;  // constant address space: s_load
            cfloat* rg = (cfloat*)(a.t.ring + static_cast<size_t>(ph) * RS);
            float dyv[NW];
#pragma unroll
            for (int jj = 0; jj < NW; ++jj) dyv[jj] = rg[jj];
#pragma unroll 1
            for (int i = 0; i < NW; ++i) {
                const int c = ex + i;
                const float2* col = s_lat + (c & 1) * LPL + ey * LP + lane + (c >> 1);
                const float dxi = rg[RDX + i];
                cfloat* dt = rg + RDT + 2 * i * NW;
                float dv[NW], tv[NW];
#pragma unroll
                for (int jj = 0; jj < NW; ++jj) {
                    dv[jj] = dt[2 * jj];
                    tv[jj] = dt[2 * jj + 1];
                }
                // lattice rows are read two ahead of their use so the LDS latency overlaps the ring arithmetic
                float2 g1 = col[0];
                float2 g2 = col[LP];
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    const float2 g = g1;
                    g1 = g2;
                    if (r + 2 < NR) g2 = col[(r + 2) * LP];
                    const float px_ = g.x * dxi;
                    const float py_ = g.y * dxi;
#pragma unroll
                    for (int jj = 0; jj < NW; ++jj) {
                        const int rq = r - jj;
                        if (rq < 0 || (rq & 1) || (rq >> 1) >= Q) continue;
                        const int q = rq >> 1;
                        const float u = fmaf(g.y, dyv[jj], px_);
                        const float v = fmaf(-g.x, dyv[jj], py_);
                        const float m = __builtin_amdgcn_fmed3f(0.f, -u, dv[jj] - fabsf(v));
                        acc[q] = fmaf(tv[jj], m, acc[q]);
                    }
                }
            }
            // (5) normalise, sensitivity, intensity weight; then store or fold into the moments. Branch-free per
            // pixel: pixels whose window crosses the border take their normaliser from the prefix-sum table.
            const float wcv[4] = {wcol.x, wcol.y, wcol.z, wcol.w};
            const float inv_np1 = 1.0f / static_cast<float>(f - f_begin + 1);
            const float wsum = ringL[RSUM];
            const int b0 = 2 * xi + ex;
            const int i0 = max(0, 1 - (b0 - K));
            const int i1 = max(0, min(NW, 2 * a.WM - 2 - (b0 - K) + 1));  // lanes past the image: empty range
            const bool col_ok = xi < a.W && cM >= a.margin && cM < a.WM - a.margin;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int yi = y0 + q;
                const int rM = M * yi + py;
                const int a0 = 2 * yi + ey;
                const int j0 = max(0, 1 - (a0 - K));
                const int j1 = max(0, min(NW, 2 * a.HM - 2 - (a0 - K) + 1));
                const bool row_ok = yi < a.H && rM >= a.margin && rM < a.HM - a.margin;  // uniform
                float den = wsum;
                if (j0 > 0 || j1 < NW || i0 > 0 || i1 < NW) {
                    const float* P = ringL + RPRE;
                    const int jl = min(j0, j1), il = min(i0, i1);
                    den = (P[j1 * (NW + 1) + i1] - P[jl * (NW + 1) + i1]) - (P[j1 * (NW + 1) + il] - P[jl * (NW + 1) + il]);
                }
                float rgc = den > 0.f ? acc[q] / den : 0.f;
                rgc = apply_pow(rgc, a.pow_mode, a.S);
                if (a.iw) {
                    const float4 wrow = s_crw[wave * Q + q];
                    const int crow = s_crb[wave * Q + q];
                    const float wrv[4] = {wrow.x, wrow.y, wrow.z, wrow.w};
                    float sacc = 0.f;
#pragma unroll
                    for (int v2 = 0; v2 < 4; ++v2) {
                        const int tc = min(max(ccol + v2, 0), TIC - 1);
                        float pp = 0.f;
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const int tr = min(max(crow + u, 0), TIR - 1);
                            pp += s_in[tr * TIC + tc] * wrv[u];
                        }
                        sacc += pp * wcv[v2];
                    }
                    rgc = rgc * sacc;
                }
                const float val = (row_ok && col_ok) ? rgc : 0.f;
                if (MODE == kModeFrames) {
                    if (xi < a.W && yi < a.H) a.out[(static_cast<size_t>(f) * a.HM + rM) * a.WM + cM] = val;
                } else {
                    float fq = val, pq = val;
                    if (f == f_begin) {
                        s_first[q * 64] = val;
                    } else {
                        fq = s_first[q * 64];
                        pq = s_prev[q * 64];
                    }
                    moments_push<WITH_M2>(val, f - f_begin, inv_np1, mu[q], m2[WITH_M2 ? q : 0], cc[q], fq, pq);
                    s_prev[q * 64] = val;
                }
                // one pixel at a time: keeps the finalise's live registers low (no interleaving across q)
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        __syncthreads();
    }

    if (MODE != kModeFrames && active && f_end > f_begin) {
        const size_t P = static_cast<size_t>(a.HM) * a.WM;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int yi = y0 + q;
            if (xi >= a.W || yi >= a.H) continue;
            const size_t pix = static_cast<size_t>(M * yi + py) * a.WM + cM;
            if (a.nchunk == 1) {
                const float n = static_cast<float>(f_end - f_begin);
                if (a.avg) a.avg[pix] = mu[q];
                if (a.var) a.var[pix] = WITH_M2 ? m2[WITH_M2 ? q : 0] / n : 0.f;
                if (a.tac2) a.tac2[pix] = cc[q] / (n - 1.0f);
            } else {
                float* base = a.part + static_cast<size_t>(chunk) * 5 * P + pix;
                base[0] = mu[q];
                base[P] = WITH_M2 ? m2[WITH_M2 ? q : 0] : 0.f;
                base[2 * P] = cc[q];
                base[3 * P] = s_first[q * 64];
                base[4 * P] = s_prev[q * 64];
            }
        }
    }
}

// Merges per-chunk moments in frame order (Chan for the mean/M2; the lag-1 co-moment is re-centred on the
// merged mean and gains the cross term of the two frames that meet at the chunk boundary).
__global__ __launch_bounds__(256) void moments_merge(const float* __restrict__ part, int nchunk, int fpc, int N,
                                                     size_t P, float* avg, float* var, float* tac2) {
    const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= P) return;
    float n = 0.f, mu = 0.f, m2 = 0.f, c = 0.f, first = 0.f, last = 0.f;
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
            const float e = d * expf((-d * d) / tSS);
            const float e2 = e * e;
            const float dw = e2 * e2;
            dws += dw;
            if (Gx * dx + Gy * dy < 0) {
                float Dk = fabsf(Gy * dx - Gx * dy) / sqrtf(Gx * Gx + Gy * Gy);
                if (isnan(Dk)) Dk = d;
                Dk = 1 - Dk / d;
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
    const size_t dyn = static_cast<size_t>(a.nwaves) * Geometry<K, kQTemporal>::RS * sizeof(float) +
                       (mode != kModeFrames ? 2 * static_cast<size_t>(a.nwaves) * kQTemporal * 64 * sizeof(float) : 0);
    if (mode == kModeFrames)
        hipLaunchKernelGGL((esrrf_fused<K, kQFrames, kModeFrames>), grid, block, dyn, st, a);
    else if (mode == kModeTemporal)
        hipLaunchKernelGGL((esrrf_fused<K, kQTemporal, kModeTemporal>), grid, block, dyn, st, a);
    else
        hipLaunchKernelGGL((esrrf_fused<K, kQTemporal, kModeTemporalVar>), grid, block, dyn, st, a);
    return hipGetLastError();
}

hipError_t launch_fused(const FusedArgs& a, int K, dim3 grid, dim3 block, int mode, hipStream_t st) {
    switch (K) {
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

hipError_t launch_merge(const float* part, int nchunk, int fpc, int N, size_t P, float* avg, float* var, float* tac2,
                        hipStream_t st) {
    const unsigned blocks = static_cast<unsigned>((P + 255) / 256);
    hipLaunchKernelGGL(moments_merge, dim3(blocks), dim3(256), 0, st, part, nchunk, fpc, N, P, avg, var, tac2);
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
