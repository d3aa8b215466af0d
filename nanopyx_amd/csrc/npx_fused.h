// Device body of the fused eSRRF kernels (included by npx_kernels.hip and npx_batched.hip).
//
// A block owns a 64-column x Q-row region of the input image and up to 16 wave slots; a wave slot is one
// (radius, output phase) pair, one lane per input column, Q rows per lane. Every lane of a wavefront shares
// the phase and the radius, so the ring geometry (offsets, distances, weights per window offset) is wave-uniform;
// it is staged per wave in LDS and read into VGPRs. Per frame the block stages the input tile + halo, the Roberts gradients and the
// normalised gradient lattice in LDS (the lattice is stored in two column-parity planes so that the 64 lanes
// of a read touch 64 consecutive float2). Ring sums accumulate in registers.
//
// Reference semantics: src/include/_c_sr_radial_gradient_convergence.c:17-90 (ring sum),
// src/include/_c_interpolation_catmull_rom.c:6-52 (taps), src/include/_c_gradients.c:53-72 (Roberts),
// src/nanopyx/core/transform/_le_esrrf.pyx:172-195 (composition),
// src/nanopyx/core/transform/sr_temporal_correlations.py:46-75,205-221 (temporal moments),
// src/nanopyx/core/analysis/parameter_sweep.py:33-37 (sweep: one eSRRF per (sensitivity, radius)).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "npx_kernels.h"

namespace npx {

using cfloat = const __attribute__((address_space(4))) float;  // constant address space: scalar loads

// Output-phase rows among the wave slots of one block: at most kRowSets distinct phase rows py (16 consecutive
// phases of one radius span at most 4 rows of M phases for every M; slots of several radii repeat rows).
constexpr int kRowSets = 4;

// The distinct phase rows of group `group`'s active slots, ascending, in rows[0..n); returns n.
__device__ __forceinline__ int group_rows(const FusedArgs& a, int group, int (&rows)[kRowSets]) {
    int n = 0;
    for (int w = 0; w < a.nwaves; ++w) {
        const int sw = group * a.nwaves + w, rw = sw / a.nphase;
        if (rw >= a.nrad) break;
        int py = (sw - rw * a.nphase) / a.M;
        bool seen = false;
        for (int k = 0; k < n; ++k) seen = seen || rows[k] == py;
        if (seen || n == kRowSets) continue;
        int k = n++;
        while (k > 0 && rows[k - 1] > py) {  // insertion keeps rows ascending
            rows[k] = rows[k - 1];
            --k;
        }
        rows[k] = py;
    }
    return n;
}

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

// apply_pow over N values with the (wave-uniform) mode tested once.
template <int N>
__device__ __forceinline__ void pow_all(float (&x)[N], int pow_mode, float S) {
    if (pow_mode == 2) {
#pragma unroll
        for (int k = 0; k < N; ++k) x[k] = x[k] * x[k];
    } else if (pow_mode == 3) {
#pragma unroll
        for (int k = 0; k < N; ++k) x[k] = x[k] * x[k] * x[k];
    } else if (pow_mode == 4) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const float s = x[k] * x[k];
            x[k] = s * s;
        }
    } else if (pow_mode == 1) {
#pragma unroll
        for (int k = 0; k < N; ++k)
            x[k] = x[k] > 0.f ? __builtin_amdgcn_exp2f(S * __builtin_amdgcn_logf(x[k])) : 0.f;
    }
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
    // d = v - mu, e = d / (n + 1), mu' = mu + e, a = prev - mu, b = first - mu. Re-centring C on mu' adds
    // e (a + b) + (n - 1) e^2 (the deviations of the first and of the last n - 1 values sum to -a and -b), the new
    // pair adds (a - e)(d - e); with n e = d - e the sum is a d + e (b - e) = a d + e (first - mu').
    const float delta = v - mu;
    const float e = delta * inv_np1;
    const float mun = mu + e;
    c = fmaf(prev - mu, delta, c);
    c = fmaf(e, first - mun, c);
    if (WITH_M2) m2 = fmaf(delta, v - mun, m2);
    mu = mun;
    prev = v;
}

__device__ __forceinline__ float load_px(const FusedArgs& a, int f, int gy, int gx) {
    if (gy < 0 || gy >= a.H || gx < 0 || gx >= a.W) return 0.f;
    const size_t o = (static_cast<size_t>(f) * a.H + gy) * a.W + gx;
    return a.u16 ? static_cast<float>(static_cast<const uint16_t*>(a.in)[o]) : static_cast<const float*>(a.in)[o];
}

// LDS carving of the fused kernels (one dynamic buffer; byte offsets, 16-byte aligned segments).
template <int K, int Q>
struct Lds {
    using G = Geometry<K, Q>;
    static constexpr int al(int b) { return (b + 15) / 16 * 16; }
    // lattice row pass (RX, RY: LR*TGC floats each), later in the frame the intensity row pass (kRowSets*Q*TIC)
    static constexpr int NXY = 2 * G::LR * G::TGC > kRowSets * Q * G::TIC ? 2 * G::LR * G::TGC
                                                                             : kRowSets * Q * G::TIC;
    static constexpr int IN = 0;                                 // float [TIR*TIC]
    static constexpr int RX = IN + al(4 * G::TIR * G::TIC);      // float [NXY]: RX, RY / intensity rows
    static constexpr int LAT = RX + al(4 * NXY);                 // float2 [2*LR*LP]
    static constexpr int LRW = LAT + al(8 * 2 * G::LR * G::LP);  // float4 [LR]
    static constexpr int LRB = LRW + al(16 * G::LR);             // int [LR]
    static constexpr int LCW = LRB + al(4 * G::LR);              // float4 [LB]
    static constexpr int LCB = LCW + al(16 * G::LB);             // int [LB]
    static constexpr int CRW = LCB + al(4 * G::LB);              // float4 [kRowSets*Q] intensity row taps
    static constexpr int CRB = CRW + al(16 * kRowSets * Q);      // int [kRowSets*Q]
    static constexpr int RING = CRB + al(4 * kRowSets * Q);      // float [nwaves*RW], then per-pixel state
    static_assert(2 * G::TGR * G::TGC <= 4 * G::LR * G::LP, "gradient tile must fit in the lattice storage");
};

template <int K, int Q>
constexpr size_t lds_base_bytes(int nwaves, int mode) {
    return Lds<K, Q>::RING + static_cast<size_t>(nwaves) * Geometry<K, Q>::RW * 4 +
           ((mode == kModeTemporal || mode == kModeTemporalVar) ? 2 * static_cast<size_t>(nwaves) * Q * 64 * 4 : 0);
}

// Whether the intensity row pass gets LDS of its own (decided for the largest block, 16 waves).
template <int K, int Q>
constexpr bool t_own(int mode) {
    return lds_base_bytes<K, Q>(16, mode) + static_cast<size_t>(kRowSets) * Q * Geometry<K, Q>::TIC * 4 <=
           static_cast<size_t>(kLdsLimit);
}

template <int K, int Q>
inline size_t lds_bytes_kq(int nwaves, int mode) {
    return lds_base_bytes<K, Q>(nwaves, mode) +
           (t_own<K, Q>(mode) ? static_cast<size_t>(kRowSets) * Q * Geometry<K, Q>::TIC * 4 : 0);
}

#ifndef NPX_RING_GROUPS
#define NPX_RING_GROUPS 1
#endif

// Window rows [J0, J1) of the ring sums for column i whose offset dx has sign SG (+1, -1, or 0 for dx = 0): constants
// of those rows in VGPRs, lattice rows streamed. rs = 1 / |dx| (1 where dx = 0).
template <int K, int KL, int Q, int J0, int J1, int SG>
__device__ __forceinline__ void ring_rows(float (&acc)[Q], const float2* col, float rs, const float* rl, int i) {
    using GL = Geometry<KL, Q>;
    constexpr int NW = 2 * K + 1;
    constexpr int LP = GL::LP;
    constexpr int NJ = J1 - J0;
    constexpr int R0 = J0, R1 = 2 * (Q - 1) + J1;  // lattice rows these window rows read: [R0, R1)
    constexpr int RDT = 2 * NW;
    float cj[NJ], tdv[NJ], tsv[NJ];
    const float2* dt = reinterpret_cast<const float2*>(rl + RDT) + i * NW + J0;
#pragma unroll
    for (int jj = 0; jj < NJ; ++jj) {
        cj[jj] = SG == 0 ? rl[J0 + jj] : rl[J0 + jj] * rs;  // -dy / |dx| (-dy where dx = 0)
        const float2 e = dt[jj];
        tdv[jj] = e.x;
        tsv[jj] = e.y;
    }
    // lattice rows are read two ahead of their use so the LDS latency overlaps the ring arithmetic
    float2 g1 = col[R0 * LP];
    float2 g2 = col[(R0 + 1 < R1 ? R0 + 1 : R0) * LP];
#pragma unroll
    for (int r = R0; r < R1; ++r) {
        const float2 g = g1;  // s g, g = G / |G|
        g1 = g2;
        if (r + 2 < R1) g2 = col[(r + 2) * LP];
#pragma unroll
        for (int jj = 0; jj < NJ; ++jj) {
            const int rq = r - (J0 + jj);
            if (rq < 0 || (rq & 1) || (rq >> 1) >= Q) continue;
            const int q = rq >> 1;
            // z = -s u / |dx| and vs = s v / |dx| (u = g.d, v = g x d): sgn(dx) enters as the sign of one operand
            float z, vs;
            if constexpr (SG > 0) {
                z = fmaf(g.y, cj[jj], -g.x);
                vs = fmaf(g.x, cj[jj], g.y);
            } else if constexpr (SG < 0) {
                z = fmaf(g.y, cj[jj], g.x);
                vs = fmaf(g.x, cj[jj], -g.y);
            } else {
                z = g.y * cj[jj];
                vs = g.x * cj[jj];
            }
            const float c = __builtin_amdgcn_fmed3f(z, 0.f, 1.f);  // clamp(-s u / |dx|, 0, 1)
            const float x = fmaf(fabsf(vs), tsv[jj], tdv[jj]);      // t (d - |v|)
            acc[q] = fmaf(x, c, acc[q]);
        }
    }
}

// One window column of the ring sums, its rows taken in NPX_RING_GROUPS groups so that only one group's constants
// are live at a time (scheduling barriers keep each group's loads inside the group).
template <int K, int KL, int Q, int SG>
__device__ __forceinline__ void ring_col(float (&acc)[Q], const float2* col, float rs, const float* rl, int i) {
    constexpr int NW = 2 * K + 1;
    constexpr int NG = NPX_RING_GROUPS < NW ? NPX_RING_GROUPS : NW;
    constexpr int JA = (NW + NG - 1) / NG;  // rows per group (the last group may be shorter)
    ring_rows<K, KL, Q, 0, (JA < NW ? JA : NW), SG>(acc, col, rs, rl, i);
    if constexpr (JA < NW) {
        __builtin_amdgcn_sched_barrier(0);
        ring_rows<K, KL, Q, JA, (2 * JA < NW ? 2 * JA : NW), SG>(acc, col, rs, rl, i);
    }
    if constexpr (2 * JA < NW) {
        __builtin_amdgcn_sched_barrier(0);
        ring_rows<K, KL, Q, 2 * JA, (3 * JA < NW ? 3 * JA : NW), SG>(acc, col, rs, rl, i);
    }
    if constexpr (3 * JA < NW) {
        __builtin_amdgcn_sched_barrier(0);
        ring_rows<K, KL, Q, 3 * JA, NW, SG>(acc, col, rs, rl, i);
    }
}

// Ring sums of one wave (radius half-width K) over a lattice built for half-width KL >= K. Sample (j, i) of
// pixel q reads lattice row 2q + ey + j + (KL - K), column 2*lane + ex + i + (KL - K). The lattice holds s g with
// g = G/|G| and s = kRingScale. With u = g.(dx, dy) and v = g x (dx, dy) (d = |(dx, dy)|, t = w/d), the
// reference's contribution [G.d < 0] * w * (1 - |G x d| / (|G| d)) is [u < 0] * t * (d - |v|). Both u and v are
// taken divided by |dx| (a positive factor of the column, 1 where dx = 0), which leaves the sign of u and, with the
// factor folded into the constant of |v|, the term unchanged: -s u / |dx| = s g.y c_j - sgn(dx) s g.x and
// s v / |dx| = s g.x c_j + sgn(dx) s g.y with c_j = -dy_j / |dx|, each one FMA of the lattice value with a
// per-sample constant, so a sample costs four FMAs and a lattice point no products of its own. The indicator is
// taken as clamp(-s u / |dx|, 0, 1), which the compiler folds into the clamp bit of the FMA: it equals [u < 0] for
// |u| >= |dx| / s, and below that the term it scales is itself below t u^2 / d (u^2 + v^2 = d^2, so
// d - |v| <= u^2 / d), under 3e-9 t / d at s = 2^16 and |dx| <= 3.5. The ring constants come from the wave's LDS copy,
// so they sit in VGPRs (on gfx950 this four-FMA sample issues about a quarter slower with its constants in SGPRs than
// with every operand a VGPR). The window columns with dx < 0, dx = 0 and dx > 0 are the ranges [0, ineg),
// [ineg, inz) and [inz, NW) (dx grows with the column), each with its own loop body.
template <int K, int KL, int Q>
__device__ __forceinline__ void ring_sum(float (&acc)[Q], const float2* s_lat, int ey, int ex, int lane,
                                         const float* rl, int ineg, int inz) {
    using GL = Geometry<KL, Q>;
    constexpr int NW = 2 * K + 1;
    constexpr int LP = GL::LP, LPL = GL::LR * GL::LP;
    constexpr int DK = KL - K;
    constexpr int RDX = NW;
    static_assert(K <= KL, "ring wider than the lattice");
#pragma unroll
    for (int q = 0; q < Q; ++q) acc[q] = 0.f;
    // Wave priority falls as the wave moves through the window columns (3 for the first third, then 2, 1, and 0 for
    // the last column). The SIMD issues from the highest-priority ready wave first and otherwise from the oldest, so
    // without this the oldest of its waves runs ahead and the youngest finishes the ring sums alone at a fraction of
    // the issue rate; with it the waves that are behind are preferred and the waves of a SIMD finish close together.
    constexpr int P2 = NW / 3, P1 = 2 * NW / 3, P0 = NW - 1;
    auto prio = [&](int i) {
        if (i == P2) __builtin_amdgcn_s_setprio(2);
        else if (i == P1) __builtin_amdgcn_s_setprio(1);
        else if (i == P0) __builtin_amdgcn_s_setprio(0);
    };
    auto column = [&](int i) {
        const int c = ex + i + DK;
        return s_lat + (c & 1) * LPL + (ey + DK) * LP + lane + (c >> 1);
    };
    __builtin_amdgcn_s_setprio(3);
#pragma unroll 1
    for (int i = 0; i < ineg; ++i) {
        prio(i);
        ring_col<K, KL, Q, -1>(acc, column(i), -rl[RDX + i], rl, i);  // rdx = 1 / dx < 0
    }
#pragma unroll 1
    for (int i = ineg; i < inz; ++i) {
        prio(i);
        ring_col<K, KL, Q, 0>(acc, column(i), 1.f, rl, i);
    }
#pragma unroll 1
    for (int i = inz; i < NW; ++i) {
        prio(i);
        ring_col<K, KL, Q, 1>(acc, column(i), rl[RDX + i], rl, i);
    }
}

// Dispatch on the wave's (uniform) half-width for lattices of half-width KL.
template <int KL, int Q>
__device__ __forceinline__ void ring_sum_any(int K, float (&acc)[Q], const float2* s_lat, int ey, int ex, int lane,
                                             const float* rg, int ineg, int inz) {
    switch (K) {
#define NPX_RING_CASE(k)                                                             \
    case k:                                                                          \
        if constexpr (k <= KL) ring_sum<(k <= KL ? k : KL), KL, Q>(acc, s_lat, ey, ex, lane, rg, ineg, inz); \
        break;
        NPX_RING_CASE(0)
        NPX_RING_CASE(1)
        NPX_RING_CASE(2)
        NPX_RING_CASE(3)
        NPX_RING_CASE(4)
        NPX_RING_CASE(5)
        NPX_RING_CASE(6)
#undef NPX_RING_CASE
        default: break;
    }
}

// One block of the fused computation: region (bx, by) of 64 input columns x Q input rows, wave-slot group
// `group`, frames of chunk `chunk`. Steps per frame: (1) input tile + halo, (2) Roberts cross, (3a/3b)
// Catmull-Rom x2M lattice along rows then columns + normalisation, (4) ring sums, (5) normalise /
// sensitivity / intensity weight and store (frames) or fold into the moments (temporal) or the per-
// sensitivity means (sweep). VARK: waves may have ring half-widths K < KL (batched launches); otherwise
// every wave has K = KL, fixed at compile time.
template <int KL, int Q, int MODE, bool VARK>
__device__ __forceinline__ void fused_block(const FusedArgs& a, int bx, int by, int group, int chunk,
                                            unsigned char* smem) {
    using G = Geometry<KL, Q>;
    using L = Lds<KL, Q>;
    constexpr int RCL = G::RC, RWL = G::RW;
    constexpr int HG = G::HG, HI = G::HI;
    constexpr int TIR = G::TIR, TIC = G::TIC;
    constexpr int TGR = G::TGR, TGC = G::TGC;
    constexpr int LR = G::LR, LP = G::LP, LPL = LR * LP, LB = G::LB;
    constexpr bool TEMPORAL = MODE == kModeTemporal || MODE == kModeTemporalVar;
    constexpr bool SWEEP = MODE == kModeSweep;
    constexpr bool WITH_M2 = MODE == kModeTemporalVar;
    constexpr bool ARRAYS = MODE == kModeArrays;                // lattice and intensity from the caller's arrays
    constexpr bool FRAMES = MODE == kModeFrames || ARRAYS;      // per-frame maps out
    constexpr int NSM = SWEEP ? kMaxSens : 1;

    float* const s_in = reinterpret_cast<float*>(smem + L::IN);
    float* const s_rx = reinterpret_cast<float*>(smem + L::RX);
    float2* const s_rxy = reinterpret_cast<float2*>(s_rx);  // row pass of both gradients [LR][TGC]
    // intensity row pass [set][q][TIC]: own storage after the per-pixel state where the LDS budget allows (then it
    // is computed beside Roberts), else the lattice-row scratch once the lattice is built
    constexpr bool T_OWN = t_own<KL, Q>(MODE);
    float* const s_irow = T_OWN ? reinterpret_cast<float*>(smem + L::RING) + a.nwaves * RWL +
                                      (TEMPORAL ? 2 * a.nwaves * Q * 64 : 0)
                                : s_rx;
    float2* const s_lat = reinterpret_cast<float2*>(smem + L::LAT);
    float2* const s_g = s_lat;  // Roberts gradients (col, row) share the lattice storage
    float4* const s_lrw = reinterpret_cast<float4*>(smem + L::LRW);
    int* const s_lrb = reinterpret_cast<int*>(smem + L::LRB);
    float4* const s_lcw = reinterpret_cast<float4*>(smem + L::LCW);
    int* const s_lcb = reinterpret_cast<int*>(smem + L::LCB);
    float4* const s_crw = reinterpret_cast<float4*>(smem + L::CRW);
    int* const s_crb = reinterpret_cast<int*>(smem + L::CRB);
    float* const s_ring = reinterpret_cast<float*>(smem + L::RING);

    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const int x0 = bx * 64;
    const int y0 = by * Q;
    const int f_begin = a.f0 + chunk * a.fpc;
    const int f_end = min(a.N, f_begin + a.fpc);
    const int slot = __builtin_amdgcn_readfirstlane(group * a.nwaves + wave);
    const int rad = slot / a.nphase;
    const int ph = slot - rad * a.nphase;
    const bool active = rad < a.nrad;
    const int M = a.M;
    const int py = active ? ph / M : 0;
    const int px = active ? ph % M : 0;
    const int ey = (2 * py + 1) / M;
    const int ex = (2 * px + 1) / M;
    const int xi = x0 + lane;
    const int cM = M * xi + px;
    const int a_lo = 2 * y0 - KL;
    const int b_lo = 2 * x0 - KL;
    const int Kw = VARK ? (active ? a.K[rad] : KL) : KL;
    const int NWw = 2 * Kw + 1;
    const int margin = active ? a.margin[rad] : 0;

    // ---- stage the frame-invariant tables. Per wave (block of RWL floats): the ring constants of its radius and
    // phase in the form the ring sums use (-dy[j], 1 / dx[i] (0 for dx = 0), (t d, -t |dx| / s)[i][j] with
    // s = kRingScale and |dx| taken as 1 for dx = 0), then the normaliser part (wsum, then the (NW+1)^2 prefix sums
    // of w). The geometry is the reference's for this block's pixels: it takes each output pixel's centre as
    // float32((x M + 0.5) / M) (_c_sr_radial_gradient_convergence.c:21-22), and the window offsets, distances and
    // weights from that rounded centre (:31-44, :4-6, :63). Every pixel of a phase in the block has the same rounded
    // fraction (columns x0 .. x0 + 63 lie in one binade for x0 >= 64, rows y0 .. y0 + Q - 1 for y0 >= Q), so the
    // offsets are those of the block's last column and row (in the first block the other pixels' centres, below 64,
    // differ by less than 2^-18). One work item per sample writes its two ring constants and its weight into a grid
    // (the lattice-row area, unused before the first frame); after a barrier the row prefix sums of each grid go to
    // the lattice storage (double), after another the column sums of those give the prefix sums.
    constexpr int NWL = 2 * KL + 1, WG = NWL * NWL;
    float* const s_wgrid = s_rx;                                      // [wave][WG] weights w[j][i], j outer
    double* const s_wrow = reinterpret_cast<double*>(s_lat);          // [wave][WG] sum over i' < i + 1 of w[j][i']
    static_assert(16 * WG <= L::NXY, "weight grids must fit the lattice-row area");
    static_assert(2 * 16 * WG <= 4 * LR * LP, "row prefix sums must fit the lattice storage");
    const int xr = min(x0 + 63, a.W - 1), yr = min(y0 + Q - 1, a.H - 1);
    // offset o (0 .. 2K) of the window of half-width K around pixel r, output phase p, along rows or columns
    auto offset = [&](int r, int p, int K, int o) {
        const float c = static_cast<float>((static_cast<double>(M * r + p) + 0.5) / M);
        const float v = static_cast<float>(static_cast<int>(2.0f * c) + (o - K)) / 2.0f;
        return v - c;
    };
    for (int k = tid; k < a.nwaves * RCL; k += nthr) {  // dy, 1 / dx and the padding (pairs below)
        const int w = k / RCL, e = k - w * RCL;
        const int sw = group * a.nwaves + w, rw = sw / a.nphase, pw = sw - rw * a.nphase;
        const int Kr = rw < a.nrad ? a.K[rw] : 0, nw = rw < a.nrad ? 2 * Kr + 1 : 0;
        if (e >= 2 * nw && e < 2 * nw + 2 * nw * nw) continue;
        float v = 0.f;
        if (e < nw) {
            v = -offset(yr, pw / M, Kr, e);  // -dy
        } else if (e < 2 * nw) {
            const float dx = offset(xr, pw % M, Kr, e - nw);
            v = dx != 0.f ? 1.0f / dx : 0.f;  // 1 / dx, 0 where dx = 0
        }
        s_ring[w * RWL + e] = v;
    }
    for (int k = tid; k < a.nwaves * WG; k += nthr) {  // sample (i, j): pair i nw + j
        const int w = k / WG, q = k - w * WG;
        const int sw = group * a.nwaves + w, rw = sw / a.nphase, pw = sw - rw * a.nphase;
        if (rw >= a.nrad) continue;
        const int Kr = a.K[rw], nw = 2 * Kr + 1;
        if (q >= nw * nw) continue;
        const int i = q / nw, j = q - i * nw;
        const float dx = offset(xr, pw % M, Kr, i), dy = offset(yr, pw / M, Kr, j);
        const float dd = __fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy));
        const float d = static_cast<float>(sqrt(static_cast<double>(dd)));
        float wt = 0.f, t = 0.f;
        if (d != 0.f && d <= a.tso[rw]) {  // :44
            const double dn = d;
            const double y = dn * exp((-dn * dn) / static_cast<double>(a.tss[rw]));
            const double y2 = y * y;
            wt = static_cast<float>(y2 * y2);  // the fourth power (:5)
            t = static_cast<float>(static_cast<double>(wt) / dn);
        }
        const float dxa = fabsf(dx);
        float* const pr = s_ring + w * RWL + 2 * nw + 2 * q;
        pr[0] = t * d;                                                 // t d
        pr[1] = t * (-1.f / kRingScale) * (dxa != 0.f ? dxa : 1.f);  // -t |dx| / s
        s_wgrid[w * WG + j * nw + i] = wt;
    }
    for (int la = tid; la < LR; la += nthr) {
        const int aa = a_lo + la;
        float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
        int b = 0;
        if (aa >= 1 && aa <= 2 * a.HM - 2) {  // rows a >= 2H read the clamped last row (:45-52)
            const int ai = min(aa, 2 * a.H);
            w = a.lrw[ai];
            b = a.lrb[ai] - (y0 - HG);
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
            w = a.lcw[bi];
            b = a.lcb[bi] - (x0 - HG);
        }
        s_lcw[lb] = w;
        s_lcb[lb] = b;
    }
    __syncthreads();  // the weight grids are complete
    for (int k = tid; k < a.nwaves * NWL; k += nthr) {  // row j of each grid: prefix sums along i
        const int w = k / NWL, j = k - w * NWL;
        const int rw = (group * a.nwaves + w) / a.nphase;
        const int nw = rw < a.nrad ? 2 * a.K[rw] + 1 : 0;
        if (j >= nw) continue;
        double acc = 0.0;
        for (int i = 0; i < nw; ++i) {
            acc += s_wgrid[w * WG + j * nw + i];
            s_wrow[w * WG + j * nw + i] = acc;
        }
    }
    __syncthreads();  // the row prefix sums are complete
    // wsum (float, j outer and i inner, :30-35,64) and the prefix sums P[jj][ii] = sum over j < jj, i < ii of w
    for (int k = tid; k < a.nwaves * G::RN; k += nthr) {
        const int w = k / G::RN, e = k - w * G::RN;
        const int rw = (group * a.nwaves + w) / a.nphase;
        const int nw = rw < a.nrad ? 2 * a.K[rw] + 1 : 0;
        float* const out = s_ring + w * RWL + RCL;
        if (e == 0) {
            float v = 0.f;
            for (int q = 0; q < nw * nw; ++q) v += s_wgrid[w * WG + q];
            out[0] = v;
        } else if (e - 1 < nw + 1) {  // column ii: P[jj][ii] for every jj
            const int ii = e - 1;
            double acc = 0.0;
            out[1 + ii] = 0.f;
            for (int jj = 1; jj <= nw; ++jj) {
                if (ii > 0) acc += s_wrow[w * WG + (jj - 1) * nw + ii - 1];
                out[1 + jj * (nw + 1) + ii] = static_cast<float>(acc);
            }
        } else if (e - 1 >= (nw + 1) * (nw + 1)) {
            out[e] = 0.f;  // padding
        }
    }

    // intensity Catmull-Rom row taps of every (phase row set, q); the wave's set is the index of its phase row
    int rows[kRowSets];
    const int nrows = group_rows(a, group, rows);
    for (int k = tid; k < nrows * Q; k += nthr) {
        const int rs = k / Q, q = k - rs * Q;
        const int rM = min(M * (y0 + q) + rows[rs], a.HM - 1);
        s_crw[k] = a.crw[rM];
        s_crb[k] = a.crb[rM] - (y0 - HI);
    }
    int rset = 0;
    for (int k = 1; k < nrows; ++k) rset += rows[k] <= py ? 1 : 0;
    const int cMs = min(cM, a.WM - 1);
    const float4 wcol = a.ccw[cMs];
    const int ccol = a.ccb[cMs] - (x0 - HI);

    // per-pixel state: temporal moments (mean, lag-1 co-moment, M2 if VAR) or one mean per sensitivity
    float mu[SWEEP ? NSM * Q : Q], m2[WITH_M2 ? Q : 1], cc[TEMPORAL ? Q : 1];
#pragma unroll
    for (int q = 0; q < (SWEEP ? NSM * Q : Q); ++q) mu[q] = 0.f;
#pragma unroll
    for (int q = 0; q < (WITH_M2 ? Q : 1); ++q) m2[q] = 0.f;
#pragma unroll
    for (int q = 0; q < (TEMPORAL ? Q : 1); ++q) cc[q] = 0.f;
    // first and previous frame's value of every pixel of the wave: [wave][q][lane] and [nwaves + wave][q][lane]
    float* const s_first = s_ring + a.nwaves * RWL + (wave * Q) * 64 + lane;
    float* const s_prev = s_first + a.nwaves * Q * 64;

    // frame slices of a served job resume the moments folded so far
    if constexpr (TEMPORAL) {
        if (a.state && a.n0 > 0 && active) {
            const size_t P = static_cast<size_t>(a.HM) * a.WM;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int yi = y0 + q;
                if (xi >= a.W || yi >= a.H) continue;
                const size_t pix = static_cast<size_t>(M * yi + py) * a.WM + cM;
                mu[q] = a.state[pix];
                if (WITH_M2) m2[WITH_M2 ? q : 0] = a.state[P + pix];
                cc[q] = a.state[2 * P + pix];
                s_first[q * 64] = a.state[3 * P + pix];
                s_prev[q * 64] = a.state[4 * P + pix];
            }
        }
    }
    const int n0 = TEMPORAL && a.state ? a.n0 : 0;

    // Input tile + halo of one frame (zero outside the image; uint16 converted as a plain cast). The first
    // kPrefetch elements of each thread come through registers loaded a frame ahead. The tile of frame f + 1 is
    // stored while frame f's ring sums run (its last reader, the intensity row pass, is behind a barrier by then).
    float pf[kPrefetch];
    // in-frame element offsets of the prefetched tile elements (-1: outside the image or the tile), fixed for the
    // block; a frame's loads add them to the frame's base address (one VGPR per element instead of a 64-bit address)
    int pf_off[kPrefetch];
#pragma unroll
    for (int k = 0; k < kPrefetch; ++k) {
        const int idx = tid + k * nthr;
        const int r = idx / TIC, c = idx - r * TIC;
        const int gy = y0 - HI + r, gx = x0 - HI + c;
        pf_off[k] = (idx < TIR * TIC && gy >= 0 && gy < a.H && gx >= 0 && gx < a.W) ? gy * a.W + gx : -1;
    }
    auto fetch = [&](int f) {
        const size_t fb = static_cast<size_t>(f) * a.H * a.W;
        if (a.u16) {
            const uint16_t* base = static_cast<const uint16_t*>(a.in) + fb;
#pragma unroll
            for (int k = 0; k < kPrefetch; ++k) pf[k] = pf_off[k] >= 0 ? static_cast<float>(base[pf_off[k]]) : 0.f;
        } else {
            const float* base = static_cast<const float*>(a.in) + fb;
#pragma unroll
            for (int k = 0; k < kPrefetch; ++k) pf[k] = pf_off[k] >= 0 ? base[pf_off[k]] : 0.f;
        }
    };
    auto store_tile = [&](int f) {
#pragma unroll
        for (int k = 0; k < kPrefetch; ++k) {
            const int idx = tid + k * nthr;
            if (idx < TIR * TIC) s_in[idx] = pf[k];
        }
        for (int idx = tid + kPrefetch * nthr; idx < TIR * TIC; idx += nthr) {
            const int r = idx / TIC, c = idx - r * TIC;
            s_in[idx] = load_px(a, f, y0 - HI + r, x0 - HI + c);
        }
    };
    if (!ARRAYS && f_begin < f_end) {
        fetch(f_begin);
        store_tile(f_begin);
        if (f_begin + 1 < f_end) fetch(f_begin + 1);
    }
    __syncthreads();
    const float* ringC = s_ring + wave * RWL;  // ring constants of this wave
    const float* ringL = ringC + RCL;          // wsum, prefix sums
    // the wave's window columns with dx < 0 are [0, ring_ineg), with dx = 0 [ring_ineg, ring_inz)
    int ring_ineg = 0, ring_inz = 0;
    for (int i = 0; i < NWw; ++i) {
        const float rdx = ringC[NWw + i];
        ring_ineg += rdx < 0.f ? 1 : 0;
        ring_inz += rdx <= 0.f ? 1 : 0;
    }
    ring_ineg = __builtin_amdgcn_readfirstlane(ring_ineg);
    ring_inz = __builtin_amdgcn_readfirstlane(ring_inz);
    // 1 / sum(w) of the wave's ring (frame-invariant): interior pixels multiply by it
    const float wsum_inv = __int_as_float(
        __builtin_amdgcn_readfirstlane(__float_as_int(ringL[0] > 0.f ? 1.0f / ringL[0] : 0.f)));

    // intensity Catmull-Rom row pass, shared by the waves of one phase row: T[set][q][c] = sum_u w_u tile[b+u][c]
    // (the 4 x 4 taps lie inside the tile for every lane: tap base floor(r - 0.5) - 1 with r in [y0, y0 + Q),
    // halo HI >= 2)
    auto intensity_rows = [&]() {
        for (int idx = tid; idx < nrows * Q * TIC; idx += nthr) {
            const int k = idx / TIC, c = idx - k * TIC;
            const float4 wr = s_crw[k];
            const float* col = s_in + s_crb[k] * TIC + c;
            const float wrv[4] = {wr.x, wr.y, wr.z, wr.w};
            float pp = 0.f;
#pragma unroll
            for (int u = 0; u < 4; ++u) pp += col[u * TIC] * wrv[u];
            s_irow[idx] = pp;
        }
    };

    // arrays mode: lattice point (a, b) is the caller's gradient at row min(a M, 2 HM - 1), column min(b M, 2 WM - 1)
    // of its x2M arrays (or row min(a, 2H), column min(b, 2W) of a compact lattice), normalised and scaled by
    // s = kRingScale (_c_sr_radial_gradient_convergence.c:45-52 reads those positions); points outside
    // a in [1, 2 HM - 2], b in [1, 2 WM - 2] are outside every window and hold 0. The first kLatPf points of each
    // thread are gathered a frame ahead into registers, so the gather of frame f + 1 is in flight during frame f's
    // ring sums (the Q = 16 bodies and the widest rings have no registers to spare and gather in place).
    constexpr int kLatPf = ARRAYS && Q <= 8 && KL <= 4 ? 4 : 0;
    float2 lpf[kLatPf > 0 ? kLatPf : 1];
    auto lat_load = [&](int f, int idx) {
        const int plane = idx / LPL;
        const int rem = idx - plane * LPL;
        const int la = rem / LP;
        const int pos = rem - la * LP;
        const int aa = a_lo + la, bb = b_lo + 2 * pos + plane;
        float2 v = make_float2(0.f, 0.f);
        if (aa >= 1 && aa <= 2 * a.HM - 2 && bb >= 1 && bb <= 2 * a.WM - 2) {
            const size_t o = a.alat ? (static_cast<size_t>(f) * (2 * a.H + 1) + min(aa, 2 * a.H)) * (2 * a.W + 1) +
                                          min(bb, 2 * a.W)
                                    : (static_cast<size_t>(f) * (2 * a.HM) + min(aa * M, 2 * a.HM - 1)) * (2 * a.WM) +
                                          min(bb * M, 2 * a.WM - 1);
            v = make_float2(a.agx[o], a.agy[o]);
        }
        return v;
    };
    // the prefetched and the in-place points must round alike whatever the block size: the sum of squares is
    // written with explicit roundings so the compiler cannot contract the two copies differently
    auto lat_norm = [](float2 v) {
        const float n2 = __fmaf_rn(v.x, v.x, __fmul_rn(v.y, v.y));
        if (!(n2 > 0.f)) return make_float2(0.f, 0.f);
        const float inv = rsqrtf(n2) * kRingScale;  // s / |G|
        return make_float2(v.x * inv, v.y * inv);
    };
    auto lat_fetch = [&](int f) {
#pragma unroll
        for (int k = 0; k < kLatPf; ++k) {
            const int idx = tid + k * nthr;
            lpf[k] = idx < 2 * LPL ? lat_load(f, idx) : make_float2(0.f, 0.f);
        }
    };
    if (ARRAYS && f_begin < f_end) lat_fetch(f_begin);

    for (int f = f_begin; f < f_end; ++f) {
        // intensity rows with their own storage are computed here, beside Roberts (both only read the tile)
        if (!ARRAYS && T_OWN && a.iw) intensity_rows();
        if constexpr (ARRAYS) {
            // the first kLatPf lattice points of each thread arrived in registers during the previous frame
#pragma unroll
            for (int k = 0; k < kLatPf; ++k) {
                const int idx = tid + k * nthr;
                if (idx < 2 * LPL) s_lat[idx] = lat_norm(lpf[k]);
            }
            for (int idx = tid + kLatPf * nthr; idx < 2 * LPL; idx += nthr) s_lat[idx] = lat_norm(lat_load(f, idx));
            __syncthreads();
            if (f + 1 < f_end) lat_fetch(f + 1);  // in flight during this frame's ring sums
        } else {
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
                s_g[idx] = make_float2(gcv, grv);
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
                    const float2 gv = s_g[tr * TGC + c];
                    sx += gv.x * wrv[u];
                    sy += gv.y * wrv[u];
                }
                s_rxy[idx] = make_float2(sx, sy);
            }
            __syncthreads();
            // (3b) ... then along columns, normalised: lattice point (a, b) holds s G/|G| (s = kRingScale) at (a/2, b/2); stored in two
            // column-parity planes so the 64 lanes of one read touch 64 consecutive float2. A thread keeps one
            // lattice column (its taps and weights) and takes every rg_n-th lattice row of it.
            {
                constexpr int NCOL = 2 * LP;  // (plane, position) columns of the two planes
                const int rg_n = max(1, nthr / NCOL);
                for (int t = tid; t < NCOL * rg_n; t += nthr) {
                    const int col = t % NCOL, rg = t / NCOL;
                    const int plane = col >= LP ? 1 : 0;
                    const int pos = col - plane * LP;
                    const int lb = 2 * pos + plane;
                    const float4 wc = s_lcw[lb];
                    const int cb = s_lcb[lb];
                    const float wcv[4] = {wc.x, wc.y, wc.z, wc.w};
                    int tcv[4];
#pragma unroll
                    for (int v = 0; v < 4; ++v) tcv[v] = min(max(cb + v, 0), TGC - 1);
                    for (int la = rg; la < LR; la += rg_n) {
                        const float2* row = s_rxy + la * TGC;
                        // the four taps are read before the first is used, so their LDS reads overlap
                        float2 rv[4];
#pragma unroll
                        for (int v = 0; v < 4; ++v) rv[v] = row[tcv[v]];
                        float gx = 0.f, gy = 0.f;
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            gx += rv[v].x * wcv[v];
                            gy += rv[v].y * wcv[v];
                        }
                        const float n2 = gx * gx + gy * gy;
                        float2 g = make_float2(0.f, 0.f);
                        if (n2 > 0.f) {
                            const float inv = rsqrtf(n2) * kRingScale;  // s / |G|
                            g = make_float2(gx * inv, gy * inv);
                        }
                        s_lat[plane * LPL + la * LP + pos] = g;
                    }
                }
            }
            __syncthreads();
        }
        // otherwise they go to the lattice-row scratch, free from here on
        if (!ARRAYS && !T_OWN && a.iw) {
            intensity_rows();
            __syncthreads();
        }
        // the next frame's tile: the tile of frame f has no reader left
        if (!ARRAYS && f + 1 < f_end) {
            store_tile(f + 1);
            if (f + 2 < f_end) fetch(f + 2);
        }

        if (active) {
            // (4) ring sums
            float acc[Q];
            // arrays mode: the caller's intensity weights of the lane's pixels, loaded before the ring sums so that
            // their latency hides behind them (the Q = 16 bodies have no registers to spare and load them after)
            constexpr bool IW_EARLY = ARRAYS && Q <= 8;
            float aiw[IW_EARLY ? Q : 1];
            if constexpr (IW_EARLY) {
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    const int rM = min(M * (y0 + q) + py, a.HM - 1);
                    aiw[q] = a.iw ? a.aimg[(static_cast<size_t>(f) * a.HM + rM) * a.WM + cMs] : 0.f;
                }
            }
            if constexpr (VARK) ring_sum_any<KL, Q>(Kw, acc, s_lat, ey, ex, lane, ringC, ring_ineg, ring_inz);
            else ring_sum<KL, KL, Q>(acc, s_lat, ey, ex, lane, ringC, ring_ineg, ring_inz);
            // (5) normalise, sensitivity, intensity weight; then store or fold into the moments. Each step runs over
            // the lane's Q pixels, so the wave-uniform decisions (power, weighting, first frame) are taken once per
            // frame; pixels whose window crosses the border take their normaliser from the prefix-sum table.
            const float wcv[4] = {wcol.x, wcol.y, wcol.z, wcol.w};
            const int nabs = n0 + (f - f_begin);  // values already folded into the moments
            const float inv_np1 = 1.0f / static_cast<float>(nabs + 1);
            const float* P = ringL + 1;
            const int b0 = 2 * xi + ex;
            const int i0 = max(0, 1 - (b0 - Kw));
            const int i1 = max(0, min(NWw, 2 * a.WM - 2 - (b0 - Kw) + 1));  // lanes past the image: empty range
            const bool col_ok = xi < a.W && cM >= margin && cM < a.WM - margin;
            auto pixel_ok = [&](int q) {
                const int yi = y0 + q, rM = M * yi + py;
                return col_ok && yi < a.H && rM >= margin && rM < a.HM - margin;
            };
            // (5a) RGC / sum(w) (:77-81) as RGC * (1 / sum(w)): within 2 ulp of the division
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int a0 = 2 * (y0 + q) + ey;
                const int j0 = max(0, 1 - (a0 - Kw));
                const int j1 = max(0, min(NWw, 2 * a.HM - 2 - (a0 - Kw) + 1));
                float den_inv = wsum_inv;
                if (j0 > 0 || j1 < NWw || i0 > 0 || i1 < NWw) {
                    const int jl = min(j0, j1), il = min(i0, i1);
                    const float den = (P[j1 * (NWw + 1) + i1] - P[jl * (NWw + 1) + i1]) -
                                      (P[j1 * (NWw + 1) + il] - P[jl * (NWw + 1) + il]);
                    den_inv = den > 0.f ? __builtin_amdgcn_rcpf(den) : 0.f;
                }
                acc[q] *= den_inv;
            }
            if constexpr (SWEEP) {
                // every sensitivity re-applies power, weight and its running mean to the same ring sums; the
                // sensitivities and their power modes are read together (one scalar load of the adjacent fields).
                // Pixels outside the image or the margin enter as 0 (their ring sum and weight set to 0 once per
                // frame), and v^2 is formed once for the integer powers: v^3 = v^2 v and v^4 = v^2 v^2, the products
                // the per-sensitivity power forms.
                float sv[NSM];
                int sp[NSM];
#pragma unroll
                for (int s = 0; s < NSM; ++s) {
                    sv[s] = a.sens[s];
                    sp[s] = a.sens_pow[s];
                }
                const int ns = a.ns;
                float inten[Q], sq[Q];
                // the intensity row values of all Q pixels are read before the first is used (their LDS reads overlap)
                float tv[Q][4];
                {
                    const float* rowp = s_irow + rset * Q * TIC + ccol;
#pragma unroll
                    for (int q = 0; q < Q; ++q)
#pragma unroll
                        for (int v2 = 0; v2 < 4; ++v2) tv[q][v2] = rowp[q * TIC + v2];
                }
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    float sacc = 0.f;
#pragma unroll
                    for (int v2 = 0; v2 < 4; ++v2) sacc += tv[q][v2] * wcv[v2];
                    const bool ok = pixel_ok(q);
                    inten[q] = ok ? (a.iw ? sacc : 1.f) : 0.f;
                    acc[q] = ok ? acc[q] : 0.f;
                    sq[q] = acc[q] * acc[q];
                }
#pragma unroll
                for (int s = 0; s < NSM; ++s) {
                    if (s < ns) {
                        float val[Q];
                        const int pm = sp[s];
                        if (pm == 2) {
#pragma unroll
                            for (int q = 0; q < Q; ++q) val[q] = sq[q];
                        } else if (pm == 3) {
#pragma unroll
                            for (int q = 0; q < Q; ++q) val[q] = sq[q] * acc[q];
                        } else if (pm == 4) {
#pragma unroll
                            for (int q = 0; q < Q; ++q) val[q] = sq[q] * sq[q];
                        } else {
#pragma unroll
                            for (int q = 0; q < Q; ++q) val[q] = acc[q];
                            pow_all<Q>(val, pm, sv[s]);
                        }
#pragma unroll
                        for (int q = 0; q < Q; ++q) {
                            // the weight product and the deviation from the mean contract into one FMA: the mean
                            // takes the unrounded weighted value (within half an ulp of the stored per-frame value)
                            const float v = val[q] * inten[q];
                            mu[s * Q + q] += (v - mu[s * Q + q]) * inv_np1;
                        }
                    }
                }
            } else {
                pow_all<Q>(acc, a.sens_pow[0], a.sens[0]);
                if (a.iw) {
                    if constexpr (ARRAYS) {
#pragma unroll
                        for (int q = 0; q < Q; ++q) {
                            const int rM = min(M * (y0 + q) + py, a.HM - 1);
                            acc[q] = acc[q] * (IW_EARLY ? aiw[q]
                                                        : a.aimg[(static_cast<size_t>(f) * a.HM + rM) * a.WM + cMs]);
                        }
                    } else {
                        // the row values of up to 8 pixels are read before the first of them is used, so that their
                        // LDS reads overlap instead of waiting one pixel at a time
                        constexpr int QG = Q < 8 ? Q : 8;
                        const float* rowp = s_irow + rset * Q * TIC + ccol;
#pragma unroll
                        for (int g = 0; g < Q; g += QG) {
                            float tv[QG][4];
#pragma unroll
                            for (int q = 0; q < QG; ++q)
#pragma unroll
                                for (int v2 = 0; v2 < 4; ++v2) tv[q][v2] = rowp[(g + q) * TIC + v2];
#pragma unroll
                            for (int q = 0; q < QG; ++q) {
                                float sacc = 0.f;
#pragma unroll
                                for (int v2 = 0; v2 < 4; ++v2) sacc += tv[q][v2] * wcv[v2];
                                acc[g + q] = acc[g + q] * sacc;
                            }
                        }
                    }
                }
#pragma unroll
                for (int q = 0; q < Q; ++q) acc[q] = pixel_ok(q) ? acc[q] : 0.f;
                if constexpr (FRAMES) {
#pragma unroll
                    for (int q = 0; q < Q; ++q) {
                        const int yi = y0 + q;
                        if (xi < a.W && yi < a.H)
                            a.out[(static_cast<size_t>(f) * a.HM + M * yi + py) * a.WM + cM] = acc[q];
                    }
                } else if (nabs == 0) {
#pragma unroll
                    for (int q = 0; q < Q; ++q) {
                        float fq = 0.f, pq = 0.f;
                        moments_push<WITH_M2>(acc[q], 0, inv_np1, mu[q], m2[WITH_M2 ? q : 0], cc[q], fq, pq);
                        s_first[q * 64] = acc[q];
                        s_prev[q * 64] = acc[q];
                    }
                } else {
#pragma unroll
                    for (int q = 0; q < Q; ++q) {
                        float fq = s_first[q * 64], pq = s_prev[q * 64];
                        moments_push<WITH_M2>(acc[q], nabs, inv_np1, mu[q], m2[WITH_M2 ? q : 0], cc[q], fq, pq);
                        s_prev[q * 64] = acc[q];
                    }
                }
            }
        }
        __syncthreads();
    }

    if (!FRAMES && active && f_end > f_begin) {
        const size_t P = static_cast<size_t>(a.HM) * a.WM;
        const int pslot = a.chunk0 + chunk;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int yi = y0 + q;
            if (xi >= a.W || yi >= a.H) continue;
            const size_t pix = static_cast<size_t>(M * yi + py) * a.WM + cM;
            if constexpr (SWEEP) {
#pragma unroll
                for (int s = 0; s < NSM; ++s) {
                    if (s >= a.ns) continue;
                    if (a.direct) a.sweep_avg[rad * kMaxSens + s][pix] = mu[s * Q + q];
                    else a.part[((static_cast<size_t>(pslot) * a.nrad + rad) * a.ns + s) * P + pix] = mu[s * Q + q];
                }
            } else if (a.direct) {
                const float n = static_cast<float>(n0 + (f_end - f_begin));
                if (a.avg) a.avg[pix] = mu[q];
                if (a.var) a.var[pix] = WITH_M2 ? m2[WITH_M2 ? q : 0] / n : 0.f;
                if (a.tac2) a.tac2[pix] = cc[q] / (n - 1.0f);
            } else if (a.state) {
                a.state[pix] = mu[q];
                if (WITH_M2) a.state[P + pix] = m2[WITH_M2 ? q : 0];
                a.state[2 * P + pix] = cc[q];
                a.state[3 * P + pix] = s_first[q * 64];
                a.state[4 * P + pix] = s_prev[q * 64];
            } else {
                float* base = a.part + static_cast<size_t>(pslot) * 5 * P + pix;
                base[0] = mu[q];
                base[P] = WITH_M2 ? m2[WITH_M2 ? q : 0] : 0.f;
                base[2 * P] = cc[q];
                base[3 * P] = s_first[q * 64];
                base[4 * P] = s_prev[q * 64];
            }
        }
    }
}

}  // namespace npx
