// Host-side eSRRF plan tables (see npx_plan.h). Compiled for the host only; float64 is used here and
// never on the GPU. The float32 roundings the reference performs are reproduced with explicit casts
// (this file is built with -ffp-contract=off so no product is fused into a sum behind our back).
#include <algorithm>
#include <cmath>
#include <string>

#include "npx_plan.h"

namespace npx {

// Keys cubic, a = 0.5 (src/include/_c_interpolation_catmull_rom.c:6-18).
static double cubic(double v) {
    const double a = 0.5;
    if (v < 0) v = -v;
    if (v < 1) return v * v * (v * (-a + 2) + (a - 3)) + 1;
    if (v < 2) return -a * v * v * v + 5 * a * v * v - 8 * a * v + 4 * a;
    return 0.0;
}

CrTap cr_taps(int i, float mag, int n) {
    CrTap t{};
    // row = i / magnification (int / float in float), _le_interpolation_catmull_rom.pyx:165-168
    const float r = static_cast<float>(i) / mag;
    t.base = 0;
    for (int u = 0; u < 4; ++u) t.w[u] = 0.0f;
    if (r < 0 || r >= static_cast<float>(n)) return t;  // _c_interpolate returns 0 outside (:23-25)
    const int r_int = static_cast<int>(std::floor(static_cast<float>(static_cast<double>(r) - 0.5)));
    t.base = r_int - 1;
    for (int u = 0; u < 4; ++u) {
        const int rn = r_int - 1 + u;
        if (rn < 0 || rn >= n) continue;  // neighbour skipped, no renormalisation (:37-45)
        t.w[u] = static_cast<float>(cubic(static_cast<double>(r) - (rn + 0.5)));
    }
    return t;
}

CrTap cr_taps_shift(int i, float mag, float shift, int n) {
    CrTap t{};
    const float r = static_cast<float>(i) / mag - shift;
    t.base = 0;
    for (int u = 0; u < 4; ++u) t.w[u] = 0.0f;
    if (r < 0 || r >= static_cast<float>(n)) return t;
    const int r_int = static_cast<int>(std::floor(static_cast<float>(static_cast<double>(r) - 0.5)));
    t.base = r_int - 1;
    for (int u = 0; u < 4; ++u) {
        const int rn = r_int - 1 + u;
        if (rn < 0 || rn >= n) continue;
        t.w[u] = static_cast<float>(cubic(static_cast<double>(r) - (rn + 0.5)));
    }
    return t;
}

bool build_plan(int H, int W, int M, double R, PlanHost& p, std::string* err) {
    if (H < 1 || W < 1 || M < 1) {
        if (err) *err = "image size and magnification must be >= 1";
        return false;
    }
    p.H = H;
    p.W = W;
    p.M = M;
    p.R = R;
    p.HM = H * M;
    p.WM = W * M;
    // _le_radial_gradient_convergence.pyx:85-98: radius is a Python float, sigma = float(radius / 2.355)
    p.sigma = static_cast<float>(R / 2.355);
    p.fwhm = static_cast<float>(R);
    p.tSS = static_cast<float>(2.0f * p.sigma) * p.sigma;
    p.tSO = static_cast<float>(2.0f * p.sigma) + 1.0f;
    p.margin = static_cast<int>(static_cast<double>(p.fwhm) * M);
    // _c_sr_radial_gradient_convergence.c:27-28 with Gx_Gy_MAGNIFICATION = 2.0f
    const int start = -static_cast<int>(2.0f * p.fwhm);
    const int end = static_cast<int>(2.0f * p.fwhm + 1.0f);
    p.K = -start;
    if (end - 1 != p.K) {
        if (err) *err = "asymmetric ring window (radius too large for float32)";
        return false;
    }
    if (p.K > 6 || !(R > 0)) {
        if (err) *err = "radius must be in (0, 3.5): ring windows wider than 13x13 lattice points are not built";
        return false;
    }
    if (static_cast<long long>(2) * p.HM >= (1LL << 24) || static_cast<long long>(2) * p.WM >= (1LL << 24)) {
        if (err) *err = "magnified image too large for exact float32 sampling coordinates";
        return false;
    }
    const int K = p.K, NW = 2 * K + 1;
    p.nphase = M * M;
    p.ey.resize(M);
    p.ex.resize(M);
    for (int q = 0; q < M; ++q) {
        p.ey[q] = (2 * q + 1) / M;
        p.ex[q] = p.ey[q];
    }
    // Every output row rM = M*y + py must reach lattice row floor(2*yc) = 2*y + ey[py] with
    // yc = float((rM + 0.5) / M) as the reference computes it (:21-22,31); check all rows and columns.
    for (int dim = 0; dim < 2; ++dim) {
        const int n = dim ? p.WM : p.HM;
        for (int r = 0; r < n; ++r) {
            const float yc = static_cast<float>((r + 0.5) / M);
            const int a0 = static_cast<int>(2.0f * yc);
            if (a0 != 2 * (r / M) + p.ey[r % M]) {
                if (err) *err = "lattice base mismatch (image too large for float32 pixel centres)";
                return false;
            }
        }
    }
    // Ring geometry per phase, in the exact phase-relative coordinates:
    // dy = (ey + j)/2 - (py + 0.5)/M, dx = (ex + i)/2 - (px + 0.5)/M.
    const RingLayout L(K);
    p.ring_stride = L.stride;
    p.ring.assign(static_cast<size_t>(p.nphase) * L.stride, 0.0f);
    p.wsum.assign(p.nphase, 0.0f);
    const double tSS = p.tSS;
    for (int ph = 0; ph < p.nphase; ++ph) {
        const int py = ph / M, px = ph % M;
        float* blk = p.ring.data() + static_cast<size_t>(ph) * L.stride;
        for (int j = -K; j <= K; ++j) blk[L.dy + j + K] = static_cast<float>((p.ey[py] + j) / 2.0 - (py + 0.5) / M);
        for (int i = -K; i <= K; ++i) blk[L.dx + i + K] = static_cast<float>((p.ex[px] + i) / 2.0 - (px + 0.5) / M);
        std::vector<float> wgrid(NW * NW, 0.0f);  // [j][i]
        for (int j = -K; j <= K; ++j) {
            for (int i = -K; i <= K; ++i) {
                const float dy = blk[L.dy + j + K];
                const float dx = blk[L.dx + i + K];
                const float dd = static_cast<float>(dx * dx) + static_cast<float>(dy * dy);
                const float d = static_cast<float>(std::sqrt(static_cast<double>(dd)));
                const int e = (i + K) * NW + (j + K);
                blk[L.dt + 2 * e] = d;
                if (!(d != 0.0f && d <= p.tSO)) continue;  // :44
                // _c_calculate_dw (:4-6) in double, stored as float (:63)
                const double dd64 = d;
                const float w = static_cast<float>(std::pow(dd64 * std::exp((-dd64 * dd64) / tSS), 4));
                blk[L.dt + 2 * e + 1] = static_cast<float>(static_cast<double>(w) / dd64);
                blk[L.w + e] = w;
                wgrid[(j + K) * NW + (i + K)] = w;
            }
        }
        float s = 0.0f;  // distanceWeightSum, float, j outer / i inner (:30-35,64)
        for (int k = 0; k < NW * NW; ++k) s += wgrid[k];
        p.wsum[ph] = s;
        blk[L.wsum] = s;
        // prefix sums for pixels whose window is cut by the image border (rows / columns outside the valid
        // lattice range form a prefix or a suffix of the window)
        for (int j = 0; j <= NW; ++j)
            for (int i = 0; i <= NW; ++i) {
                double acc = 0.0;
                for (int jj = 0; jj < j; ++jj)
                    for (int ii = 0; ii < i; ++ii) acc += wgrid[jj * NW + ii];
                blk[L.prefix + j * (NW + 1) + i] = static_cast<float>(acc);
            }
    }
    // Gradient lattice taps: lattice index a reads row pos = a*M of the x2M image, clamped to 2*H*M-1
    // (:45-52); that row samples the gradient at pos / (2M) (_le_esrrf.pyx:192-193).
    const float mag2 = static_cast<float>(2 * M);
    p.lat_row.resize(2 * H + 1);
    p.lat_col.resize(2 * W + 1);
    for (int a = 0; a <= 2 * H; ++a) p.lat_row[a] = cr_taps(a < 2 * H ? a * M : 2 * H * M - 1, mag2, H);
    for (int b = 0; b <= 2 * W; ++b) p.lat_col[b] = cr_taps(b < 2 * W ? b * M : 2 * W * M - 1, mag2, W);
    // Intensity Catmull-Rom taps at x M (_le_esrrf.pyx:190).
    const float mag = static_cast<float>(M);
    p.cr_row.resize(p.HM);
    p.cr_col.resize(p.WM);
    for (int r = 0; r < p.HM; ++r) p.cr_row[r] = cr_taps(r, mag, H);
    for (int c = 0; c < p.WM; ++c) p.cr_col[c] = cr_taps(c, mag, W);
    return true;
}

// Taps of one ring-sample coordinate v = (p + 0.5 + ring) / M of the gradient image of size n, as
// _c_interpolate uses them: 0 weights when v is outside [0, n) or the tap is outside the image.
static CrTap ring_taps(float v, int n) {
    CrTap t{};
    const int v_int = static_cast<int>(std::floor(static_cast<float>(static_cast<double>(v) - 0.5)));
    t.base = v_int - 1;
    for (int u = 0; u < 4; ++u) t.w[u] = 0.0f;
    if (v < 0 || v >= static_cast<float>(n)) return t;
    for (int u = 0; u < 4; ++u) {
        const int vn = v_int - 1 + u;
        if (vn < 0 || vn >= n) continue;
        t.w[u] = static_cast<float>(cubic(static_cast<double>(v) - (vn + 0.5)));
    }
    return t;
}

bool build_srrf_plan(int H, int W, int M, double ringRadius, int border, int TJ, int TI, SrrfPlan& p,
                     std::string* err) {
    if (H < 1 || W < 1 || M < 1 || TJ < 1 || TI < 1) {
        if (err) *err = "image size and magnification must be >= 1";
        return false;
    }
    if (!(ringRadius >= 0) || !std::isfinite(ringRadius)) {
        if (err) *err = "ringRadius must be a finite number >= 0";
        return false;
    }
    p.H = H;
    p.W = W;
    p.M = M;
    p.HM = H * M;
    p.WM = W * M;
    p.border = border;
    p.TJ = TJ;
    p.TI = TI;
    // _le_radiality.pyx:95-103: _ringRadius = float(ringRadius * M); angleStep = float(2 pi / 12);
    // coordinate = float(cos(float(angleStep * k)) * _ringRadius) (double product)
    p.R = static_cast<float>(ringRadius * M);
    const float step = static_cast<float>((M_PI * 2.) / 12);
    for (int k = 0; k < 12; ++k) {
        const float a = step * static_cast<float>(k);
        p.xr[k] = static_cast<float>(std::cos(static_cast<double>(a)) * p.R);
        p.yr[k] = static_cast<float>(std::sin(static_cast<double>(a)) * p.R);
    }
    // computed range (_le_radiality.pyx:117-118)
    p.j_lo = (1 + border) * M;
    p.j_hi = (H - 1 - border) * M;
    p.i_lo = (1 + border) * M;
    p.i_hi = (W - 1 - border) * M;
    const float magf = static_cast<float>(M);
    // entries in absolute image coordinates first
    auto fill = [&](int n_out, int n_img, const float* ring, std::vector<float>& tab, std::vector<int>& lo_abs,
                    std::vector<int>& hi_abs) {
        tab.assign(static_cast<size_t>(n_out) * SrrfEntry::kFloats, 0.0f);
        lo_abs.assign(n_out, 0);
        hi_abs.assign(n_out, 0);
        for (int q = 0; q < n_out; ++q) {
            float* e = &tab[static_cast<size_t>(q) * SrrfEntry::kFloats];
            const CrTap it = cr_taps(q, magf, n_img);
            // xc = i + 0.5 (double, stored float); x0 = xc + ring (float); coordinate x0 / M (float)
            const float c = static_cast<float>(q + 0.5);
            int lo = it.base, hi = it.base + 3;
            for (int k = 0; k < 12; ++k) {
                const float x0 = c + ring[k];
                const float v = x0 / magf;
                CrTap t = ring_taps(v, n_img);
                bool any = false;
                for (int u = 0; u < 4; ++u) any = any || t.w[u] != 0.0f;
                if (!any) t.base = it.base;  // all-zero sample: park its taps on the intensity taps
                reinterpret_cast<int*>(e)[SrrfEntry::kBase + k] = t.base;
                for (int u = 0; u < 4; ++u) e[SrrfEntry::kW + 4 * k + u] = t.w[u];
                // offset over the ring radius: the kernel forms |G x d| / R and the sign of G . ring from it
                e[SrrfEntry::kD + k] = static_cast<float>(static_cast<double>(c - x0) / static_cast<double>(p.R));
                lo = std::min(lo, t.base);
                hi = std::max(hi, t.base + 3);
            }
            reinterpret_cast<int*>(e)[SrrfEntry::kIBase] = it.base;
            for (int u = 0; u < 4; ++u) e[SrrfEntry::kIW + u] = it.w[u];
            lo_abs[q] = lo;
            hi_abs[q] = hi;
        }
    };
    std::vector<int> rlo, rhi, clo, chi;
    fill(p.HM, H, p.yr, p.row, rlo, rhi);
    fill(p.WM, W, p.xr, p.col, clo, chi);
    // per block: gradient tile origin = lowest tap, extent = max over blocks; bases made tile-relative
    auto tiles = [&](int n_out, int T, std::vector<float>& tab, const std::vector<int>& lo, const std::vector<int>& hi,
                     std::vector<int>& t0, int& ext) {
        const int nt = (n_out + T - 1) / T;
        t0.assign(nt, 0);
        ext = 0;
        for (int b = 0; b < nt; ++b) {
            int a = lo[b * T], z = hi[b * T];
            for (int q = b * T; q < std::min(n_out, (b + 1) * T); ++q) {
                a = std::min(a, lo[q]);
                z = std::max(z, hi[q]);
            }
            t0[b] = a;
            ext = std::max(ext, z - a + 1);
            for (int q = b * T; q < std::min(n_out, (b + 1) * T); ++q) {
                int* e = reinterpret_cast<int*>(&tab[static_cast<size_t>(q) * SrrfEntry::kFloats]);
                for (int k = 0; k < 12; ++k) e[SrrfEntry::kBase + k] -= a;
                e[SrrfEntry::kIBase] -= a - 1;  // the input tile starts one before the gradient tile
            }
        }
    };
    tiles(p.HM, TJ, p.row, rlo, rhi, p.tile_r0, p.TH);
    tiles(p.WM, TI, p.col, clo, chi, p.tile_c0, p.TW);
    // every tap the kernel reads lies inside the block's tiles: gradient taps in [0, T) and intensity taps in
    // [0, T + 2) of the input tile
    auto inside = [&](const std::vector<float>& tab, int n_out, int ext) {
        for (int q = 0; q < n_out; ++q) {
            const int* e = reinterpret_cast<const int*>(&tab[static_cast<size_t>(q) * SrrfEntry::kFloats]);
            for (int k = 0; k < 12; ++k)
                if (e[SrrfEntry::kBase + k] < 0 || e[SrrfEntry::kBase + k] + 3 >= ext) return false;
            if (e[SrrfEntry::kIBase] < 0 || e[SrrfEntry::kIBase] + 3 >= ext + 2) return false;
        }
        return true;
    };
    if (!inside(p.row, p.HM, p.TH) || !inside(p.col, p.WM, p.TW)) {
        if (err) *err = "internal: a radiality tap falls outside its tile";
        return false;
    }
    return true;
}

}  // namespace npx
