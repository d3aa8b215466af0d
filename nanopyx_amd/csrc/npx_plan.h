// Host-side eSRRF plan: every quantity of the reference algorithm that depends only on the image size,
// the magnification and the radius, computed once in float64 on the host exactly as the reference
// derives it, then handed to the gfx950 kernels as small FP32 tables.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace npx {

// One tap row of a separable 4-tap Catmull-Rom evaluation: first source index and four weights
// (weights of taps outside the image are 0, the reference skips them).
struct CrTap {
    int base;
    float w[4];
};

struct PlanHost {
    int H = 0, W = 0, M = 0;
    double R = 0;
    // reference-derived constants (src/nanopyx/core/transform/_le_radial_gradient_convergence.pyx:85-98)
    float fwhm = 0, sigma = 0, tSS = 0, tSO = 0;
    int K = 0;            // window half-width int(2*fwhm) (_c_sr_radial_gradient_convergence.c:27-28)
    int margin = 0;       // int(fwhm * M)
    int HM = 0, WM = 0;
    int nphase = 0;       // M * M output phases
    // ring tables, one block of ring_stride floats per phase (layout in ring_layout below)
    std::vector<float> ring;
    int ring_stride = 0;
    std::vector<float> wsum;        // [phase] float32 sum of w in the reference's j-outer/i-inner order
    std::vector<int> ey, ex;        // [phase row], [phase col]: floor((2p+1)/M)
    // Catmull-Rom taps for the x2M gradient lattice (index a in [0, 2H], 2H = the clamped last row)
    std::vector<CrTap> lat_row, lat_col;
    // Catmull-Rom taps for the xM intensity image (index rM in [0, HM), cM in [0, WM))
    std::vector<CrTap> cr_row, cr_col;
};

// Per-phase ring block (float offsets from the block start), NW = 2K+1, window offsets i (columns), j (rows):
//   dy[j]          at 0            offset of lattice row j from the pixel centre, float32
//   dx[i]          at NW           offset of lattice column i
//   (d, t)[i][j]   at 2*NW         distance and t = w / d (t = 0 outside the ring)
//   w[i][j]        at 2*NW+2*NW^2  distance weight (0 outside the ring)
//   wsum           at 2*NW+3*NW^2  sum of w (the interior normaliser)
//   P[j][i]        at 2*NW+3*NW^2+1  (NW+1)^2 prefix sums of w over rows < j, columns < i (edge normalisers)
struct RingLayout {
    int NW, dy, dx, dt, w, wsum, prefix, stride;
    explicit RingLayout(int K)
        : NW(2 * K + 1), dy(0), dx(NW), dt(2 * NW), w(2 * NW + 2 * NW * NW), wsum(2 * NW + 3 * NW * NW),
          prefix(2 * NW + 3 * NW * NW + 1),
          stride(((2 * NW + 3 * NW * NW + 1 + (NW + 1) * (NW + 1)) + 15) / 16 * 16) {}
};

// Builds the tables; returns false with a message on unsupported input.
bool build_plan(int H, int W, int M, double R, PlanHost& p, std::string* err);

// ---- SRRF radiality -------------------------------------------------------------------------------------------
// One table entry per output column (or row): for each of the 12 ring samples the first of the 4 Catmull-Rom taps
// of the gradient (relative to the gradient tile of the block that owns the column/row), the 4 tap weights (0 for
// taps outside the image and for samples outside the image, where the reference returns 0) and the offset
// (xc - x0) / R (or (yc - y0) / R) with R = ringRadius * M; then the intensity taps (first tap relative to the block's input tile, which
// starts one row/column before the gradient tile) for the xM Catmull-Rom magnification.
struct SrrfEntry {
    static constexpr int kRing = 12;
    static constexpr int kBase = 0, kW = 12, kD = 60, kIBase = 72, kIW = 73, kFloats = 80;
};

struct SrrfPlan {
    int H = 0, W = 0, M = 0, HM = 0, WM = 0, border = 0;
    float R = 0;                            // ringRadius * M, float32
    float xr[12] = {}, yr[12] = {};         // ring coordinates
    int j_lo = 0, j_hi = 0, i_lo = 0, i_hi = 0;  // computed output range [lo, hi); everything else is 0
    int TJ = 0, TI = 0;                     // output rows / columns per block
    int TH = 0, TW = 0;                     // gradient tile rows / columns per block (max over blocks)
    std::vector<int> tile_r0, tile_c0;      // [row band], [column tile]: image row / column of gradient tile 0
    std::vector<float> row, col;            // [HM][80], [WM][80]
};

// Reference-exact ring, positions and Catmull-Rom taps of Radiality (src/nanopyx/core/transform/_le_radiality.pyx:
// 86-126, src/include/_c_sr_radiality.c:15-51, src/include/_c_interpolation_catmull_rom.c:21-52). TJ output rows and
// TI output columns per block.
bool build_srrf_plan(int H, int W, int M, double ringRadius, int border, int TJ, int TI, SrrfPlan& p,
                     std::string* err);

// Taps of one output coordinate of ShiftAndMagnify: float(i) / mag - shift, in float32 as
// src/nanopyx/core/transform/_le_interpolation_catmull_rom.pyx:165-168 forms it (0 weights outside [0, n)).
CrTap cr_taps_shift(int i, float mag, float shift, int n);

// Reference-exact Catmull-Rom taps of one output coordinate i sampled at float(i) / float(mag)
// (src/nanopyx/core/transform/_le_interpolation_catmull_rom.pyx:165-168, _c_interpolation_catmull_rom.c:21-52).
CrTap cr_taps(int i, float mag, int n);

}  // namespace npx
