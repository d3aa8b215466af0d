/*
 * ORACLE — test infrastructure only.
 *
 * FP32 CPU restatement of NanoPyx's threaded Cython eSRRF chain (v2.1.2), used as the checker for the
 * HIP path and as the timed CPU baseline. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product (nanopyx_amd) never does.
 *
 * Every function follows a reference file:line (paths relative to the reference repo root). Storage is
 * float32 and the per-pixel helpers compute in double exactly where the reference C does, so that this
 * file and the reference's own C kernels (built into oracle/_ref by oracle/Makefile) give bit-identical
 * results under the same compiler flags; tests/test_oracle.py checks that.
 *
 * The per-pixel kernels can be swapped for the reference's own compiled ones (ora_use_kernels), so the
 * same loop drivers run either "port" (this file) or "reference" (reference C + these loops).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

/* ---------------------------------------------------------------- per-pixel kernels (restated) */

/* src/include/_c_interpolation_catmull_rom.c:6-18 — Keys cubic, a = 0.5, in double. */
static double ora_cubic(double v) {
    double a = 0.5;
    double z = 0;
    if (v < 0) v = -v;
    if (v < 1) {
        z = v * v * (v * (-a + 2) + (a - 3)) + 1;
    } else if (v < 2) {
        z = -a * v * v * v + 5 * a * v * v - 8 * a * v + 4 * a;
    }
    return z;
}

/* src/include/_c_interpolation_catmull_rom.c:21-52 — 4x4 taps, pixel centres at n+0.5, neighbours
 * outside the image skipped (no renormalisation), columns outer / rows inner, double accumulators. */
static float ora_interpolate(float *image, float r, float c, int rows, int cols) {
    if (r < 0 || r >= rows || c < 0 || c >= cols) return 0;
    const int r_int = (int)floor((float)(r - 0.5));
    const int c_int = (int)floor((float)(c - 0.5));
    double q = 0;
    double p = 0;
    for (int j = 0; j < 4; j++) {
        int cn = c_int - 1 + j;
        p = 0;
        if (cn < 0 || cn >= cols) continue;
        for (int i = 0; i < 4; i++) {
            int rn = r_int - 1 + i;
            if (rn < 0 || rn >= rows) continue;
            p = p + image[rn * cols + cn] * ora_cubic(r - (rn + 0.5));
        }
        q = q + p * ora_cubic(c - (cn + 0.5));
    }
    return q;
}

/* src/include/_c_gradients.c:53-72 — Roberts cross with the previous row/column clamped at 0. */
static void ora_gradient_roberts_cross(float *image, float *imGc, float *imGr, int rows, int cols) {
    for (int r1 = 0; r1 < rows; r1++) {
        for (int c1 = 0; c1 < cols; c1++) {
            int c0 = c1 > 0 ? c1 - 1 : 0;
            int r0 = r1 > 0 ? r1 - 1 : 0;
            float a_r0c1 = image[r0 * cols + c1];
            float a_r1c0 = image[r1 * cols + c0];
            float a_r0c0 = image[r0 * cols + c0];
            float a_r1c1 = image[r1 * cols + c1];
            imGc[r1 * cols + c1] = a_r0c1 - a_r1c0 + a_r1c1 - a_r0c0;
            imGr[r1 * cols + c1] = -a_r0c1 + a_r1c0 + a_r1c1 - a_r0c0;
        }
    }
}

/* src/include/_c_sr_radial_gradient_convergence.c:4-6 */
static double ora_dw(double distance, double tSS) {
    return pow((distance * exp((-distance * distance) / tSS)), 4);
}

/* src/include/_c_sr_radial_gradient_convergence.c:8-15 */
static double ora_dk(float Gx, float Gy, float dx, float dy, float distance) {
    float Dk = fabs(Gy * dx - Gx * dy) / sqrt(Gx * Gx + Gy * Gy);
    if (isnan(Dk)) Dk = distance;
    Dk = 1 - Dk / distance;
    return Dk;
}

/* src/include/_c_sr_radial_gradient_convergence.c:17-90 — ring sum for one output pixel. */
static float ora_calculate_rgc(int xM, int yM, float *imIntGx, float *imIntGy, int colsM, int rowsM,
                               int magnification, float G2, float fwhm, float tSO, float tSS,
                               float sensitivity) {
    float vx, vy, Gx, Gy, dx, dy, distance, distanceWeight, GdotR, Dk, pos_x, pos_y;
    float xc = (xM + 0.5) / magnification;
    float yc = (yM + 0.5) / magnification;
    float RGC = 0;
    float distanceWeightSum = 0;
    int _start = -(int)(G2 * fwhm);
    int _end = (int)(G2 * fwhm + 1);
    for (int j = _start; j < _end; j++) {
        vy = (int)(G2 * yc) + j;
        vy /= G2;
        if (0 < vy && vy <= rowsM - 1) {
            for (int i = _start; i < _end; i++) {
                vx = (int)(G2 * xc) + i;
                vx /= G2;
                if (0 < vx && vx <= colsM - 1) {
                    dx = vx - xc;
                    dy = vy - yc;
                    distance = sqrt(dx * dx + dy * dy);
                    if (distance != 0 && distance <= tSO) {
                        pos_y = vy * magnification * G2;
                        pos_x = vx * magnification * G2;
                        if (pos_y >= rowsM * G2) pos_y = rowsM * G2 - 1;
                        if (pos_x >= colsM * G2) pos_x = colsM * G2 - 1;
                        if (pos_y < 0) pos_y = 0;
                        if (pos_x < 0) pos_x = 0;
                        int grad_width = (int)(colsM * G2);
                        Gx = imIntGx[(int)pos_y * grad_width + (int)pos_x];
                        Gy = imIntGy[(int)pos_y * grad_width + (int)pos_x];
                        distanceWeight = ora_dw(distance, tSS);
                        distanceWeightSum += distanceWeight;
                        GdotR = Gx * dx + Gy * dy;
                        if (GdotR < 0) {
                            Dk = ora_dk(Gx, Gy, dx, dy, distance);
                            RGC += Dk * distanceWeight;
                        }
                    }
                }
            }
        }
    }
    if (distanceWeightSum > 0) RGC /= distanceWeightSum;
    else RGC = 0;
    if (RGC >= 0 && sensitivity > 1) RGC = pow(RGC, sensitivity);
    else if (RGC < 0) RGC = 0;
    return RGC;
}

/* ---------------------------------------------------------------- kernel selection */

typedef float (*rgc_fn_t)(int, int, float *, float *, int, int, int, float, float, float, float, float);
typedef float (*interp_fn_t)(float *, float, float, int, int);
typedef void (*roberts_fn_t)(float *, float *, float *, int, int);

static rgc_fn_t k_rgc = ora_calculate_rgc;
static interp_fn_t k_interp = ora_interpolate;
static roberts_fn_t k_roberts = ora_gradient_roberts_cross;

/* Swap the per-pixel kernels (NULL restores this file's restatement). The pointers are the
 * reference's own _c_calculate_rgc / _c_interpolate / _c_gradient_roberts_cross from oracle/_ref. */
void ora_use_kernels(void *rgc, void *interp, void *roberts) {
    k_rgc = rgc ? (rgc_fn_t)rgc : ora_calculate_rgc;
    k_interp = interp ? (interp_fn_t)interp : ora_interpolate;
    k_roberts = roberts ? (roberts_fn_t)roberts : ora_gradient_roberts_cross;
}

/* Direct entry points to one per-pixel kernel (for the bit-exactness tests). */
float ora_px_interpolate(float *image, float r, float c, int rows, int cols) {
    return k_interp(image, r, c, rows, cols);
}
float ora_px_rgc(int xM, int yM, float *gx, float *gy, int colsM, int rowsM, int M, float G2, float fwhm,
                 float tSO, float tSS, float S) {
    return k_rgc(xM, yM, gx, gy, colsM, rowsM, M, G2, fwhm, tSO, tSS, S);
}
void ora_px_roberts(float *image, float *gc, float *gr, int rows, int cols) {
    k_roberts(image, gc, gr, rows, cols);
}

int ora_num_threads(void) { return omp_get_max_threads(); }

/* ---------------------------------------------------------------- stage drivers (Cython loops) */

/* ShiftAndMagnify._run_threaded: src/nanopyx/core/transform/_le_interpolation_catmull_rom.pyx:143-170.
 * Output (N, int(rows*mr), int(cols*mc)); prange over output columns; coordinates are float divisions. */
void ora_shift_magnify(const float *in, int nFrames, int rows, int cols, float shift_row, float shift_col,
                       float magnification_row, float magnification_col, float *out) {
    int rowsM = (int)(rows * magnification_row);
    int colsM = (int)(cols * magnification_col);
    for (int f = 0; f < nFrames; f++) {
        float *img = (float *)in + (size_t)f * rows * cols;
        float *o = out + (size_t)f * rowsM * colsM;
#pragma omp parallel for schedule(static)
        for (int j = 0; j < colsM; j++) {
            float col = j / magnification_col - shift_col;
            for (int i = 0; i < rowsM; i++) {
                float row = i / magnification_row - shift_row;
                o[(size_t)i * colsM + j] = k_interp(img, row, col, rows, cols);
            }
        }
    }
}

/* GradientRobertsCross._run_threaded: src/nanopyx/core/transform/_le_roberts_cross_gradients.pyx:67-83
 * (prange over frames). */
void ora_roberts(const float *in, int nFrames, int rows, int cols, float *gc, float *gr) {
#pragma omp parallel for schedule(static)
    for (int n = 0; n < nFrames; n++) {
        size_t off = (size_t)n * rows * cols;
        k_roberts((float *)in + off, gc + off, gr + off, rows, cols);
    }
}

/* Parameter derivation of RadialGradientConvergence._run_threaded
 * (src/nanopyx/core/transform/_le_radial_gradient_convergence.pyx:85-98): `radius` arrives as a Python
 * float, so sigma = float(radius / 2.355) with the division in double; tSS, tSO are float arithmetic;
 * margin = int(float(fwhm) * magnification) computed by Python in double. */
void ora_rgc_params(double radius, int magnification, int pad_edges, float *fwhm, float *sigma, float *tSS,
                    float *tSO, int *margin) {
    float s = (float)(radius / 2.355);
    float fw = (float)radius;
    *sigma = s;
    *fwhm = fw;
    *tSS = 2 * s * s;
    *tSO = 2 * s + 1;
    *margin = pad_edges ? 0 : (int)((double)fw * (double)magnification);
}

/* RadialGradientConvergence._run_threaded: _le_radial_gradient_convergence.pyx:79-115.
 * Frames sequential, prange over rows in [margin, rowsM - margin), zeros elsewhere. */
void ora_rgc(const float *gci, const float *gri, const float *img_i, int nFrames, int rowsM, int colsM,
             int magnification, double radius, double sensitivity, int doIntensityWeighting, int pad_edges,
             float *out) {
    float fwhm, sigma, tSS, tSO;
    int margin;
    ora_rgc_params(radius, magnification, pad_edges, &fwhm, &sigma, &tSS, &tSO, &margin);
    float G2 = 2.0f;
    float S = (float)sensitivity;
    size_t plane = (size_t)rowsM * colsM;
    size_t gplane = (size_t)(2 * rowsM) * (2 * colsM);
    memset(out, 0, sizeof(float) * plane * nFrames);
    for (int f = 0; f < nFrames; f++) {
        float *gx = (float *)gci + f * gplane;
        float *gy = (float *)gri + f * gplane;
        const float *im = img_i + f * plane;
        float *o = out + f * plane;
#pragma omp parallel for schedule(static)
        for (int rM = margin; rM < rowsM - margin; rM++) {
            for (int cM = margin; cM < colsM - margin; cM++) {
                float v = k_rgc(cM, rM, gx, gy, colsM, rowsM, magnification, G2, fwhm, tSO, tSS, S);
                if (doIntensityWeighting) v = v * im[(size_t)rM * colsM + cM];
                o[(size_t)rM * colsM + cM] = v;
            }
        }
    }
}

/* eSRRF._run_threaded (non-FHT branch): src/nanopyx/core/transform/_le_esrrf.pyx:172-195.
 * CR xM, Roberts, CR x2M of both gradients, RGC; pad_edges is not forwarded (:194). Frames are processed in
 * chunks of `chunk` so the full-size intermediates stay bounded; each chunk runs the reference's stages. */
int ora_esrrf(const float *in, int nFrames, int rows, int cols, int magnification, double radius,
              double sensitivity, int doIntensityWeighting, int chunk, float *out) {
    if (chunk <= 0) chunk = nFrames;
    int M = magnification;
    size_t p = (size_t)rows * cols, pM = (size_t)(rows * M) * (cols * M), p2 = 4 * pM;
    float *mag = (float *)malloc(sizeof(float) * pM * chunk);
    float *gc = (float *)malloc(sizeof(float) * p * chunk);
    float *gr = (float *)malloc(sizeof(float) * p * chunk);
    float *gci = (float *)malloc(sizeof(float) * p2 * chunk);
    float *gri = (float *)malloc(sizeof(float) * p2 * chunk);
    if (!mag || !gc || !gr || !gci || !gri) {
        free(mag); free(gc); free(gr); free(gci); free(gri);
        return -1;
    }
    for (int f0 = 0; f0 < nFrames; f0 += chunk) {
        int n = nFrames - f0 < chunk ? nFrames - f0 : chunk;
        const float *x = in + (size_t)f0 * p;
        ora_shift_magnify(x, n, rows, cols, 0, 0, (float)M, (float)M, mag);
        ora_roberts(x, n, rows, cols, gc, gr);
        ora_shift_magnify(gc, n, rows, cols, 0, 0, (float)(M * 2), (float)(M * 2), gci);
        ora_shift_magnify(gr, n, rows, cols, 0, 0, (float)(M * 2), (float)(M * 2), gri);
        ora_rgc(gci, gri, mag, n, rows * M, cols * M, M, radius, sensitivity, doIntensityWeighting, 0,
                out + (size_t)f0 * pM);
    }
    free(mag); free(gc); free(gr); free(gci); free(gri);
    return 0;
}


/* ---------------------------------------------------------------- SRRF radiality (restated) */

/* src/include/_c_gradients.c:3-25 — central differences on the interior, borders left untouched (0). */
static void ora_gradient_radiality(float *image, float *imGc, float *imGr, int rows, int cols) {
    for (int j = 1; j < rows - 1; j++) {
        int r0 = j * cols, r_m1 = (j - 1) * cols, r_p1 = (j + 1) * cols;
        for (int i = 1; i < cols - 1; i++) {
            imGc[r0 + i] = -image[r0 + i - 1] + image[r0 + i + 1];
            imGr[r0 + i] = -image[r_m1 + i] + image[r_p1 + i];
        }
    }
}

/* src/include/_c_sr_radiality.c:6-13 — squared alignment term; the division runs in double. */
static float ora_radiality_dk(float x, float y, float xc, float yc, float vGx, float vGy, float GMag,
                              float ringRadius) {
    float Dk = 0;
    if (GMag != 0) {
        Dk = 1 - (fabs(vGy * (xc - x) - vGx * (yc - y)) / GMag) / ringRadius;
        Dk = Dk * Dk;
    }
    return Dk;
}

/* src/include/_c_sr_radiality.c:15-51 — 12 ring samples of the Catmull-Rom interpolated gradient at
 * ((i + 0.5 + xRing) / M, (j + 0.5 + yRing) / M); +Dk for samples whose gradient points inwards, -Dk otherwise;
 * mean over the ring, clamped at 0 with the positivity constraint. */
static float ora_radiality_per_subpixel(int i, int j, float *imGx, float *imGy, float *xRing, float *yRing,
                                        int magnification, float ringRadius, int nRing, int positivity, int h,
                                        int w) {
    float x0, y0, xc, yc, xr, yr, vGx, vGy, GMag, Dk, DivDFactor = 0, CGH = 0;
    xc = i + 0.5;
    yc = j + 0.5;
    for (int s = 0; s < nRing; s++) {
        xr = xRing[s];
        yr = yRing[s];
        x0 = xc + xr;
        y0 = yc + yr;
        vGx = k_interp(imGx, y0 / magnification, x0 / magnification, h, w);
        vGy = k_interp(imGy, y0 / magnification, x0 / magnification, h, w);
        GMag = sqrt(vGx * vGx + vGy * vGy);
        Dk = ora_radiality_dk(x0, y0, xc, yc, vGx, vGy, GMag, ringRadius);
        if ((vGx * xr + vGy * yr) > 0) {
            DivDFactor -= Dk;
        } else {
            DivDFactor += Dk;
        }
    }
    DivDFactor /= nRing;
    if (positivity == 1) {
        CGH = fmaxf(DivDFactor, 0);
    } else {
        CGH = DivDFactor;
    }
    return CGH;
}

typedef float (*radiality_fn_t)(int, int, float *, float *, float *, float *, int, float, int, int, int, int);
typedef void (*grad_fn_t)(float *, float *, float *, int, int);
static radiality_fn_t k_radiality = ora_radiality_per_subpixel;
static grad_fn_t k_grad_radiality = ora_gradient_radiality;

/* Swap the radiality kernels for the reference's _c_calculate_radiality_per_subpixel / _c_gradient_radiality
 * (NULL restores the restatement). The reference's radiality calls its own _c_interpolate. */
void ora_use_radiality_kernels(void *radiality, void *grad) {
    k_radiality = radiality ? (radiality_fn_t)radiality : ora_radiality_per_subpixel;
    k_grad_radiality = grad ? (grad_fn_t)grad : ora_gradient_radiality;
}

float ora_px_radiality(int i, int j, float *gx, float *gy, float *xRing, float *yRing, int M, float ringRadius,
                       int nRing, int positivity, int h, int w) {
    return k_radiality(i, j, gx, gy, xRing, yRing, M, ringRadius, nRing, positivity, h, w);
}
void ora_px_gradient_radiality(float *image, float *gc, float *gr, int rows, int cols) {
    k_grad_radiality(image, gc, gr, rows, cols);
}

/* Ring coordinates of Radiality._run_threaded (src/nanopyx/core/transform/_le_radiality.pyx:95-103): 12 angles,
 * angleStep = float(2 pi / 12), coordinates cos/sin(angleStep * k) * ringRadius * M in double, stored as float. */
void ora_radiality_ring(double ringRadius, int magnification, float *xRing, float *yRing, float *ringR) {
    float R = (float)(ringRadius * magnification);
    float angleStep = (float)((M_PI * 2.) / 12);
    for (int k = 0; k < 12; k++) {
        xRing[k] = (float)(cos(angleStep * k) * R);
        yRing[k] = (float)(sin(angleStep * k) * R);
    }
    *ringR = R;
}

/* Radiality._run_threaded: src/nanopyx/core/transform/_le_radiality.pyx:86-126. Per frame: gradients, then
 * every output pixel of [(1+border)M, (h-1-border)M)^2 (prange over rows), times image_interp when
 * doIntensityWeighting; everything else stays 0. */
void ora_radiality(const float *image, const float *image_interp, int nFrames, int h, int w, int magnification,
                   double ringRadius, int border, int positivity, int doIntensityWeighting, float *out) {
    float xRing[12], yRing[12], R;
    ora_radiality_ring(ringRadius, magnification, xRing, yRing, &R);
    int M = magnification, hM = h * M, wM = w * M;
    size_t p = (size_t)h * w, pM = (size_t)hM * wM;
    float *gx = (float *)calloc(p, sizeof(float));
    float *gy = (float *)calloc(p, sizeof(float));
    for (int f = 0; f < nFrames; f++) {
        k_grad_radiality((float *)image + f * p, gx, gy, h, w);
        float *o = out + f * pM;
        const float *ii = image_interp + f * pM;
#pragma omp parallel for schedule(static)
        for (int j = (1 + border) * M; j < (h - 1 - border) * M; j++) {
            for (int i = (1 + border) * M; i < (w - 1 - border) * M; i++) {
                float v = k_radiality(i, j, gx, gy, xRing, yRing, M, R, 12, positivity, h, w);
                o[(size_t)j * wM + i] = doIntensityWeighting ? v * ii[(size_t)j * wM + i] : v;
            }
        }
    }
    free(gx);
    free(gy);
}

/* ---------------------------------------------------------------- eSRRF 3D (restated) */

/* src/include/_c_gradients.c:75-116 — per voxel: x and y central differences / 2, z backward difference / 2,
 * neighbours clamped to the volume. */
static void ora_gradient_3d(float *image, float *imGc, float *imGr, float *imGs, int slices, int rows, int cols) {
    for (int z = 0; z < slices; z++) {
        for (int y = 0; y < rows; y++) {
            for (int x = 0; x < cols; x++) {
                int zm = z <= 0 ? 0 : z - 1;
                int yp = y >= rows - 1 ? rows - 1 : y + 1, ym = y <= 0 ? 0 : y - 1;
                int xp = x >= cols - 1 ? cols - 1 : x + 1, xm = x <= 0 ? 0 : x - 1;
                size_t plane = (size_t)rows * cols, o = z * plane + (size_t)y * cols + x;
                float c = image[o];
                float yplus = image[z * plane + (size_t)yp * cols + x];
                float xplus = image[z * plane + (size_t)y * cols + xp];
                float yminus = image[z * plane + (size_t)ym * cols + x];
                float xminus = image[z * plane + (size_t)y * cols + xm];
                float zminus = image[zm * plane + (size_t)y * cols + x];
                imGc[o] = (xplus - xminus) / 2;
                imGr[o] = (yplus - yminus) / 2;
                imGs[o] = (c - zminus) / 2;
            }
        }
    }
}

/* src/include/_c_sr_radial_gradient_convergence.c:96-102 — the factors are float, the power is double. */
static double ora_dw3d(double distance, double distance_xy, double distance_z, double tSS, double tSS_z) {
    float wxy, wz, w;
    wxy = (exp((-distance_xy * distance_xy) / tSS));
    wz = (exp((-distance_z * distance_z) / tSS_z));
    w = distance * (wxy * wz);
    return pow(w, 4);
}

/* src/include/_c_sr_radial_gradient_convergence.c:112-119 — 1 - |G x d| / (|G| |d|); NaN (G = 0) gives 0. */
static double ora_dk3d(float Gx, float Gy, float Gz, float dx, float dy, float dz, float distance) {
    float Dk = sqrt((Gy * dz - Gz * dy) * (Gy * dz - Gz * dy) + (Gz * dx - Gx * dz) * (Gz * dx - Gx * dz) +
                    (Gx * dy - Gy * dx) * (Gx * dy - Gy * dx)) / sqrt(Gx * Gx + Gy * Gy + Gz * Gz);
    if (isnan(Dk)) Dk = distance;
    Dk = 1 - Dk / distance;
    return Dk;
}

/* src/include/_c_sr_radial_gradient_convergence.c:132-204 — window [-int(fwhm), int(fwhm + 1)] in x and y,
 * [-int(fwhm_z), int(fwhm_z + 1)] in z around (xM/Mxy, yM/Mxy, sliceM/Mz); samples inside the volume (strictly
 * positive coordinates), d != 0, d_xy <= tSO and (signed) d_z <= tSO_z; gradients read from the magnified volumes at
 * (int(vz Mz), int(Mxy vy), int(Mxy vx)); RGC = sum(Dk w over G.d < 0) / sum(w), then ^S when S > 1, 0 if < 0. */
static float ora_rgc3d(int xM, int yM, int sliceM, float *imIntGx, float *imIntGy, float *imIntGz, int colsM,
                       int rowsM, int slicesM, int magnification_xy, int magnification_z, float voxel_ratio,
                       float fwhm, float fwhm_z, float tSO, float tSO_z, float tSS, float tSS_z, float sensitivity) {
    float vx, vy, vz, Gx, Gy, Gz, dx, dy, dz, distance, distance_xy, w, GdotR, Dk;
    (void)voxel_ratio;
    float xc = (float)(xM) / magnification_xy;
    float yc = (float)(yM) / magnification_xy;
    float zc = (float)(sliceM) / magnification_z;
    float RGC = 0;
    float wsum = 0;
    int j0 = -(int)(fwhm), j1 = (int)(fwhm + 1);
    int k0 = -(int)(fwhm_z), k1 = (int)(fwhm_z + 1);
    for (int j = j0; j <= j1; j++) {
        vy = yc + j;
        if (!(0 < vy && vy <= (rowsM / magnification_xy) - 1)) continue;
        for (int i = j0; i <= j1; i++) {
            vx = xc + i;
            if (!(0 < vx && vx <= (colsM / magnification_xy) - 1)) continue;
            for (int k = k0; k <= k1; k++) {
                vz = zc + k;
                if (!(0 < vz && vz <= (slicesM / magnification_z) - 1)) continue;
                dx = vx - xc;
                dy = vy - yc;
                dz = vz - zc;
                distance = sqrt(dx * dx + dy * dy + dz * dz);
                distance_xy = sqrt(dx * dx + dy * dy);
                if (distance != 0 && distance_xy <= tSO && dz <= tSO_z) {
                    int idx = (int)(vz * magnification_z) * rowsM * colsM + (int)(magnification_xy * vy) * colsM +
                              (int)(magnification_xy * vx);
                    Gx = imIntGx[idx];
                    Gy = imIntGy[idx];
                    Gz = imIntGz[idx];
                    w = ora_dw3d(distance, distance_xy, dz, tSS, tSS_z);
                    wsum += w;
                    GdotR = Gx * dx + Gy * dy + Gz * dz;
                    if (GdotR < 0) {
                        Dk = ora_dk3d(Gx, Gy, Gz, dx, dy, dz, distance);
                        RGC += (Dk * w);
                    }
                }
            }
        }
    }
    RGC /= wsum;
    if (RGC >= 0 && sensitivity > 1) {
        RGC = pow(RGC, sensitivity);
    } else if (RGC < 0) {
        RGC = 0;
    }
    return RGC;
}

typedef float (*rgc3d_fn_t)(int, int, int, float *, float *, float *, int, int, int, int, int, float, float, float,
                            float, float, float, float, float);
typedef void (*grad3d_fn_t)(float *, float *, float *, float *, int, int, int);
static rgc3d_fn_t k_rgc3d = ora_rgc3d;
static grad3d_fn_t k_grad3d = ora_gradient_3d;

/* Swap the 3D kernels for the reference's _c_calculate_rgc3D / _c_gradient_3d (NULL restores the restatement). */
void ora_use_3d_kernels(void *rgc3d, void *grad3d) {
    k_rgc3d = rgc3d ? (rgc3d_fn_t)rgc3d : ora_rgc3d;
    k_grad3d = grad3d ? (grad3d_fn_t)grad3d : ora_gradient_3d;
}

/* interpolate_3d_zlinear (src/nanopyx/core/transform/_interpolation.pyx:91-103, :54-88): Catmull-Rom xMxy of every
 * slice (skipped when Mxy = 1), then the z step, whose integer slice position makes it a repetition of each slice
 * Mz times (weight1 = slc - slice0 = 0). */
static void ora_interpolate_3d_zlinear(const float *vol, int slices, int rows, int cols, int Mxy, int Mz,
                                       float *xy_scratch, float *out) {
    int rM = rows * Mxy, cM = cols * Mxy;
    size_t pM = (size_t)rM * cM;
    if (Mxy > 1)
        ora_shift_magnify(vol, slices, rows, cols, 0, 0, (float)Mxy, (float)Mxy, xy_scratch);
    else
        memcpy(xy_scratch, vol, sizeof(float) * pM * slices);
    for (int sM = 0; sM < slices * Mz; sM++) memcpy(out + sM * pM, xy_scratch + (size_t)(sM / Mz) * pM, sizeof(float) * pM);
}

/* eSRRF3D._run_threaded: src/nanopyx/core/transform/_le_esrrf3d.pyx:73-175. Parameters derived as there (float32
 * sigma, tSS, tSO; margins int(radius * Mxy), int(radius_z * Mz)); per frame: magnified intensity, 3D gradients,
 * magnified gradients, RGC over [margin_z, SM - margin_z) x [margin, RM - margin) x [margin, CM - margin), times the
 * magnified intensity, folded into a running mean (mode 0) or Welford M2 (mode 1, then sqrt(M2 / n)). */
int ora_esrrf3d(const float *in, int nFrames, int slices, int rows, int cols, int Mxy, int Mz, double radius,
                double radius_z, double voxel_ratio, double sensitivity, int mode, int doIntensityWeighting,
                float *out) {
    float sigma = (float)(radius / 2.355);
    int margin = (int)(radius * Mxy);
    float tSS = 2 * sigma * sigma;
    float tSO = 2 * sigma + 1;
    float sigma_z = (float)(radius_z / 2.355);
    int margin_z = (int)(radius_z * Mz);
    float tSS_z = 2 * sigma_z * sigma_z;
    float tSO_z = 2 * sigma_z + 1;
    float fwhm = (float)radius, fwhm_z = (float)radius_z, vr = (float)voxel_ratio, S = (float)sensitivity;
    int SM = slices * Mz, RM = rows * Mxy, CM = cols * Mxy;
    size_t p = (size_t)slices * rows * cols, pxy = (size_t)slices * RM * CM, pM = (size_t)SM * RM * CM;
    float *scratch = (float *)malloc(sizeof(float) * pxy);
    float *interp = (float *)malloc(sizeof(float) * pM);
    float *g[3], *gm[3];
    int ok = scratch && interp;
    for (int c = 0; c < 3; c++) {
        g[c] = (float *)calloc(p, sizeof(float));
        gm[c] = (float *)malloc(sizeof(float) * pM);
        ok = ok && g[c] && gm[c];
    }
    float *mean = mode == 1 ? (float *)calloc(pM, sizeof(float)) : NULL;
    if (!ok || (mode == 1 && !mean)) return -1;
    memset(out, 0, sizeof(float) * pM);
    for (int f = 0; f < nFrames; f++) {
        const float *vol = in + f * p;
        ora_interpolate_3d_zlinear(vol, slices, rows, cols, Mxy, Mz, scratch, interp);
        k_grad3d((float *)vol, g[0], g[1], g[2], slices, rows, cols);
        for (int c = 0; c < 3; c++) ora_interpolate_3d_zlinear(g[c], slices, rows, cols, Mxy, Mz, scratch, gm[c]);
        for (int sM = margin_z; sM < SM - margin_z; sM++) {
#pragma omp parallel for schedule(static)
            for (int rM = margin; rM < RM - margin; rM++) {
                for (int cM = margin; cM < CM - margin; cM++) {
                    float v = k_rgc3d(cM, rM, sM, gm[0], gm[1], gm[2], CM, RM, SM, Mxy, Mz, vr, fwhm, fwhm_z, tSO,
                                      tSO_z, tSS, tSS_z, S);
                    size_t o = (size_t)sM * RM * CM + (size_t)rM * CM + cM;
                    if (doIntensityWeighting) v = v * interp[o];
                    if (mode == 0) {
                        out[o] = out[o] + (v - out[o]) / (f + 1);
                    } else {
                        float delta = v - mean[o];
                        mean[o] += delta / (f + 1);
                        float delta2 = v - mean[o];
                        out[o] += delta * delta2;
                    }
                }
            }
        }
    }
    if (mode == 1)
        for (size_t o = 0; o < pM; o++) out[o] = sqrtf(out[o] / (float)nFrames);
    free(scratch); free(interp); free(mean);
    for (int c = 0; c < 3; c++) { free(g[c]); free(gm[c]); }
    return 0;
}
