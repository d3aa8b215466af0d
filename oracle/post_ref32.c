/*
 * ORACLE — test infrastructure only (see esrrf_ref32.c for the rules on who may load it).
 *
 * NumPy-float32 emulation of the eSRRF post-processing steps. This file is compiled with
 * -ffp-contract=off: NumPy rounds every product before it is summed, so no multiply-add may be fused.
 */
#include <stdlib.h>
#include <omp.h>


/* calculate_eSRRF_temporal_correlations (src/nanopyx/core/transform/sr_temporal_correlations.py:46-75,
 * calculate_tac2 :205-221) as NumPy evaluates it on a C-contiguous float32 (N, P) stack: reductions over
 * axis 0 add frame after frame into a float32 accumulator; the division by the item count (an np.intp)
 * promotes to float64 and is cast back to float32. mode 0 = AVG (np.mean), 1 = VAR (np.var, ddof 0),
 * 2 = TAC2. */
void ora_temporal(const float *stack, int N, long P, int mode, float *out) {
#pragma omp parallel for schedule(static)
    for (long k = 0; k < P; k++) {
        float s = 0;
        for (int t = 0; t < N; t++) s += stack[(size_t)t * P + k];
        float mean = (float)((double)s / (double)N);
        if (mode == 0) {
            out[k] = mean;
        } else if (mode == 1) {
            float acc = 0;
            for (int t = 0; t < N; t++) {
                float d = stack[(size_t)t * P + k] - mean;
                acc += d * d;
            }
            out[k] = (float)((double)acc / (double)N);
        } else {
            float acc = 0;
            for (int t = 0; t + 1 < N; t++) {
                float a = stack[(size_t)t * P + k] - mean;
                float b = stack[(size_t)(t + 1) * P + k] - mean;
                acc += a * b;
            }
            out[k] = (float)((double)acc / (double)(N - 1));
        }
    }
}

/* macro_pixel_corrector: src/nanopyx/core/transform/mpcorrector.pyx:7-73, in place on a float32 copy.
 * Per slice, an M x M phase map of float32 sums of img/(rows*cols) in row-block order, then each pixel is
 * divided by its phase value when that value is non-zero. */
void ora_macro_pixel(float *img, int slices, int rowsM, int colsM, int M) {
    int rows = rowsM / M, cols = colsM / M;
    float *cm = (float *)malloc(sizeof(float) * M * M);
    for (int s = 0; s < slices; s++) {
        float *im = img + (size_t)s * rowsM * colsM;
        for (int k = 0; k < M * M; k++) cm[k] = 0;
        for (int ry = 0; ry < rows; ry++)
            for (int rx = 0; rx < cols; rx++)
                for (int y = 0; y < M; y++)
                    for (int x = 0; x < M; x++) {
                        int gx = rx * M + x, gy = ry * M + y;
                        cm[y * M + x] += im[(size_t)gy * colsM + gx] / (rows * cols);
                    }
        for (int yM = 0; yM < rowsM; yM++)
            for (int xM = 0; xM < colsM; xM++) {
                float c = cm[(yM % M) * M + (xM % M)];
                if (c != 0) im[(size_t)yM * colsM + xM] /= c;
            }
    }
    free(cm);
}
