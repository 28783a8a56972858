/*
 * lzr_find_oracle.c -- FP64 CPU oracle of candidate generation.  TEST INFRASTRUCTURE
 * ONLY: the tests and the benchmark's CPU leg use it to check the GPU path; the product
 * never calls it.
 *
 *   orc_glrt_mask       ThresholdImage_GLRT_FFT::do_thresholding (source/Segmentation.cpp:241-430):
 *                       generalized likelihood-ratio test of "Gaussian spot + offset" against
 *                       "flat" in every window of side ws = ceil(4 w) made odd.
 *   orc_find_particles  ParticleFinder_adjacent8::findPositions (source/ParticleFinding.cpp:283-427):
 *                       8-connected components of the mask.
 *
 * GLRT as the reference computes it (formula order kept): window sums S1, S2 (the
 * reference takes them from an integral image, exact integers in FP64), average
 * m = S1/N, null-hypothesis variance V = S2 - m*m*N, the correlation G of the window with
 * the zero-mean Gaussian gc (gc = g - mean(g), g = exp(-(dx^2+dy^2)/(2w^2))/(1.77245385 w)),
 * alpha = G / sum(gc^2), test T = alpha*alpha / V * sum(gc^2), pixel segmented iff
 * 1 - exp(-PFA/N) < T < 1, for x in [h+1, X-h), y in [h+1, Y-h) (h = ws/2; X, Y the frame
 * size rounded down to even, :339-357).  The reference evaluates G by FFT for
 * power-of-two frames or ws > 7 and by a direct sum otherwise; the oracle always sums
 * directly (the same value up to rounding: gc is symmetric, and no window of the tested
 * range wraps around the frame).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "lzr_oracle.h"

int orc_glrt_window(double width) {
    int ws = (int)ceil(4.0 * width);
    if (ws % 2 == 0) ws += 1;
    return ws;
}

int orc_glrt_mask(const uint16_t *frame, int nx, int ny, double width, double pfa, uint8_t *mask,
                  double *test_out) {
    memset(mask, 0, (size_t)nx * ny);
    if (test_out)
        for (size_t k = 0; k < (size_t)nx * ny; ++k) test_out[k] = 0.0;
    const int ws = orc_glrt_window(width);
    if (ws > 31) return -1;
    const int h = ws / 2;
    const double N = (double)ws * (double)ws;
    const int X = nx - (nx % 2), Y = ny - (ny % 2);
    double gc[31 * 31];
    double sum = 0.0;
    for (int j = -h; j <= h; ++j)
        for (int i = -h; i <= h; ++i) {
            const double v = 1.0 / (1.77245385 * width) *
                             exp(-1.0 / (2.0 * width * width) * ((double)i * i + (double)j * j));
            gc[(i + h) + ws * (j + h)] = v;
            sum += v;
        }
    sum /= N;
    double ssg = 0.0;
    for (int k = 0; k < ws * ws; ++k) {
        gc[k] -= sum;
        ssg += gc[k] * gc[k];
    }
    const double thr = 1.0 - exp(-pfa / N);
    for (int y = h + 1; y < Y - h; ++y)
        for (int x = h + 1; x < X - h; ++x) {
            double S1 = 0.0, S2 = 0.0, G = 0.0;
            for (int b = 0; b < ws; ++b)
                for (int a = 0; a < ws; ++a) {
                    const double v = frame[(size_t)(y - h + b) * nx + (x - h + a)];
                    S1 += v;
                    S2 += v * v;
                    G += v * gc[a + ws * b];
                }
            const double m = S1 / N;
            const double V = S2 - m * m * N;
            const double alpha = G / ssg;
            const double T = alpha * alpha / V * ssg;
            if (test_out) test_out[(size_t)y * nx + x] = T;
            if (T < 1.0 && T > thr) mask[(size_t)y * nx + x] = 1;
        }
    return 0;
}

int orc_find_particles(const uint16_t *frame, const uint8_t *mask, int nx, int ny, double *out, int max_out) {
    const size_t npx = (size_t)nx * ny;
    int32_t *label = (int32_t *)malloc(npx * sizeof(int32_t));
    int32_t *stack = (int32_t *)malloc(npx * sizeof(int32_t));
    if (!label || !stack) {
        free(label);
        free(stack);
        return -1;
    }
    for (size_t k = 0; k < npx; ++k) label[k] = -1;
    double bg = 0.0;
    size_t nbg = 0;
    int count = 0;
    /* raster order, y outer (source/ParticleFinding.cpp:303-309) */
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const size_t p = (size_t)y * nx + x;
            if (!mask[p]) {
                bg += frame[p];
                ++nbg;
                continue;
            }
            if (label[p] != -1) continue;
            /* flood the 8-connected component of p */
            double sx = 0.0, sy = 0.0, mx = 0.0;
            size_t n = 0, top = 0;
            label[p] = count;
            stack[top++] = (int32_t)p;
            while (top > 0) {
                const int32_t q = stack[--top];
                const int qx = q % nx, qy = q / nx;
                sx += qx;
                sy += qy;
                if (frame[q] > mx) mx = frame[q];
                ++n;
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int ux = qx + dx, uy = qy + dy;
                        if ((dx == 0 && dy == 0) || ux < 0 || uy < 0 || ux >= nx || uy >= ny) continue;
                        const size_t u = (size_t)uy * nx + ux;
                        if (!mask[u] || label[u] != -1) continue;
                        label[u] = count;
                        stack[top++] = (int32_t)u;
                    }
            }
            if (count < max_out) {
                out[4 * count + 0] = sx / (double)n;
                out[4 * count + 1] = sy / (double)n;
                out[4 * count + 2] = mx;
                out[4 * count + 3] = 0.0;
            }
            ++count;
        }
    bg /= (double)nbg; /* NaN when every pixel is segmented, as in the reference */
    const int m = count < max_out ? count : max_out;
    for (int k = 0; k < m; ++k) {
        out[4 * k + 2] -= bg;
        out[4 * k + 3] = bg;
    }
    free(label);
    free(stack);
    return count;
}
