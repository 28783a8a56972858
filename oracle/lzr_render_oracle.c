/*
 * lzr_render_oracle.c -- FP64 CPU oracle of the localization bitmap.  TEST
 * INFRASTRUCTURE ONLY: the tests use it to check the GPU renderer; the product never
 * calls it.
 *
 *   orc_render  PALMBitmapImageCalculator::CalculateImage (source/PALMImages.cpp:68-157)
 *               with the deviation calculators of source/PALMImages.h:106-160 and
 *               source/PALMImages.cpp:159-177.
 *
 * Every localization adds a Gaussian of standard deviation dev, integrated over each
 * output pixel, to an image of (int)(xSize*scale) x (int)(ySize*scale) pixels (x fast):
 *   sf = (width-1)/(xSize-1), pixel i spans [i/sf - 1/(2 sf), i/sf + 1/(2 sf)),
 *   contribution 0.25 (E(minX)E(minY) - E(maxX)E(minY) - E(minX)E(maxY) + E(maxX)E(maxY))
 *   with E(v) = erf((v - x0)/s), s = sqrt(2 dev^2), over the pixels
 *   [floor(c - 4 dev sf), ceil(c + 4 dev sf)] (c = x0 sf) clamped to the image,
 *   times 1 (weighing "same") or the fitted integral (weighing "integral", positions
 *   with integral 0 skipped).  E is the reference's lookup table (ERFLookupTable,
 *   source/PALMImages.cpp:34-66): erf at -4 + i 1e-4, i = 0..80000, linear interpolation,
 *   -1 / +1 outside [-4, 4].  Pixels accumulate in position order.
 * Deviation: constant p0; fit uncertainty: xDev*p0 if xDev < p1 else -1 (no pixel);
 * Gaussian mask: sqrt((w^2 + 1/12) / I') with I' = max(0, floor(I/m + 0.5)) (p0 = w,
 * p1 = camera offset, p2 = m; an error if the integral or background is -1).
 * A bound that is not a finite int becomes INT_MIN, the value the reference's
 * double-to-int conversion gives on x86-64: such a position adds no pixel.
 */
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "lzr_oracle.h"

#define ERF_LOWER (-4.0)
#define ERF_UPPER 4.0
#define ERF_STRIDE 0.0001

int orc_erf_table_size(void) { return (int)((ERF_UPPER - ERF_LOWER) / ERF_STRIDE + 1); }

/* table[0 .. size-1] = erf(lower + i*stride); table[size] = table[size-1] (read with
 * weight 0 at x = upper exactly). */
void orc_erf_table(double *table) {
    const int n = orc_erf_table_size();
    for (int i = 0; i < n; ++i) table[i] = erf(ERF_LOWER + (double)i * ERF_STRIDE);
    table[n] = table[n - 1];
}

double orc_erf_lookup(const double *table, double x) {
    if (x != x) return x; /* NaN (a zero deviation): no table index */
    if (x < ERF_LOWER) return -1.0;
    if (x > ERF_UPPER) return 1.0;
    const double fi = (x - ERF_LOWER) / ERF_STRIDE;
    const int index = (int)floor(fi);
    const double fraction = fi - (double)index;
    const double a = table[index] * (1.0 - fraction);
    const double b = table[index + 1] * fraction;
    return a + b;
}

static int x86_int(double v) { return (v >= -2147483648.0 && v < 2147483648.0) ? (int)v : INT_MIN; }

int orc_render_size(int x_size, int y_size, double scale, int *width, int *height) {
    if (!(scale > 0.0) || x_size < 2 || y_size < 2) return -1;
    *width = (int)((double)x_size * scale);
    *height = (int)((double)y_size * scale);
    return (*width >= 2 && *height >= 2) ? 0 : -1;
}

/* pos: n rows of (x, y, integral, x_deviation, background).  image: width*height, x fast. */
int orc_render(const double *pos, int n, int x_size, int y_size, double scale, int method, const double *p,
               int weighing, double *image) {
    int width, height;
    if (orc_render_size(x_size, y_size, scale, &width, &height) != 0) return -1;
    if (method < 0 || method > 2 || weighing < 0 || weighing > 1) return -1;
    static double *table = NULL;
    if (!table) {
        double *t = (double *)malloc(sizeof(double) * (size_t)(orc_erf_table_size() + 1));
        orc_erf_table(t);
        table = t;
    }
    for (size_t k = 0; k < (size_t)width * height; ++k) image[k] = 0.0;
    const double sf = (double)(width - 1) / (double)(x_size - 1);
    const double sps = 1.0 / sf;
    const double half = 0.5 * sps;
    for (int k = 0; k < n; ++k) {
        const double x0 = pos[5 * k + 0], y0 = pos[5 * k + 1], integral = pos[5 * k + 2];
        const double xdev = pos[5 * k + 3], background = pos[5 * k + 4];
        if (integral == 0.0 && weighing == 1) continue;
        double dev;
        if (method == 0) {
            dev = p[0];
        } else if (method == 1) {
            dev = xdev < p[1] ? xdev * p[0] : -1.0;
        } else {
            if (integral == -1.0 || background == -1.0) return -1;
            double bg = floor((background - p[1]) / p[2] + 0.5);
            bg = bg < 0.0 ? 0.0 : bg;
            (void)bg; /* computed by the reference; its term is commented out there */
            double I = floor(integral / p[2] + 0.5);
            I = I < 0.0 ? 0.0 : I;
            const double w2 = p[0] * p[0];
            dev = sqrt((w2 + 1.0 / 12.0) / I);
        }
        const double s = sqrt(2.0 * dev * dev);
        const double weight = weighing == 0 ? 1.0 : integral;
        const double cr = x0 * sf, cc = y0 * sf;
        int r0 = x86_int(floor(cr - 4.0 * dev * sf));
        int c0 = x86_int(floor(cc - 4.0 * dev * sf));
        int r1 = x86_int(ceil(cr + 4.0 * dev * sf));
        int c1 = x86_int(ceil(cc + 4.0 * dev * sf));
        if (r0 < 0) r0 = 0;
        if (r1 >= width) r1 = width - 1;
        if (c0 < 0) c0 = 0;
        if (c1 >= height) c1 = height - 1;
        for (int j = c0; j <= c1; ++j) {
            for (int i = r0; i <= r1; ++i) {
                const double minX = (double)i * sps - half;
                const double maxX = minX + sps;
                const double minY = (double)j * sps - half;
                const double maxY = minY + sps;
                const double eminx = orc_erf_lookup(table, (minX - x0) / s);
                const double emaxx = orc_erf_lookup(table, (maxX - x0) / s);
                const double eminy = orc_erf_lookup(table, (minY - y0) / s);
                const double emaxy = orc_erf_lookup(table, (maxY - y0) / s);
                const double t1 = eminx * eminy, t2 = emaxx * eminy, t3 = eminx * emaxy, t4 = emaxx * emaxy;
                const double v = 0.25 * (((t1 - t2) - t3) + t4);
                const double add = v * weight;
                image[(size_t)j * width + i] = image[(size_t)j * width + i] + add;
            }
        }
    }
    return 0;
}
