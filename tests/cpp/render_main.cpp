// Driver for the C++ drop-in renderer (localizer_amd/dropin/PALMImagesHIP.h), used by
// tests/test_dropin.py.  TEST INFRASTRUCTURE.
//
// usage: render_main <method> <p0> <p1> <p2> <weighing> <xSize> <ySize> <scale> <in.bin> <out.bin>
//   in.bin : int32 n; n x (x, y, integral, xDeviation, background) doubles, stored in a
//            LocalizedPositionsContainer_2DGauss
//   out.bin: int32 rows, cols; doubles image (column-major)
#include <cstdio>
#include <cstdlib>

#include "../../localizer_amd/dropin/PALMImagesHIP.h"

int main(int argc, char **argv) {
    if (argc != 11) {
        fprintf(stderr, "usage\n");
        return 2;
    }
    const int method = atoi(argv[1]);
    const double p0 = atof(argv[2]), p1 = atof(argv[3]), p2 = atof(argv[4]);
    const int weighing = atoi(argv[5]);
    const size_t xs = (size_t)atol(argv[6]), ys = (size_t)atol(argv[7]);
    const double scale = atof(argv[8]);
    FILE *in = fopen(argv[9], "rb");
    if (!in) return 2;
    int32_t n = 0;
    if (fread(&n, 4, 1, in) != 1) return 2;
    std::shared_ptr<LocalizedPositionsContainer_2DGauss> c(new LocalizedPositionsContainer_2DGauss());
    for (int k = 0; k < n; ++k) {
        double v[5];
        if (fread(v, 8, 5, in) != 5) return 2;
        std::shared_ptr<LocalizedPosition_2DGauss> p(new LocalizedPosition_2DGauss());
        p->xPosition = v[0];
        p->yPosition = v[1];
        p->integral = v[2];
        p->xPositionDeviation = v[3];
        p->background = v[4];
        c->addPosition(p);
    }
    fclose(in);
    std::shared_ptr<PALMBitmapImageDeviationCalculatorHIP> dev;
    if (method == LZR_DEVIATION_CONSTANT)
        dev.reset(new PALMBitmapImageDeviationCalculator_ConstantHIP(p0));
    else if (method == LZR_DEVIATION_FIT_UNCERTAINTY)
        dev.reset(new PALMBitmapImageDeviationCalculator_FitUncertaintyHIP(p0, p1));
    else
        dev.reset(new PALMBitmapImageDeviationCalculator_GaussianMaskHIP(p0, p1, p2));
    ImagePtr img;
    try {
        PALMBitmapImageCalculatorHIP calc(dev, weighing);
        img = calc.CalculateImage(c, xs, ys, scale);
    } catch (const std::exception &e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 3;
    }
    FILE *out = fopen(argv[10], "wb");
    int32_t r = (int32_t)img->rows(), cc = (int32_t)img->cols();
    fwrite(&r, 4, 1, out);
    fwrite(&cc, 4, 1, out);
    fwrite(img->data(), 8, (size_t)r * cc, out);
    fclose(out);
    return 0;
}
