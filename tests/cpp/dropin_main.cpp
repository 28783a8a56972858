// Driver for the C++ drop-in classes (localizer_amd/dropin/LocalizationHIP.h), used by
// tests/test_dropin.py.  TEST INFRASTRUCTURE.
//
// usage: dropin_main <model> <initialPSFWidth> <sigma> <mode> <in.bin> <out.bin>
//   mode: single (one fit_positions call per frame), threads (4 threads sharing one
//         fitter instance), batch (fit_positions_batch over all frames); a "scaled_" prefix
//         fits the frames with pixel values 1.375 v + 0.25 (non-integer: the float path)
//   in.bin : int32 nx, ny, nframes; uint16 frames[nframes][ny][nx];
//            per frame: int32 count, count x (x, y, intensity, background) doubles
//   out.bin: per frame: int32 rows, cols, doubles matrix (column-major);
//            int32 survivors, survivors x (x, y) doubles
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>

#include "../../localizer_amd/dropin/LocalizationHIP.h"

static void rd(FILE *f, void *p, size_t n) {
    if (fread(p, 1, n, f) != n) {
        fprintf(stderr, "short read\n");
        exit(2);
    }
}

int main(int argc, char **argv) {
    if (argc != 7) {
        fprintf(stderr, "usage\n");
        return 2;
    }
    const int model = atoi(argv[1]);
    const double w = atof(argv[2]), sigma = atof(argv[3]);
    std::string mode = argv[4];
    const bool scaled = mode.rfind("scaled_", 0) == 0;
    if (scaled) mode = mode.substr(7);
    FILE *in = fopen(argv[5], "rb");
    if (!in) return 2;
    int32_t nx, ny, nf;
    rd(in, &nx, 4);
    rd(in, &ny, 4);
    rd(in, &nf, 4);
    std::vector<uint16_t> raw((size_t)nx * ny * nf);
    rd(in, raw.data(), raw.size() * 2);
    std::vector<ImagePtr> images;
    std::vector<std::shared_ptr<std::list<Particle> > > lists;
    for (int f = 0; f < nf; ++f) {
        ImagePtr im(new Image(nx, ny));
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const double v = raw[((size_t)f * ny + y) * nx + x];
                (*im)(x, y) = scaled ? 1.375 * v + 0.25 : v;
            }
        images.push_back(im);
        int32_t cnt;
        rd(in, &cnt, 4);
        std::shared_ptr<std::list<Particle> > l(new std::list<Particle>());
        for (int i = 0; i < cnt; ++i) {
            double v[4];
            rd(in, v, 32);
            Particle p(v[0], v[1], v[2]);
            p.background = v[3];
            l->push_back(p);
        }
        lists.push_back(l);
    }
    fclose(in);

    std::vector<std::shared_ptr<LocalizedPositionsContainer> > res(nf);
    try {
        std::unique_ptr<FitPositions> fitter;
        if (model == LZR_MODEL_SYMMETRIC)
            fitter.reset(new FitPositions_SymmetricGaussianHIP(w, sigma));
        else if (model == LZR_MODEL_ELLIPSOIDAL)
            fitter.reset(new FitPositions_EllipsoidalGaussianHIP(w, sigma));
        else if (model == LZR_MODEL_FIXED_WIDTH)
            fitter.reset(new FitPositions_FixedWidthGaussianHIP(w, sigma));
        else
            fitter.reset(new FitPositions_EllipsoidalGaussian_SymmetricPSFHIP(w, sigma));
        if (mode == "single") {
            for (int f = 0; f < nf; ++f) res[f] = fitter->fit_positions(images[f], lists[f]);
        } else if (mode == "threads") {
            std::vector<std::thread> th;
            for (int t = 0; t < 4; ++t)
                th.emplace_back([&, t]() {
                    for (int f = t; f < nf; f += 4) res[f] = fitter->fit_positions(images[f], lists[f]);
                });
            for (auto &x : th) x.join();
        } else {
            if (model == LZR_MODEL_SYMMETRIC)
                res = static_cast<FitPositions_SymmetricGaussianHIP *>(fitter.get())->fit_positions_batch(images, lists);
            else if (model == LZR_MODEL_ELLIPSOIDAL)
                res = static_cast<FitPositions_EllipsoidalGaussianHIP *>(fitter.get())->fit_positions_batch(images, lists);
            else if (model == LZR_MODEL_FIXED_WIDTH)
                res = static_cast<FitPositions_FixedWidthGaussianHIP *>(fitter.get())->fit_positions_batch(images, lists);
            else
                res = static_cast<FitPositions_EllipsoidalGaussian_SymmetricPSFHIP *>(fitter.get())
                          ->fit_positions_batch(images, lists);
        }
        for (int f = 0; f < nf; ++f) res[f]->setFrameNumbers(f);
    } catch (const std::exception &e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 3;
    }
    FILE *out = fopen(argv[6], "wb");
    for (int f = 0; f < nf; ++f) {
        ImagePtr m = res[f]->getLocalizedPositionsAsMatrix();
        int32_t r = (int32_t)m->rows(), c = (int32_t)m->cols();
        fwrite(&r, 4, 1, out);
        fwrite(&c, 4, 1, out);
        fwrite(m->data(), 8, (size_t)r * c, out);
        int32_t s = (int32_t)lists[f]->size();
        fwrite(&s, 4, 1, out);
        for (const Particle &p : *lists[f]) {
            double v[2] = {p.x, p.y};
            fwrite(v, 8, 2, out);
        }
    }
    fclose(out);
    return 0;
}
