// Golden-vector harness for the particle finder: calls the reference's
// ParticleFinder_adjacent8::findPositions (source/ParticleFinding.cpp:283-373, compiled
// from /root/reference by oracle/ref/Makefile) on given frames and threshold masks.
// TEST INFRASTRUCTURE ONLY.
//
// usage: ref_finder <in.bin> <out.bin>
//   in.bin : int32 nframes, nx, ny; per frame uint16 pixels[ny][nx] (x fastest), then
//            uint8 mask[ny][nx] (non-zero = segmented)
//   out.bin: per frame int32 count, then count x (x, y, intensity, background) doubles
#include "ParticleFinding.h"
#include <cstdio>
#include <cstdlib>
#include <vector>

static void rd(FILE *f, void *p, size_t n) {
    if (fread(p, 1, n, f) != n) { fprintf(stderr, "short read\n"); exit(2); }
}

int main(int argc, char **argv) {
    if (argc != 3) { fprintf(stderr, "usage: %s in out\n", argv[0]); return 2; }
    FILE *in = fopen(argv[1], "rb"), *out = fopen(argv[2], "wb");
    if (!in || !out) return 2;
    int32_t nf, nx, ny;
    rd(in, &nf, 4); rd(in, &nx, 4); rd(in, &ny, 4);
    ParticleFinder_adjacent8 finder;
    std::vector<uint16_t> px((size_t)nx * ny);
    std::vector<uint8_t> mk((size_t)nx * ny);
    for (int f = 0; f < nf; ++f) {
        rd(in, px.data(), px.size() * 2);
        rd(in, mk.data(), mk.size());
        ImagePtr image(new Image(nx, ny));
        std::shared_ptr<Eigen::Matrix<int, Eigen::Dynamic, Eigen::Dynamic> > thr(
            new Eigen::Matrix<int, Eigen::Dynamic, Eigen::Dynamic>(nx, ny));
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                (*image)(x, y) = px[(size_t)y * nx + x];
                (*thr)(x, y) = mk[(size_t)y * nx + x] ? 255 : 0;
            }
        std::shared_ptr<std::list<Particle> > parts = finder.findPositions(image, thr);
        int32_t c = (int32_t)parts->size();
        fwrite(&c, 4, 1, out);
        for (const Particle &p : *parts) {
            double v[4] = {p.x, p.y, p.intensity, p.background};
            fwrite(v, 8, 4, out);
        }
    }
    fclose(out);
    return 0;
}
