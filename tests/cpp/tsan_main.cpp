// Thread-sanitizer driver for the oracle's frame-parallel pool (oracle/lzr_oracle.c,
// orc_fit_batch: one frame per task on pthreads, as the reference's analysis controller)
// and for concurrent oracle calls from several threads (the finder and the renderer are
// called per frame by the reference's workers).  TEST INFRASTRUCTURE: built with
// -fsanitize=thread by `python -m localizer_amd.build sanitize`, run by
// tests/test_sanitizers.py.  Exit 0 when the pooled fit equals the single-threaded fit and
// the concurrent calls equal their serial results.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "../../oracle/lzr_oracle.h"

static uint32_t lcg(uint32_t &s) {
    s = s * 1664525u + 1013904223u;
    return s >> 8;
}

int main() {
    const int nx = 64, ny = 64, nframes = 24, per = 12;
    std::vector<uint16_t> frames((size_t)nx * ny * nframes);
    std::vector<LzrCandidate> cands;
    uint32_t seed = 12345u;
    for (int f = 0; f < nframes; ++f) {
        std::vector<double> img((size_t)nx * ny, 100.0);
        for (int e = 0; e < per; ++e) {
            const double x = 8.0 + (lcg(seed) % 4800) / 100.0, y = 8.0 + (lcg(seed) % 4800) / 100.0;
            const double amp = 200.0 + lcg(seed) % 1500;
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const double d2 = (i - x) * (i - x) + (j - y) * (j - y);
                    img[(size_t)j * nx + i] += amp * std::exp(-d2 / (2.0 * 1.3 * 1.3));
                }
            LzrCandidate c;
            std::memset(&c, 0, sizeof c);
            c.x = std::floor(x) + 0.5;
            c.y = std::floor(y) + 0.5;
            c.intensity = amp;
            c.background = 100.0;
            c.frame = f;
            c.half_width = 0;
            c.model = -1;
            cands.push_back(c);
        }
        for (size_t k = 0; k < img.size(); ++k)
            frames[(size_t)f * nx * ny + k] = (uint16_t)std::lround(img[k] + (double)(lcg(seed) % 21) - 10.0);
    }
    const int n = (int)cands.size();
    int bad = 0;
    for (int model : {LZR_MODEL_SYMMETRIC, LZR_MODEL_ELLIPSOIDAL}) {
        OrcSettings s;
        std::memset(&s, 0, sizeof s);
        s.model = model;
        s.initial_psf_width = 1.3;
        s.sigma = 1.0;
        s.default_half_width = 0;
        orc_default_options(&s.opt);
        std::vector<LzrFit> one(n), pool(n);
        orc_fit_batch(&s, frames.data(), nx, ny, nframes, cands.data(), n, one.data(), 1);
        orc_fit_batch(&s, frames.data(), nx, ny, nframes, cands.data(), n, pool.data(), 8);
        if (std::memcmp(one.data(), pool.data(), sizeof(LzrFit) * n) != 0) {
            std::fprintf(stderr, "model %d: pooled fit differs from the single-threaded fit\n", model);
            ++bad;
        }
    }
    /* the finder and the renderer from several threads at once, each on its own frame */
    std::vector<std::vector<uint8_t>> masks(nframes, std::vector<uint8_t>((size_t)nx * ny));
    std::vector<std::vector<uint8_t>> masks1 = masks;
    std::vector<int> counts(nframes), counts1(nframes);
    for (int f = 0; f < nframes; ++f)
        counts1[f] = orc_glrt_mask(frames.data() + (size_t)f * nx * ny, nx, ny, 1.3, 20.0, masks1[f].data(), nullptr);
    std::vector<std::thread> th;
    for (int t = 0; t < 6; ++t)
        th.emplace_back([&, t]() {
            for (int f = t; f < nframes; f += 6)
                counts[f] = orc_glrt_mask(frames.data() + (size_t)f * nx * ny, nx, ny, 1.3, 20.0, masks[f].data(),
                                          nullptr);
        });
    for (auto &x : th) x.join();
    for (int f = 0; f < nframes; ++f)
        if (counts[f] != counts1[f] || masks[f] != masks1[f]) {
            std::fprintf(stderr, "frame %d: concurrent GLRT mask differs\n", f);
            ++bad;
        }
    std::printf("tsan driver: %d fits x 2 models, %d frames; %s\n", n, nframes, bad ? "MISMATCH" : "ok");
    return bad ? 1 : 0;
}
