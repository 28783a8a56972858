/*
 * lzr_pipeline.hip -- the whole localization pipeline behind one C call: GLRT
 * segmentation, 8-way particle finding and the LM fit on the device, returning the
 * reference's positions matrix.  Host code over the C ABI of include/localizer_hip.h.
 *
 * lzr_fit_emitters is LocalizerFitEmitters (source/LocalizerDLL.cpp:89-129): uint16
 * frames (nRows = x extent, nCols = y extent), ThresholdImage_GLRT_FFT(glrtInsensitivity,
 * psfWidth), ParticleFinder_adjacent8, FitPositions_EllipsoidalGaussian(psfWidth, 1.0),
 * positions sorted by frame (PALMAnalysisController::DoPALMAnalysis,
 * source/PALMAnalysis.cpp:79-190) and returned as the container's matrix
 * (getLocalizedPositionsAsMatrix, column-major, callee-allocated, freed by lzr_free).
 * Unlike the reference entry point it counts the frames it is given (the reference's
 * ImageLoaderRawPointer leaves the base-class image count at 0, so its controller
 * returns no container).
 */
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "../../include/localizer_hip.h"

#define PIPE_CHECK(call)                                                                    \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) throw std::string(#call ": ") + hipGetErrorString(e_);        \
    } while (0)

static thread_local std::string g_pipeline_error;

namespace {

/* One matrix row of the model's container (source/Storage.cpp:268-290, :405-425, :614-640). */
int columns_for(int model) {
    switch (model) {
    case LZR_MODEL_SYMMETRIC: return 12;
    case LZR_MODEL_FIXED_WIDTH: return 10;
    default: return 16;
    }
}

void row_for(int model, const LzrFit &r, double frame, double *row) {
    if (model == LZR_MODEL_SYMMETRIC) {
        const double v[12] = {frame, r.integral, r.params[1], r.params[2], r.params[3], r.params[4],
                              r.integral_deviation, r.deviations[1], r.deviations[2], r.deviations[3],
                              r.deviations[4], 1.0};
        memcpy(row, v, sizeof v);
    } else if (model == LZR_MODEL_FIXED_WIDTH) {
        const double v[10] = {frame, r.integral, r.params[1], r.params[2], r.params[3], r.integral_deviation,
                              r.deviations[1], r.deviations[2], r.deviations[3], 1.0};
        memcpy(row, v, sizeof v);
    } else {
        const double v[16] = {frame, r.integral, r.sd1, r.sd2, r.params[3], r.params[4], r.theta, r.params[6],
                              r.integral_deviation, r.sd1_deviation, r.sd2_deviation, r.deviations[3],
                              r.deviations[4], 0.0, r.deviations[6], 1.0};
        memcpy(row, v, sizeof v);
    }
}

std::string ctx_error(const LzrContext *ctx) {
    char buf[512];
    lzr_last_error(ctx, buf, (int)sizeof buf);
    return buf;
}

std::string finder_error(const LzrFinder *f) {
    char buf[512];
    lzr_finder_last_error(f, buf, (int)sizeof buf);
    return buf;
}

}  // namespace

LZR_EXPORT int lzr_localize(const uint16_t *image_data, int n_rows, int n_cols, int n_images, double psf_width,
                            double glrt_insensitivity, int model, int *n_emitters, int *n_columns, double **matrix) {
    if (!n_emitters || !n_columns || !matrix) {
        g_pipeline_error = "lzr_localize: null output pointer";
        return -1;
    }
    *n_emitters = 0;
    *n_columns = columns_for(model);
    *matrix = nullptr;
    if (n_rows < 1 || n_cols < 1 || n_images < 0 || (n_images > 0 && !image_data)) {
        g_pipeline_error = "lzr_localize: invalid image dimensions";
        return -1;
    }
    LzrContext *ctx = nullptr;
    LzrFinder *finder = nullptr;
    uint16_t *d_frames = nullptr;
    LzrCandidate *d_cands = nullptr;
    LzrFit *d_fits = nullptr;
    int *d_count = nullptr;
    hipStream_t s = nullptr;
    std::vector<double> rows;
    try {
        if (lzr_create(&ctx, model, psf_width, 1.0, 0) != 0) throw "lzr_create: " + ctx_error(nullptr);
        if (lzr_finder_create(&finder, psf_width, glrt_insensitivity, 0) != 0)
            throw "lzr_finder_create: " + finder_error(nullptr);
        const size_t npx = (size_t)n_rows * n_cols;
        const int block = (int)std::max<size_t>(1, std::min<size_t>((size_t)n_images, (size_t)(64u << 20) / npx));
        const int max_cands = block * 16384;
        PIPE_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        PIPE_CHECK(hipMalloc((void **)&d_frames, (size_t)block * npx * sizeof(uint16_t)));
        PIPE_CHECK(hipMalloc((void **)&d_cands, (size_t)max_cands * sizeof(LzrCandidate)));
        PIPE_CHECK(hipMalloc((void **)&d_fits, (size_t)max_cands * sizeof(LzrFit)));
        PIPE_CHECK(hipMalloc((void **)&d_count, sizeof(int)));
        const int cols = columns_for(model);
        std::vector<LzrFit> fits;
        std::vector<LzrCandidate> cands;
        for (int f0 = 0; f0 < n_images; f0 += block) {
            const int nb = std::min(block, n_images - f0);
            PIPE_CHECK(hipMemcpyAsync(d_frames, image_data + (size_t)f0 * npx, (size_t)nb * npx * sizeof(uint16_t),
                                      hipMemcpyHostToDevice, s));
            if (lzr_find_candidates_device(finder, d_frames, n_rows, n_cols, nb, d_cands, max_cands, d_count, s) != 0)
                throw "segmentation: " + finder_error(finder);
            int count = 0;
            PIPE_CHECK(hipMemcpyAsync(&count, d_count, sizeof(int), hipMemcpyDeviceToHost, s));
            PIPE_CHECK(hipStreamSynchronize(s));
            if (count < 0) throw std::string("segmentation: too many segmented pixels in a frame");
            if (count == 0) continue;
            if (lzr_fit_batch_device(ctx, d_frames, n_rows, n_cols, nb, d_cands, count, d_fits, s) != 0)
                throw "fit: " + ctx_error(ctx);
            fits.resize(count);
            cands.resize(count);
            PIPE_CHECK(hipMemcpyAsync(fits.data(), d_fits, (size_t)count * sizeof(LzrFit), hipMemcpyDeviceToHost, s));
            PIPE_CHECK(hipMemcpyAsync(cands.data(), d_cands, (size_t)count * sizeof(LzrCandidate),
                                      hipMemcpyDeviceToHost, s));
            PIPE_CHECK(hipStreamSynchronize(s));
            /* candidates come frame-major, so accepted rows are already sorted by frame */
            for (int k = 0; k < count; ++k) {
                if (fits[k].status != LZR_ACCEPTED) continue;
                rows.resize(rows.size() + cols);
                row_for(model, fits[k], (double)(f0 + cands[k].frame), rows.data() + rows.size() - cols);
            }
        }
        const int n = (int)(rows.size() / cols);
        double *m = new double[(size_t)std::max(n, 1) * cols];
        for (int r = 0; r < n; ++r)
            for (int c = 0; c < cols; ++c) m[(size_t)c * n + r] = rows[(size_t)r * cols + c]; /* column-major */
        *matrix = m;
        *n_emitters = n;
    } catch (const std::string &e) {
        g_pipeline_error = "lzr_localize: " + e;
    }
    if (d_frames) (void)hipFree(d_frames);
    if (d_cands) (void)hipFree(d_cands);
    if (d_fits) (void)hipFree(d_fits);
    if (d_count) (void)hipFree(d_count);
    if (s) (void)hipStreamDestroy(s);
    if (finder) lzr_finder_destroy(finder);
    if (ctx) lzr_destroy(ctx);
    return *matrix ? 0 : -1;
}

LZR_EXPORT int lzr_fit_emitters(const uint16_t *image_data, int n_rows, int n_cols, int n_images, double psf_width,
                                double glrt_insensitivity, int *n_emitters, double **emitter_coordinates) {
    int cols = 0;
    return lzr_localize(image_data, n_rows, n_cols, n_images, psf_width, glrt_insensitivity, LZR_MODEL_ELLIPSOIDAL,
                        n_emitters, &cols, emitter_coordinates);
}

LZR_EXPORT void lzr_free(double *ptr) { delete[] ptr; }

LZR_EXPORT int lzr_pipeline_last_error(char *buf, int len) {
    const std::string &e = g_pipeline_error;
    if (buf && len > 0) {
        const int n = std::min((int)e.size(), len - 1);
        memcpy(buf, e.data(), n);
        buf[n] = 0;
    }
    return (int)e.size();
}
