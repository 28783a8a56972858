/*
 * LocalizationHIP.h -- drop-in replacements for the reference's LM fitters, running on
 * the MI355X through the C ABI of include/localizer_hip.h.
 *
 *   FitPositions_SymmetricGaussianHIP(initialPSFWidth, sigma)      <- FitPositions_SymmetricGaussian
 *   FitPositions_EllipsoidalGaussianHIP(initialPSFWidth, sigma)    <- FitPositions_EllipsoidalGaussian
 *   FitPositions_EllipsoidalGaussian_SymmetricPSFHIP(w, sigma)     <- FitPositions_EllipsoidalGaussian_SymmetricPSF
 *   FitPositions_FixedWidthGaussianHIP(initialPSFWidth, sigma)     <- FitPositions_FixedWidthGaussian
 *   (source/Localization.h:55-121)
 *
 * Same constructor arguments, the same virtual fit_positions(ImagePtr, list<Particle>)
 * returning the same concrete container type, the same side effect (rejected particles
 * are erased from the caller's list; for the symmetric-PSF variant only the ellipsoidal
 * fit's rejections are, source/Localization.cpp:606-631), the same error convention
 * (std::runtime_error, source/Errors.h).  Instances are safe to call from several
 * threads at once (PALMAnalysisController shares one fitter between its workers,
 * source/PALMAnalysis.cpp:234): concurrent fit_positions calls on one instance are
 * gathered into device batches by the library's host batcher (lzr_batcher_*), which
 * returns each frame the records a one-frame call would.
 *
 * Compile inside the reference code base with LZR_WITH_LOCALIZER_HEADERS defined (the
 * reference's Localization.h supplies FitPositions, Particle, Image and the
 * containers); without it, the mirror types of StorageHIP.h are used.
 *
 * fit_positions_batch() is the throughput form: many frames, one device launch.
 */
#ifndef LZR_LOCALIZATION_HIP_H
#define LZR_LOCALIZATION_HIP_H

#include <cmath>
#include <cstdint>
#include <list>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef LZR_WITH_LOCALIZER_HEADERS
#include "Localization.h"
#else
#include "StorageHIP.h"
#endif

#include "../../include/localizer_hip.h"

namespace lzr_dropin {

inline std::string last_error(const LzrContext *ctx) {
    char buf[512];
    lzr_last_error(ctx, buf, (int)sizeof buf);
    return std::string(buf);
}

/* True when every pixel of an FP64 frame is an integer camera count in [0, 65535]: such
 * frames travel as uint16, others as float (lzr_fit_batch_f32). */
inline bool counts_only(const Image &img) {
    const long nx = img.rows(), ny = img.cols();
    for (long y = 0; y < ny; ++y)
        for (long x = 0; x < nx; ++x) {
            const double v = img(x, y);
            if (!(v >= 0.0 && v <= 65535.0) || v != std::floor(v)) return false;
        }
    return true;
}

/* Appends an FP64 frame (column-major, rows = x, i.e. x fast) to a frame stack of T. */
template <class T>
inline void append_frame(const Image &img, std::vector<T> &dst) {
    const long nx = img.rows(), ny = img.cols();
    const size_t base = dst.size();
    dst.resize(base + (size_t)nx * ny);
    for (long y = 0; y < ny; ++y)
        for (long x = 0; x < nx; ++x) dst[base + (size_t)y * nx + x] = (T)img(x, y);
}

class FitterCore {
public:
    FitterCore(int model, double initialPSFWidth, double sigma) : model_(model) {
        if (lzr_create(&ctx_, model, initialPSFWidth, sigma, 0) != 0)
            throw std::runtime_error("LocalizationHIP: " + last_error(nullptr));
    }
    ~FitterCore() {
        if (batcher_) lzr_batcher_destroy(batcher_);
        if (ctx_) lzr_destroy(ctx_);
    }
    FitterCore(const FitterCore &) = delete;
    FitterCore &operator=(const FitterCore &) = delete;

    LzrContext *context() { return ctx_; }

    /* Fits several frames (same size) with their particle lists in one launch. */
    template <class Container>
    std::vector<std::shared_ptr<LocalizedPositionsContainer> > fit_many(
        const std::vector<ImagePtr> &images, const std::vector<std::shared_ptr<std::list<Particle> > > &lists) {
        if (images.size() != lists.size()) throw std::runtime_error("LocalizationHIP: frames and lists differ in count");
        std::vector<std::shared_ptr<LocalizedPositionsContainer> > out;
        std::vector<uint16_t> stack;
        std::vector<float> stackf;
        std::vector<LzrCandidate> cands;
        long nx = -1, ny = -1;
        bool counts = true;
        for (size_t f = 0; f < images.size(); ++f)
            if (!lists[f]->empty() && !counts_only(*images[f])) counts = false;
        for (size_t f = 0; f < images.size(); ++f) {
            out.push_back(std::shared_ptr<LocalizedPositionsContainer>(new Container()));
            if (lists[f]->empty()) continue;
            if (nx < 0) {
                nx = images[f]->rows();
                ny = images[f]->cols();
            } else if (images[f]->rows() != nx || images[f]->cols() != ny) {
                throw std::runtime_error("LocalizationHIP: frames in one batch must share their size");
            }
            const int fi = (int)((counts ? stack.size() : stackf.size()) / ((size_t)nx * ny));
            if (counts)
                append_frame(*images[f], stack);
            else
                append_frame(*images[f], stackf);
            for (const Particle &p : *lists[f]) cands.push_back(candidate(p, fi));
        }
        if (cands.empty()) return out;
        std::vector<LzrFit> res(cands.size());
        const int nframes = (int)((counts ? stack.size() : stackf.size()) / ((size_t)nx * ny));
        const int rc = counts ? lzr_fit_batch(ctx_, stack.data(), (int)nx, (int)ny, nframes, cands.data(),
                                              (int)cands.size(), res.data())
                              : lzr_fit_batch_f32(ctx_, stackf.data(), (int)nx, (int)ny, nframes, cands.data(),
                                                  (int)cands.size(), res.data());
        if (rc < 0) throw std::runtime_error("LocalizationHIP: " + last_error(ctx_));
        size_t k = 0;
        for (size_t f = 0; f < images.size(); ++f) {
            auto it = lists[f]->begin();
            while (it != lists[f]->end()) {
                const LzrFit &r = res[k++];
                if (r.status == LZR_ACCEPTED) {
                    add_record(*out[f], r);
                    ++it;
                } else if (erases(r.status)) {
                    it = lists[f]->erase(it);
                } else {
                    ++it;
                }
            }
        }
        return out;
    }

    /* One frame (the reference's fit_positions): through the batcher, created for the
     * first frame's size; frames of another size, more particles than one batch holds, or
     * windows holding non-integer pixel values take the one-call path (float frames). */
    template <class Container>
    std::shared_ptr<LocalizedPositionsContainer> fit_one(const ImagePtr image,
                                                         std::shared_ptr<std::list<Particle> > particles) {
        LzrBatcher *b = batcher_for(image->rows(), image->cols());
        if (!b || particles->size() > (size_t)kBatchRois) {
            std::vector<ImagePtr> im(1, image);
            std::vector<std::shared_ptr<std::list<Particle> > > li(1, particles);
            return fit_many<Container>(im, li)[0];
        }
        std::shared_ptr<LocalizedPositionsContainer> out(new Container());
        if (particles->empty()) return out;
        std::vector<LzrCandidate> cands;
        cands.reserve(particles->size());
        for (const Particle &p : *particles) cands.push_back(candidate(p, 0));
        std::vector<LzrFit> res(cands.size());
        if (lzr_batcher_fit_frame_f64(b, image->data(), (int)image->rows(), (int)image->cols(), cands.data(),
                                      (int)cands.size(), res.data()) < 0) {
            std::vector<ImagePtr> im(1, image);
            std::vector<std::shared_ptr<std::list<Particle> > > li(1, particles);
            return fit_many<Container>(im, li)[0];
        }
        size_t k = 0;
        for (auto it = particles->begin(); it != particles->end();) {
            const LzrFit &r = res[k++];
            if (r.status == LZR_ACCEPTED) {
                add_record(*out, r);
                ++it;
            } else if (erases(r.status)) {
                it = particles->erase(it);
            } else {
                ++it;
            }
        }
        return out;
    }

private:
    static constexpr int kBatchRois = 1 << 16;

    static LzrCandidate candidate(const Particle &p, int frame) {
        LzrCandidate c;
        c.x = p.x;
        c.y = p.y;
        c.intensity = p.intensity;
        c.background = p.background;
        c.frame = frame;
        c.half_width = 0;
        c.model = -1;
        return c;
    }

    LzrBatcher *batcher_for(long nx, long ny) {
        std::lock_guard<std::mutex> lock(mu_);
        if (!batcher_ && !batcher_failed_) {
            if (lzr_batcher_create(&batcher_, ctx_, (int)nx, (int)ny, 64, kBatchRois, 500, 4) != 0) {
                batcher_ = nullptr;
                batcher_failed_ = true;
            } else {
                bnx_ = nx;
                bny_ = ny;
            }
        }
        return batcher_ && nx == bnx_ && ny == bny_ ? batcher_ : nullptr;
    }

    /* The symmetric-PSF filter's rejections stay in the caller's list (the reference
     * erases only the ellipsoidal fit's rejections). */
    bool erases(int status) const {
        if (model_ == LZR_MODEL_ELLIPSOIDAL_SYMMETRIC_PSF)
            return status != LZR_REJECT_ANISOTROPIC && status != LZR_REJECT_PSF_WIDTH;
        return true;
    }

    void add_record(LocalizedPositionsContainer &c, const LzrFit &r) const {
        if (model_ == LZR_MODEL_SYMMETRIC) {
            std::shared_ptr<LocalizedPosition_2DGauss> p(new LocalizedPosition_2DGauss());
            p->integral = r.integral;
            p->width = r.params[1];
            p->xPosition = r.params[2];
            p->yPosition = r.params[3];
            p->background = r.params[4];
            p->integralDeviation = r.integral_deviation;
            p->widthDeviation = r.deviations[1];
            p->xPositionDeviation = r.deviations[2];
            p->yPositionDeviation = r.deviations[3];
            p->backgroundDeviation = r.deviations[4];
            c.addPosition(p);
        } else if (model_ == LZR_MODEL_FIXED_WIDTH) {
            std::shared_ptr<LocalizedPosition_2DGaussFixedWidth> p(new LocalizedPosition_2DGaussFixedWidth());
            p->integral = r.integral;
            p->xPosition = r.params[1];
            p->yPosition = r.params[2];
            p->background = r.params[3];
            p->integralDeviation = r.integral_deviation;
            p->xPositionDeviation = r.deviations[1];
            p->yPositionDeviation = r.deviations[2];
            p->backgroundDeviation = r.deviations[3];
            c.addPosition(p);
        } else {
            std::shared_ptr<LocalizedPosition_Ellipsoidal2DGauss> p(new LocalizedPosition_Ellipsoidal2DGauss());
            p->integral = r.integral;
            p->stdDev1 = r.sd1;
            p->stdDev2 = r.sd2;
            p->xPosition = r.params[3];
            p->yPosition = r.params[4];
            p->theta = r.theta;
            p->background = r.params[6];
            p->integralDeviation = r.integral_deviation;
            p->stdDev1Deviation = r.sd1_deviation;
            p->stdDev2Deviation = r.sd2_deviation;
            p->xPositionDeviation = r.deviations[3];
            p->yPositionDeviation = r.deviations[4];
            p->thetaDeviation = 0.0;
            p->backgroundDeviation = r.deviations[6];
            c.addPosition(p);
        }
    }

    LzrContext *ctx_ = nullptr;
    int model_;
    std::mutex mu_;
    LzrBatcher *batcher_ = nullptr;
    bool batcher_failed_ = false;
    long bnx_ = 0, bny_ = 0;
};

}  // namespace lzr_dropin

class FitPositions_SymmetricGaussianHIP : public FitPositions {
public:
    FitPositions_SymmetricGaussianHIP(double initialPSFWidth_rhs, double sigma_rhs)
        : core_(LZR_MODEL_SYMMETRIC, initialPSFWidth_rhs, sigma_rhs) {}
    std::shared_ptr<LocalizedPositionsContainer> fit_positions(const ImagePtr image,
                                                               std::shared_ptr<std::list<Particle> > particles) {
        return core_.fit_one<LocalizedPositionsContainer_2DGauss>(image, particles);
    }
    std::vector<std::shared_ptr<LocalizedPositionsContainer> > fit_positions_batch(
        const std::vector<ImagePtr> &images, const std::vector<std::shared_ptr<std::list<Particle> > > &lists) {
        return core_.fit_many<LocalizedPositionsContainer_2DGauss>(images, lists);
    }
    LzrContext *context() { return core_.context(); }

private:
    lzr_dropin::FitterCore core_;
};

class FitPositions_EllipsoidalGaussianHIP : public FitPositions {
public:
    FitPositions_EllipsoidalGaussianHIP(double initialPSFWidth_rhs, double sigma_rhs)
        : core_(LZR_MODEL_ELLIPSOIDAL, initialPSFWidth_rhs, sigma_rhs) {}
    std::shared_ptr<LocalizedPositionsContainer> fit_positions(const ImagePtr image,
                                                               std::shared_ptr<std::list<Particle> > particles) {
        return core_.fit_one<LocalizedPositionsContainer_Ellipsoidal2DGaussian>(image, particles);
    }
    std::vector<std::shared_ptr<LocalizedPositionsContainer> > fit_positions_batch(
        const std::vector<ImagePtr> &images, const std::vector<std::shared_ptr<std::list<Particle> > > &lists) {
        return core_.fit_many<LocalizedPositionsContainer_Ellipsoidal2DGaussian>(images, lists);
    }
    LzrContext *context() { return core_.context(); }

private:
    lzr_dropin::FitterCore core_;
};

class FitPositions_EllipsoidalGaussian_SymmetricPSFHIP : public FitPositions {
public:
    FitPositions_EllipsoidalGaussian_SymmetricPSFHIP(double initialPSFWidth_rhs, double sigma_rhs)
        : core_(LZR_MODEL_ELLIPSOIDAL_SYMMETRIC_PSF, initialPSFWidth_rhs, sigma_rhs) {}
    std::shared_ptr<LocalizedPositionsContainer> fit_positions(const ImagePtr image,
                                                               std::shared_ptr<std::list<Particle> > particles) {
        return core_.fit_one<LocalizedPositionsContainer_Ellipsoidal2DGaussian>(image, particles);
    }
    std::vector<std::shared_ptr<LocalizedPositionsContainer> > fit_positions_batch(
        const std::vector<ImagePtr> &images, const std::vector<std::shared_ptr<std::list<Particle> > > &lists) {
        return core_.fit_many<LocalizedPositionsContainer_Ellipsoidal2DGaussian>(images, lists);
    }
    LzrContext *context() { return core_.context(); }

private:
    lzr_dropin::FitterCore core_;
};

class FitPositions_FixedWidthGaussianHIP : public FitPositions {
public:
    FitPositions_FixedWidthGaussianHIP(double initialPSFWidth_rhs, double sigma_rhs)
        : core_(LZR_MODEL_FIXED_WIDTH, initialPSFWidth_rhs, sigma_rhs) {}
    std::shared_ptr<LocalizedPositionsContainer> fit_positions(const ImagePtr image,
                                                               std::shared_ptr<std::list<Particle> > particles) {
        return core_.fit_one<LocalizedPositionsContainer_2DGaussFixedWidth>(image, particles);
    }
    std::vector<std::shared_ptr<LocalizedPositionsContainer> > fit_positions_batch(
        const std::vector<ImagePtr> &images, const std::vector<std::shared_ptr<std::list<Particle> > > &lists) {
        return core_.fit_many<LocalizedPositionsContainer_2DGaussFixedWidth>(images, lists);
    }
    LzrContext *context() { return core_.context(); }

private:
    lzr_dropin::FitterCore core_;
};

#endif
