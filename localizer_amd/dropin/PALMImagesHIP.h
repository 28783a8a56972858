/*
 * PALMImagesHIP.h -- drop-in replacement for the reference's localization bitmap,
 * running on the MI355X through the C ABI of include/localizer_hip.h.
 *
 *   PALMBitmapImageCalculatorHIP(deviationCalculator, emitterWeighingMethod[, progressReporter])
 *       <- PALMBitmapImageCalculator (source/PALMImages.h:70-88)
 *   PALMBitmapImageDeviationCalculator_ConstantHIP(deviation)
 *       <- PALMBitmapImageDeviationCalculator_Constant (source/PALMImages.h:122-133)
 *   PALMBitmapImageDeviationCalculator_FitUncertaintyHIP(scaleFactor, upperLimit)
 *       <- PALMBitmapImageDeviationCalculator_FitUncertainty (source/PALMImages.h:106-120)
 *   PALMBitmapImageDeviationCalculator_GaussianMaskHIP(PSFWidth, cameraOffset, cameraMultiplicationFactor)
 *       <- PALMBitmapImageDeviationCalculator_GaussianMask (source/PALMImages.h:138-157)
 *
 * CalculateImage(positions, xSize, ySize, outputScaleFactor) returns the same Image
 * (xSize*scale rows x ySize*scale columns, FP64) with the same pixel sums as the
 * reference (source/PALMImages.cpp:68-157) and the same std::runtime_error texts.  The
 * deviation calculators are descriptors (the GPU evaluates them per position), so they
 * take the reference constructors' arguments but are not the reference classes.  The
 * progress reporter is accepted and not called: the image is computed in one launch.
 *
 * Compile inside the reference code base with LZR_WITH_LOCALIZER_HEADERS defined;
 * without it, the mirror types of StorageHIP.h are used.
 */
#ifndef LZR_PALMIMAGES_HIP_H
#define LZR_PALMIMAGES_HIP_H

#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef LZR_WITH_LOCALIZER_HEADERS
#include "PALMImages.h"
#else
#include "StorageHIP.h"
#endif

#include "../../include/localizer_hip.h"

class PALMBitmapImageDeviationCalculatorHIP {
public:
    PALMBitmapImageDeviationCalculatorHIP(int method, double p0, double p1, double p2) : method_(method) {
        p_[0] = p0;
        p_[1] = p1;
        p_[2] = p2;
    }
    virtual ~PALMBitmapImageDeviationCalculatorHIP() {}
    int method() const { return method_; }
    const double *params() const { return p_; }

private:
    int method_;
    double p_[3];
};

class PALMBitmapImageDeviationCalculator_ConstantHIP : public PALMBitmapImageDeviationCalculatorHIP {
public:
    explicit PALMBitmapImageDeviationCalculator_ConstantHIP(double deviation)
        : PALMBitmapImageDeviationCalculatorHIP(LZR_DEVIATION_CONSTANT, deviation, 0.0, 0.0) {}
};

class PALMBitmapImageDeviationCalculator_FitUncertaintyHIP : public PALMBitmapImageDeviationCalculatorHIP {
public:
    PALMBitmapImageDeviationCalculator_FitUncertaintyHIP(double scaleFactor, double upperLimit)
        : PALMBitmapImageDeviationCalculatorHIP(LZR_DEVIATION_FIT_UNCERTAINTY, scaleFactor, upperLimit, 0.0) {}
};

class PALMBitmapImageDeviationCalculator_GaussianMaskHIP : public PALMBitmapImageDeviationCalculatorHIP {
public:
    PALMBitmapImageDeviationCalculator_GaussianMaskHIP(double PSFWidth, double cameraOffset,
                                                       double cameraMultiplicationFactor)
        : PALMBitmapImageDeviationCalculatorHIP(LZR_DEVIATION_GAUSSIAN_MASK, PSFWidth, cameraOffset,
                                                cameraMultiplicationFactor) {}
};

class PALMBitmapImageCalculatorHIP {
public:
    PALMBitmapImageCalculatorHIP(std::shared_ptr<PALMBitmapImageDeviationCalculatorHIP> deviationCalculator,
                                 int emitterWeighingMethod)
        : dev_(deviationCalculator), weighing_(emitterWeighingMethod) {
        if (lzr_renderer_create(&r_) != 0) throw std::runtime_error("PALMImagesHIP: " + error(nullptr));
    }
    template <class ProgressReporterPtr>
    PALMBitmapImageCalculatorHIP(std::shared_ptr<PALMBitmapImageDeviationCalculatorHIP> deviationCalculator,
                                 int emitterWeighingMethod, ProgressReporterPtr)
        : PALMBitmapImageCalculatorHIP(deviationCalculator, emitterWeighingMethod) {}
    ~PALMBitmapImageCalculatorHIP() {
        if (r_) lzr_renderer_destroy(r_);
    }
    PALMBitmapImageCalculatorHIP(const PALMBitmapImageCalculatorHIP &) = delete;
    PALMBitmapImageCalculatorHIP &operator=(const PALMBitmapImageCalculatorHIP &) = delete;

    ImagePtr CalculateImage(std::shared_ptr<LocalizedPositionsContainer> positions, size_t xSize, size_t ySize,
                            double outputScaleFactor) {
        if (outputScaleFactor <= 0.0) throw std::runtime_error("found negative or zero outputScaleFactor");
        int w = 0, h = 0;
        if (lzr_render_image_size((int)xSize, (int)ySize, outputScaleFactor, &w, &h) != 0)
            throw std::runtime_error("PALMImagesHIP: the output image would be smaller than 2x2 pixels");
        const size_t n = positions->getNPositions();
        std::vector<LzrPosition> pos(n);
        for (size_t k = 0; k < n; ++k) {
            pos[k].x = positions->getXPosition(k);
            pos[k].y = positions->getYPosition(k);
            pos[k].integral = positions->getIntegral(k);
            pos[k].x_deviation = positions->getXPositionDeviation(k);
            pos[k].background = positions->getBackground(k);
        }
        ImagePtr image(new Image(w, h));
        std::lock_guard<std::mutex> lock(mu_);
        if (lzr_render(r_, pos.data(), (int)n, (int)xSize, (int)ySize, outputScaleFactor, dev_->method(),
                       dev_->params(), weighing_, image->data()) != 0)
            throw std::runtime_error(error(r_));
        return image;
    }

private:
    static std::string error(const LzrRenderer *r) {
        char buf[512];
        lzr_renderer_last_error(r, buf, (int)sizeof buf);
        return std::string(buf);
    }
    std::shared_ptr<PALMBitmapImageDeviationCalculatorHIP> dev_;
    int weighing_;
    LzrRenderer *r_ = nullptr;
    std::mutex mu_;
};

#endif /* LZR_PALMIMAGES_HIP_H */
