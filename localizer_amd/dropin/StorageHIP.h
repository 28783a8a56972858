/*
 * StorageHIP.h -- standalone mirror of the reference's candidate and result types.
 *
 * When the drop-in is compiled inside the reference code base (LZR_WITH_LOCALIZER_HEADERS
 * defined, reference include paths set), the reference's own Storage.h / Defines.h
 * types are used instead and this header is not included.  Names, fields and member
 * semantics follow:
 *   Image / ImagePtr                        source/Defines.h:44-45 (column-major, rows = x)
 *   Particle                                source/Storage.h:90-101
 *   LocalizedPosition_2DGauss               source/Storage.h:117-138
 *   LocalizedPosition_2DGaussFixedWidth     source/Storage.h:140-159
 *   LocalizedPosition_Ellipsoidal2DGauss    source/Storage.h:161-186
 *   LocalizedPositionsContainer(_2DGauss, _2DGaussFixedWidth, _Ellipsoidal2DGaussian)
 *                                           source/Storage.h:268-480, source/Storage.cpp:174-345,
 *                                           :405-476, :614-700
 * Only the members the fitting path and its callers use are provided.
 */
#ifndef LZR_STORAGE_HIP_H
#define LZR_STORAGE_HIP_H

#include <algorithm>
#include <cstddef>
#include <fstream>
#include <list>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

/* Column-major FP64 matrix; element (r, c) at data()[r + c*rows()] (Eigen::MatrixXd layout). */
class Image {
public:
    Image() : rows_(0), cols_(0) {}
    Image(long rows, long cols) : rows_(rows), cols_(cols), v_((size_t)(rows * cols), 0.0) {}
    long rows() const { return rows_; }
    long cols() const { return cols_; }
    double &operator()(long r, long c) { return v_[(size_t)(r + c * rows_)]; }
    double operator()(long r, long c) const { return v_[(size_t)(r + c * rows_)]; }
    double *data() { return v_.data(); }
    const double *data() const { return v_.data(); }
    void setConstant(double x) { std::fill(v_.begin(), v_.end(), x); }

private:
    long rows_, cols_;
    std::vector<double> v_;
};
typedef std::shared_ptr<Image> ImagePtr;

const double PI = 3.1415926535897932384626433;

const size_t LOCALIZED_POSITIONS_TYPE_2DGAUSS = 0;
const size_t LOCALIZED_POSITIONS_TYPE_2DGAUSS_FIXED_WIDTH = 1;
const size_t LOCALIZED_POSITIONS_TYPE_ELLIPSOIDAL2DGAUSS = 5;

class Particle {
public:
    Particle() : x(0), y(0), intensity(0), background(0) {}
    Particle(double xLoc, double yLoc) : x(xLoc), y(yLoc), intensity(0), background(0) {}
    Particle(double xLoc, double yLoc, double intensity_rhs)
        : x(xLoc), y(yLoc), intensity(intensity_rhs), background(0) {}
    double x;
    double y;
    double intensity;
    double background;
};

class LocalizedPosition {
public:
    virtual ~LocalizedPosition() {}
    virtual size_t getPositionType() const = 0;
};

class LocalizedPosition_2DGauss : public LocalizedPosition {
public:
    LocalizedPosition_2DGauss() : frameNumber(0), nFramesPresent(1) {}
    size_t frameNumber;
    size_t nFramesPresent;
    double integral;
    double width;
    double xPosition;
    double yPosition;
    double background;
    double integralDeviation;
    double widthDeviation;
    double xPositionDeviation;
    double yPositionDeviation;
    double backgroundDeviation;
    size_t getPositionType() const { return LOCALIZED_POSITIONS_TYPE_2DGAUSS; }
};

class LocalizedPosition_2DGaussFixedWidth : public LocalizedPosition {
public:
    LocalizedPosition_2DGaussFixedWidth() : frameNumber(0), nFramesPresent(1) {}
    size_t frameNumber;
    size_t nFramesPresent;
    double integral;
    double xPosition;
    double yPosition;
    double background;
    double integralDeviation;
    double xPositionDeviation;
    double yPositionDeviation;
    double backgroundDeviation;
    size_t getPositionType() const { return LOCALIZED_POSITIONS_TYPE_2DGAUSS_FIXED_WIDTH; }
};

class LocalizedPosition_Ellipsoidal2DGauss : public LocalizedPosition {
public:
    LocalizedPosition_Ellipsoidal2DGauss() : frameNumber(0), nFramesPresent(1) {}
    size_t frameNumber;
    size_t nFramesPresent;
    double integral;
    double stdDev1;
    double stdDev2;
    double xPosition;
    double yPosition;
    double theta;
    double background;
    double integralDeviation;
    double stdDev1Deviation;
    double stdDev2Deviation;
    double xPositionDeviation;
    double yPositionDeviation;
    double thetaDeviation;
    double backgroundDeviation;
    size_t getPositionType() const { return LOCALIZED_POSITIONS_TYPE_ELLIPSOIDAL2DGAUSS; }
};

class LocalizedPositionsContainer {
public:
    virtual ~LocalizedPositionsContainer() {}
    virtual size_t getNPositions() const = 0;
    virtual size_t getFrameNumber(size_t index) const = 0;
    virtual double getIntegral(size_t index) const { return 0; }
    virtual double getXPosition(size_t index) const { return 0; }
    virtual double getYPosition(size_t index) const { return 0; }
    virtual double getBackground(size_t index) const { return 0; }
    virtual double getXPositionDeviation(size_t index) const { return 0; }
    virtual void addPosition(std::shared_ptr<LocalizedPosition> newPosition) = 0;
    virtual void addPositions(std::shared_ptr<LocalizedPositionsContainer> other) = 0;
    virtual void setFrameNumbers(size_t frameNumber) = 0;
    virtual ImagePtr getLocalizedPositionsAsMatrix() const = 0;
    virtual void writePositionsToFile(std::string filePath, std::string header) const = 0;
    virtual void sortPositionsByFrameNumber() = 0;
    virtual size_t getPositionsType() const = 0;
};

/* Shared implementation of the typed containers. */
template <class Pos, size_t TYPE, int NCOLS>
class LzrTypedContainer : public LocalizedPositionsContainer {
public:
    size_t getNPositions() const { return positionsVector.size(); }
    size_t getFrameNumber(size_t index) const { return positionsVector.at(index).frameNumber; }
    double getXPosition(size_t index) const { return positionsVector.at(index).xPosition; }
    double getYPosition(size_t index) const { return positionsVector.at(index).yPosition; }
    double getBackground(size_t index) const { return positionsVector.at(index).background; }
    double getIntegral(size_t index) const { return positionsVector.at(index).integral; }
    double getXPositionDeviation(size_t index) const { return positionsVector.at(index).xPositionDeviation; }
    const Pos &at(size_t index) const { return positionsVector.at(index); }

    void addPosition(std::shared_ptr<LocalizedPosition> p) {
        if (p->getPositionType() != TYPE)
            throw std::runtime_error("Trying to append a position of a different type to a LocalizedPositionsContainer");
        positionsVector.push_back(*std::static_pointer_cast<Pos>(p));
    }
    void addPositions(std::shared_ptr<LocalizedPositionsContainer> other) {
        if (this == other.get())
            throw std::runtime_error("Trying to append a LocalizedPositionsContainer to itself");
        if (other->getPositionsType() != TYPE)
            throw std::runtime_error("Trying to append a container of a different type to a LocalizedPositionsContainer");
        const LzrTypedContainer *o = static_cast<const LzrTypedContainer *>(other.get());
        positionsVector.insert(positionsVector.end(), o->positionsVector.begin(), o->positionsVector.end());
    }
    void setFrameNumbers(size_t frameNumber) {
        for (auto &p : positionsVector) p.frameNumber = frameNumber;
    }
    void sortPositionsByFrameNumber() {
        std::sort(positionsVector.begin(), positionsVector.end(),
                  [](const Pos &a, const Pos &b) { return a.frameNumber < b.frameNumber; });
    }
    size_t getPositionsType() const { return TYPE; }

protected:
    std::vector<Pos> positionsVector;
};

class LocalizedPositionsContainer_2DGauss
    : public LzrTypedContainer<LocalizedPosition_2DGauss, LOCALIZED_POSITIONS_TYPE_2DGAUSS, 12> {
public:
    double getXWidth(size_t i) const { return positionsVector.at(i).width; }
    double getYWidth(size_t i) const { return positionsVector.at(i).width; }
    /* N x 12 matrix, columns of source/Storage.cpp:268-290 */
    ImagePtr getLocalizedPositionsAsMatrix() const {
        const size_t n = positionsVector.size();
        ImagePtr m(new Image((long)n, 12));
        for (size_t i = 0; i < n; ++i) {
            const LocalizedPosition_2DGauss &p = positionsVector[i];
            const double row[12] = {(double)p.frameNumber, p.integral, p.width, p.xPosition, p.yPosition,
                                    p.background, p.integralDeviation, p.widthDeviation, p.xPositionDeviation,
                                    p.yPositionDeviation, p.backgroundDeviation, (double)p.nFramesPresent};
            for (int c = 0; c < 12; ++c) (*m)((long)i, c) = row[c];
        }
        return m;
    }
    /* tab-separated text, layout of source/Storage.cpp:292-345 */
    void writePositionsToFile(std::string filePath, std::string header) const {
        if (positionsVector.empty()) throw std::runtime_error("No positions localized!");
        std::ofstream f(filePath.c_str(), std::ios::trunc);
        if (f.fail()) throw std::runtime_error("Unable to create the output file");
        f << "Localized positions using symmetric 2D Gauss fitting\n" << header << "\nDATA FOLLOWS\n";
        f << "First frame\tIntegrated intensity\tFitted PSF standard deviation (pixel)\tX position (pixel)\t"
             "Y position (pixel)\tBackground\tIntensity deviation\tPSF width deviation (pixel)\t"
             "X position deviation (pixel)\tY position deviation (pixel)\tBackground deviation (pixel)\t"
             "Number of frames where this emitter is present\n";
        for (const auto &p : positionsVector) {
            f << p.frameNumber << "\t" << p.integral << "\t" << p.width << "\t" << p.xPosition << "\t" << p.yPosition
              << "\t" << p.background << "\t" << p.integralDeviation << "\t" << p.widthDeviation << "\t"
              << p.xPositionDeviation << "\t" << p.yPositionDeviation << "\t" << p.backgroundDeviation << "\t"
              << p.nFramesPresent << "\n";
            if (f.fail()) throw std::runtime_error("Error writing to the localized positions file");
        }
    }
};

class LocalizedPositionsContainer_2DGaussFixedWidth
    : public LzrTypedContainer<LocalizedPosition_2DGaussFixedWidth, LOCALIZED_POSITIONS_TYPE_2DGAUSS_FIXED_WIDTH, 10> {
public:
    double getIntegralDeviation(size_t i) const { return positionsVector.at(i).integralDeviation; }
    double getXPositionDeviation(size_t i) const { return positionsVector.at(i).xPositionDeviation; }
    double getYPositionDeviation(size_t i) const { return positionsVector.at(i).yPositionDeviation; }
    double getBackgroundDeviation(size_t i) const { return positionsVector.at(i).backgroundDeviation; }
    /* N x 10 matrix, columns of source/Storage.cpp:405-425 */
    ImagePtr getLocalizedPositionsAsMatrix() const {
        const size_t n = positionsVector.size();
        ImagePtr m(new Image((long)n, 10));
        for (size_t i = 0; i < n; ++i) {
            const LocalizedPosition_2DGaussFixedWidth &p = positionsVector[i];
            const double row[10] = {(double)p.frameNumber, p.integral, p.xPosition, p.yPosition, p.background,
                                    p.integralDeviation, p.xPositionDeviation, p.yPositionDeviation,
                                    p.backgroundDeviation, (double)p.nFramesPresent};
            for (int c = 0; c < 10; ++c) (*m)((long)i, c) = row[c];
        }
        return m;
    }
    /* tab-separated text, layout of source/Storage.cpp:427-476 */
    void writePositionsToFile(std::string filePath, std::string header) const {
        if (positionsVector.empty()) throw std::runtime_error("No positions localized!");
        std::ofstream f(filePath.c_str(), std::ios::trunc);
        if (f.fail()) throw std::runtime_error("Unable to create the output file");
        f << "Localized positions using symmetric 2D Gauss fitting with fixed PSF width\n" << header
          << "\nDATA FOLLOWS\n";
        f << "First frame\tIntegrated intensity\tX position (pixel)\tY position (pixel)\tBackground\t"
             "Intensity deviation\tX position deviation (pixel)\tY position deviation (pixel)\t"
             "Background deviation\tNumber of frames where this emitter is present\n";
        for (const auto &p : positionsVector) {
            f << p.frameNumber << "\t" << p.integral << "\t" << p.xPosition << "\t" << p.yPosition << "\t"
              << p.background << "\t" << p.integralDeviation << "\t" << p.xPositionDeviation << "\t"
              << p.yPositionDeviation << "\t" << p.backgroundDeviation << "\t" << p.nFramesPresent << "\n";
            if (f.fail()) throw std::runtime_error("Error writing to the localized positions file");
        }
    }
};

class LocalizedPositionsContainer_Ellipsoidal2DGaussian
    : public LzrTypedContainer<LocalizedPosition_Ellipsoidal2DGauss, LOCALIZED_POSITIONS_TYPE_ELLIPSOIDAL2DGAUSS, 16> {
public:
    double getXWidth(size_t i) const { return positionsVector.at(i).stdDev1; }
    double getYWidth(size_t i) const { return positionsVector.at(i).stdDev2; }
    double getRotation(size_t i) const { return positionsVector.at(i).theta; }
    LocalizedPosition_Ellipsoidal2DGauss getLocalizedPositionAtIndex(int i) { return positionsVector.at(i); }
    /* N x 16 matrix, columns of source/Storage.cpp:614-640 */
    ImagePtr getLocalizedPositionsAsMatrix() const {
        const size_t n = positionsVector.size();
        ImagePtr m(new Image((long)n, 16));
        for (size_t i = 0; i < n; ++i) {
            const LocalizedPosition_Ellipsoidal2DGauss &p = positionsVector[i];
            const double row[16] = {(double)p.frameNumber, p.integral, p.stdDev1, p.stdDev2, p.xPosition,
                                    p.yPosition, p.theta, p.background, p.integralDeviation, p.stdDev1Deviation,
                                    p.stdDev2Deviation, p.xPositionDeviation, p.yPositionDeviation,
                                    p.thetaDeviation, p.backgroundDeviation, (double)p.nFramesPresent};
            for (int c = 0; c < 16; ++c) (*m)((long)i, c) = row[c];
        }
        return m;
    }
    void writePositionsToFile(std::string filePath, std::string header) const {
        if (positionsVector.empty()) throw std::runtime_error("No positions localized!");
        std::ofstream f(filePath.c_str(), std::ios::trunc);
        if (f.fail()) throw std::runtime_error("Unable to create the output file");
        f << "Localized positions using ellipsoidal 2D Gauss fitting\n" << header << "\nDATA FOLLOWS\n";
        f << "First frame\tIntegrated intensity\tFitted PSF standard deviation along first principal axis (pixel)\t"
             "Fitted PSF standard deviation along second principal axis (pixel)\tX position (pixel)\t"
             "Y position (pixel)\tAngular rotation (radians)\tBackground\tIntensity deviation\t"
             "PSF width deviation along first principal axis (pixel)\t"
             "PSF width deviation along second principal axis (pixel)\tX position deviation (pixel)\t"
             "Y position deviation (pixel)\tRotation deviation\tBackground deviation\t"
             "Number of frames where this emitter is present\n";
        for (const auto &p : positionsVector) {
            f << p.frameNumber << "\t" << p.integral << "\t" << p.stdDev1 << "\t" << p.stdDev2 << "\t" << p.xPosition
              << "\t" << p.yPosition << "\t" << p.theta << "\t" << p.background << "\t" << p.integralDeviation << "\t"
              << p.stdDev1Deviation << "\t" << p.stdDev2Deviation << "\t" << p.xPositionDeviation << "\t"
              << p.yPositionDeviation << "\t" << p.thetaDeviation << "\t" << p.backgroundDeviation << "\t"
              << p.nFramesPresent << "\n";
            if (f.fail()) throw std::runtime_error("Error writing to the localized positions file");
        }
    }
};

/* The fitter interface of source/Localization.h:55-61. */
class FitPositions {
public:
    virtual ~FitPositions() {}
    virtual std::shared_ptr<LocalizedPositionsContainer> fit_positions(const ImagePtr image,
                                                                       std::shared_ptr<std::list<Particle> > particles) = 0;
};

#endif
