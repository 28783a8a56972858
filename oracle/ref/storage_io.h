// Shared I/O of the positions-container harnesses (the reference's containers in
// oracle/ref/ref_storage.cpp, the drop-in's in tests/cpp/storage_main.cpp): the same input
// file fills either type set, since both use the reference's field names.  TEST INFRASTRUCTURE.
#ifndef LZR_STORAGE_IO_H
#define LZR_STORAGE_IO_H
#include <cstdint>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace lzr_storage_io {

template <typename T>
T get(std::ifstream &f) {
    T v;
    f.read(reinterpret_cast<char *>(&v), sizeof v);
    if (!f) throw std::runtime_error("short input");
    return v;
}

inline std::shared_ptr<LocalizedPositionsContainer> read(const char *path) {
    std::ifstream f(path, std::ios::binary);
    const int32_t type = get<int32_t>(f), n = get<int32_t>(f);
    std::shared_ptr<LocalizedPositionsContainer> c;
    if (type == 0) c.reset(new LocalizedPositionsContainer_2DGauss());
    else if (type == 1) c.reset(new LocalizedPositionsContainer_2DGaussFixedWidth());
    else if (type == 5) c.reset(new LocalizedPositionsContainer_Ellipsoidal2DGaussian());
    else throw std::runtime_error("unknown type");
    for (int k = 0; k < n; ++k) {
        const uint64_t frame = get<uint64_t>(f), nf = get<uint64_t>(f);
        if (type == 0) {
            std::shared_ptr<LocalizedPosition_2DGauss> p(new LocalizedPosition_2DGauss());
            p->frameNumber = frame;
            p->nFramesPresent = nf;
            double *d[] = {&p->integral, &p->width, &p->xPosition, &p->yPosition, &p->background,
                           &p->integralDeviation, &p->widthDeviation, &p->xPositionDeviation,
                           &p->yPositionDeviation, &p->backgroundDeviation};
            for (double *q : d) *q = get<double>(f);
            c->addPosition(p);
        } else if (type == 1) {
            std::shared_ptr<LocalizedPosition_2DGaussFixedWidth> p(new LocalizedPosition_2DGaussFixedWidth());
            p->frameNumber = frame;
            p->nFramesPresent = nf;
            double *d[] = {&p->integral, &p->xPosition, &p->yPosition, &p->background, &p->integralDeviation,
                           &p->xPositionDeviation, &p->yPositionDeviation, &p->backgroundDeviation};
            for (double *q : d) *q = get<double>(f);
            c->addPosition(p);
        } else {
            std::shared_ptr<LocalizedPosition_Ellipsoidal2DGauss> p(new LocalizedPosition_Ellipsoidal2DGauss());
            p->frameNumber = frame;
            p->nFramesPresent = nf;
            double *d[] = {&p->integral, &p->stdDev1, &p->stdDev2, &p->xPosition, &p->yPosition, &p->theta,
                           &p->background, &p->integralDeviation, &p->stdDev1Deviation, &p->stdDev2Deviation,
                           &p->xPositionDeviation, &p->yPositionDeviation, &p->thetaDeviation,
                           &p->backgroundDeviation};
            for (double *q : d) *q = get<double>(f);
            c->addPosition(p);
        }
    }
    return c;
}

inline int run(const char *in, const char *mat, const char *txt, const char *header) {
    std::shared_ptr<LocalizedPositionsContainer> c = read(in);
    ImagePtr m = c->getLocalizedPositionsAsMatrix();
    std::ofstream o(mat, std::ios::binary);
    const int64_t rows = m->rows(), cols = m->cols();
    o.write(reinterpret_cast<const char *>(&rows), sizeof rows);
    o.write(reinterpret_cast<const char *>(&cols), sizeof cols);
    for (int64_t j = 0; j < cols; ++j)
        for (int64_t i = 0; i < rows; ++i) {
            const double v = (*m)(i, j);
            o.write(reinterpret_cast<const char *>(&v), sizeof v);
        }
    c->writePositionsToFile(txt, header);
    return 0;
}

}  // namespace lzr_storage_io
#endif
