// Golden-vector harness: calls the reference's own model and Jacobian functions
// (FitFunction_*/Jacobian_* in source/Localization.cpp, compiled from
// /root/reference by oracle/ref/Makefile) and Eigen's SelfAdjointEigenSolver (vendored
// Eigen 3.3.7, used by the reference at source/Localization.cpp:525-547).
// TEST INFRASTRUCTURE ONLY.
//
// usage: ref_models <cases.bin> <out.bin>
//   cases.bin: int32 ncases, then per case
//     int32 kind (0 symmetric model+J, 1 ellipsoidal model+J, 2 eigen2), int32 n,
//     double xoff, yoff, sigma, theta[7], data[n*n]   (kind 2: theta = sx, sy, rho)
//   out.bin per case: int32 status_f, int32 status_J, double f[N], double J[N*p]
//     (kind 2: int32 info, int32 0, double sd1, sd2, theta)
#include "Localization.h"
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
    int32_t ncases; rd(in, &ncases, 4);
    for (int c = 0; c < ncases; ++c) {
        int32_t kind, n; double xoff, yoff, sigma, theta[7];
        rd(in, &kind, 4); rd(in, &n, 4); rd(in, &xoff, 8); rd(in, &yoff, 8); rd(in, &sigma, 8);
        rd(in, theta, 56);
        const int N = n * n;
        std::vector<double> data(N);
        rd(in, data.data(), 8 * N);
        if (kind == 2) {
            Eigen::Matrix2d m;
            const double sx = theta[0], sy = theta[1], rho = theta[2];
            m << sx * sx, rho * sx * sy, rho * sx * sy, sy * sy;
            Eigen::SelfAdjointEigenSolver<Eigen::Matrix2d> es(m);
            int32_t info = es.info() == Eigen::Success ? 0 : 1, z = 0;
            double res[3] = {0, 0, 0};
            if (info == 0) {
                Eigen::Vector2d ev = es.eigenvalues();
                Eigen::Matrix2d V = es.eigenvectors();
                res[0] = std::sqrt(ev[0]);
                res[1] = std::sqrt(ev[1]);
                res[2] = V(1, 0) == 0 ? 0.0 : std::atan(V(0, 0) / V(1, 0));
            }
            fwrite(&info, 4, 1, out); fwrite(&z, 4, 1, out); fwrite(res, 8, 3, out);
            continue;
        }
        const int p = kind == 0 ? 5 : 7;
        measured_data_Gauss_fits d;
        d.xOffset = xoff; d.yOffset = yoff; d.sigma = sigma;
        d.imageSubset = ImagePtr(new Image(n, n));
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) (*d.imageSubset)(i, j) = data[i + n * j];
        std::vector<double> f(N), J(N * p);
        gsl_vector pv; pv.size = p; pv.stride = 1; pv.data = theta; pv.block = 0; pv.owner = 0;
        gsl_vector fv; fv.size = N; fv.stride = 1; fv.data = f.data(); fv.block = 0; fv.owner = 0;
        gsl_matrix Jm; Jm.size1 = N; Jm.size2 = p; Jm.tda = p; Jm.data = J.data(); Jm.block = 0; Jm.owner = 0;
        int32_t sf, sj;
        if (kind == 0) {
            sf = FitFunction_SymmetricGaussian(&pv, &d, &fv);
            sj = Jacobian_SymmetricGaussian(&pv, &d, &Jm);
        } else {
            sf = FitFunction_EllipsoidalGaussian(&pv, &d, &fv);
            sj = Jacobian_EllipsoidalGaussian(&pv, &d, &Jm);
        }
        fwrite(&sf, 4, 1, out); fwrite(&sj, 4, 1, out);
        fwrite(f.data(), 8, N, out); fwrite(J.data(), 8, N * p, out);
    }
    fclose(out);
    return 0;
}
