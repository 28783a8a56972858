/*
 * lzr_oracle.h -- FP64 CPU oracle for the localization-fitting path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (localizer_amd/) links, loads or
 * calls this library.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg use it, as the checker and as the timed CPU baseline.
 *
 * It restates, in plain C and double precision:
 *   - the reference model and Jacobian functions, quirks included
 *       FitFunction_SymmetricGaussian    source/Localization.cpp:1007-1048
 *       Jacobian_SymmetricGaussian       source/Localization.cpp:1128-1181
 *       FitFunction_EllipsoidalGaussian  source/Localization.cpp:1078-1126
 *       Jacobian_EllipsoidalGaussian     source/Localization.cpp:1227-1295
 *   - GSL 1.16 gsl_multifit_fdfsolver_lmsder (called at source/Localization.cpp:75,
 *     :145-155), written from the published MINPACK lmder/lmpar/qrsolv algorithm
 *     (More 1978; More, Garbow, Hillstrom 1980) with the GSL 1.16 choices listed in
 *     lzr_oracle.c; gsl_multifit_test_delta (:154), gsl_multifit_covar (:185) and
 *     gsl_blas_dnrm2 (:186);
 *   - the per-spot drivers FitPositions_SymmetricGaussian::fit_positions
 *     (source/Localization.cpp:43-220), FitPositions_EllipsoidalGaussian (:391-604) and
 *     FitPositions_EllipsoidalGaussian_SymmetricPSF (:606-631);
 *   - the frame-parallel CPU driver of PALMAnalysisController
 *     (source/PALMAnalysis.cpp:79-280): one frame per task on a thread pool.
 *
 * Parity pins (see tests/): model/Jacobian against golden vectors written by the
 * reference's own functions; the LM trajectory against MINPACK lmder (SciPy);
 * covariance against a direct inverse; eigen closed form against numpy/Eigen.
 * The GSL 1.16 deviations from MINPACK (lzr_oracle.c, "GSL 1.16 choices") are
 * "parity unpinned": libgsl is absent, so nothing here can run them.
 */
#ifndef LZR_ORACLE_H
#define LZR_ORACLE_H

#include <stdint.h>
#include "../include/localizer_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OrcOptions {
    int max_iterations;   /* reference loop cap, 200                                  */
    double epsabs;        /* test_delta absolute tolerance, 1e-4                      */
    double epsrel;        /* test_delta relative tolerance, 1e-4                      */
    int diag_update_rows; /* rows of J used by the scale update after a step:
                             p = GSL 1.16 lmutil.c update_diag, 0 = all rows (MINPACK) */
    int inner_cap;        /* trial cap inside one iterate: 10 (GSL), 0 = none (MINPACK) */
} OrcOptions;

typedef struct OrcSettings {
    int model;               /* LZR_MODEL_*                          */
    double initial_psf_width;/* initialPSFWidth                      */
    double sigma;            /* residual weight sigma                */
    int default_half_width;  /* <= 0: ceil(4*initialPSFWidth)        */
    OrcOptions opt;
} OrcSettings;

/* Fills the reference defaults (200, 1e-4, 1e-4, GSL 1.16 scale update, cap 10). */
void orc_default_options(OrcOptions *opt);

/* Model residuals f_k, k = i + n*j, for pixel (xoff+i, yoff+j).  Returns 0 or -1
 * (GSL_FAILURE for zero width, as the reference). model: LZR_MODEL_SYMMETRIC or
 * LZR_MODEL_ELLIPSOIDAL. */
int orc_model(int model, const double *theta, const double *data, int n, double xoff,
              double yoff, double sigma, double *f);
/* Jacobian rows (N x p, row-major), reference formulas including its quirks. */
int orc_jacobian(int model, const double *theta, int n, double xoff, double yoff,
                 double sigma, double *J);

/* Fixed-width model: theta = (A, x0, y0, b), r = width; J is N x 4. */
int orc_model_fixed_width(const double *theta, double width, const double *data, int n, double xoff,
                          double yoff, double sigma, double *f);
int orc_jacobian_fixed_width(const double *theta, double width, int n, double xoff, double yoff,
                             double sigma, double *J);

/* BLAS dnrm2 (scaled sum of squares). */
double orc_nrm2(int n, const double *x, int stride);

/* gsl_multifit_test_delta: 0 (GSL_SUCCESS) iff |dx_i| < epsabs + epsrel*|x_i| for all i,
 * else -2 (GSL_CONTINUE). */
int orc_test_delta(int p, const double *dx, const double *x, double epsabs, double epsrel);

/* gsl_multifit_covar(J, 0.0, C): C = (J^T J)^-1 via QR with column pivoting; columns
 * beyond the first exactly-zero pivot get zero covariance. C is p x p row-major. */
void orc_covar(int N, int p, const double *J, double *C);

/* Runs the LM loop of the reference driver on one window (model, data n x n at
 * (xoff, yoff)) from theta0.  Writes the final theta, and if trace != NULL the accepted
 * point after every trial evaluation (trace[t*p + i], t < max_trace), and if trial_trace !=
 * NULL the trial point of every evaluation (same layout) and its residual norm
 * (trial_trace[max_trace*p + t]).  Returns the number of trials; *iterations, *lm_status,
 * *accepted receive the loop counters.  fixed_width is the width of LZR_MODEL_FIXED_WIDTH
 * (ignored by the other models). */
int orc_lm_run(int model, const double *data, int n, double xoff, double yoff, double sigma,
               const double *theta0, const OrcOptions *opt, double *theta, double *f_out,
               double *J_out, int *iterations, int *lm_status, int *accepted, double *trace,
               int max_trace, double fixed_width, double *trial_trace);

/* Eigen-decomposition of [[sx^2, rho sx sy], [rho sx sy, sy^2]] as used at
 * source/Localization.cpp:525-547: ascending eigenvalues, sd = sqrt(lambda),
 * theta = atan(v0/v1) of the first eigenvector (0 when v1 == 0).  Returns 0, or -1
 * when the input is not finite (the eigen-solver failure branch). */
int orc_eigen2(double sx, double sy, double rho, double *sd1, double *sd2, double *theta);

/* Fits one candidate; frame is the candidate's nx x ny uint16 frame. */
void orc_fit_roi(const OrcSettings *s, const uint16_t *frame, int nx, int ny,
                 const LzrCandidate *cand, LzrFit *out);

/* Fits all candidates with a frame-parallel thread pool (nthreads <= 0: hardware
 * concurrency).  Candidates of one frame are fitted by one task, in list order, as
 * PALMAnalysisController's workers do.  Returns the number of accepted fits. */
/* ---- candidate generation (lzr_find_oracle.c) ---- */
/* GLRT window side for a PSF width: ceil(4 w), made odd. */
int orc_glrt_window(double width);
/* mask[y*nx + x] = 1 where the GLRT segments the pixel; test_out (may be NULL) receives
 * the test statistic of every tested pixel (0 elsewhere).  Returns 0 or -1. */
int orc_glrt_mask(const uint16_t *frame, int nx, int ny, double width, double pfa, uint8_t *mask,
                  double *test_out);
/* 8-connected components of the mask in the reference's order: out[4k..4k+3] =
 * (x, y, intensity, background) of particle k (at most max_out written).  Returns the
 * number of particles or -1. */
int orc_find_particles(const uint16_t *frame, const uint8_t *mask, int nx, int ny, double *out, int max_out);

/* ---- localization bitmap (lzr_render_oracle.c) ---- */
int orc_erf_table_size(void);
void orc_erf_table(double *table);
double orc_erf_lookup(const double *table, double x);
int orc_render_size(int x_size, int y_size, double scale, int *width, int *height);
int orc_render(const double *pos, int n, int x_size, int y_size, double scale, int method, const double *p,
               int weighing, double *image);

int orc_fit_batch(const OrcSettings *s, const uint16_t *frames, int nx, int ny, int nframes,
                  const LzrCandidate *cands, int ncands, LzrFit *out, int nthreads);
/* The same on FP64 frames (the reference's Image; any pixel values). */
int orc_fit_batch_f64(const OrcSettings *s, const double *frames, int nx, int ny, int nframes,
                      const LzrCandidate *cands, int ncands, LzrFit *out, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
