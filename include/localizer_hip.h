/*
 * localizer_hip.h -- C ABI of the MI355X localization-fitting engine.
 *
 * This is the drop-in boundary for the per-spot 2D Gaussian Levenberg-Marquardt
 * fit of Localizer (source/Localization.cpp).  The reference fits one frame per
 * call through the C++ virtual
 *     FitPositions::fit_positions(ImagePtr, shared_ptr<list<Particle>>)
 *     (source/Localization.h:55-61)
 * and exposes its pipeline to other languages through extern "C" functions
 * returning an int status (source/LocalizerDLL.h:38-58).  This header keeps that
 * style: plain C types, an int return (count >= 0 on success, -1 on error), no
 * exception crosses the boundary, uint16 frames with x as the fast axis
 * (source/FileIO.h:96-116), consecutive frames in one buffer.
 *
 * Per-spot semantics follow the reference fitters:
 *   LZR_MODEL_SYMMETRIC    FitPositions_SymmetricGaussian   (source/Localization.cpp:43-220)
 *   LZR_MODEL_ELLIPSOIDAL  FitPositions_EllipsoidalGaussian (source/Localization.cpp:391-604)
 *   LZR_MODEL_ELLIPSOIDAL_SYMMETRIC_PSF
 *                          FitPositions_EllipsoidalGaussian_SymmetricPSF
 *                                                           (source/Localization.cpp:606-631)
 *   LZR_MODEL_FIXED_WIDTH  FitPositions_FixedWidthGaussian  (source/Localization.cpp:222-388)
 * The C++ classes in localizer_amd/dropin/LocalizationHIP.h wrap this ABI with the
 * reference constructors and container types.
 */
#ifndef LOCALIZER_HIP_H
#define LOCALIZER_HIP_H

#include <stdint.h>

#ifdef __cplusplus
#define LZR_EXTERN_C extern "C"
#else
#define LZR_EXTERN_C
#endif
#define LZR_EXPORT LZR_EXTERN_C __attribute__((visibility("default")))

#define LZR_ABI_VERSION 1

/* Fit models (method codes of source/Defines.h:157-163 are 0 and 5; the ABI uses
 * its own small enum because the symmetric-PSF ellipsoidal variant has no code). */
#define LZR_MODEL_SYMMETRIC 0
#define LZR_MODEL_ELLIPSOIDAL 1
#define LZR_MODEL_ELLIPSOIDAL_SYMMETRIC_PSF 2
#define LZR_MODEL_FIXED_WIDTH 3 /* FitPositions_FixedWidthGaussian (source/Localization.cpp:222-388) */

/* Maximum window half-width: n = 2h+1 <= 31. */
#define LZR_MAX_HALF_WIDTH 15
#define LZR_MAX_PARAMS 7

/* Per-spot outcome.  0 = accepted.  The non-zero codes name the check of the
 * reference driver that erased the particle from the caller's list. */
#define LZR_ACCEPTED 0
#define LZR_REJECT_EDGE 1      /* window leaves the frame   (Localization.cpp:115-124, :467-471) */
#define LZR_REJECT_AMPLITUDE 2 /* fitted amplitude <= 0     (Localization.cpp:160, :510)         */
#define LZR_REJECT_POSITION 3  /* position outside window   (Localization.cpp:166, :516)         */
#define LZR_REJECT_WIDTH 4     /* width band [0.4w, 1.5w]   (Localization.cpp:179, :550)         */
#define LZR_REJECT_EIGEN 5     /* 2x2 eigen-solve failed    (Localization.cpp:533-536)           */
#define LZR_REJECT_ANISOTROPIC 6 /* symmetric-PSF filter: sd ratio (Localization.cpp:617-619)    */
#define LZR_REJECT_PSF_WIDTH 7   /* symmetric-PSF filter: widths   (Localization.cpp:622-625)    */
#define LZR_REJECT_INVALID 8   /* candidate refers to a missing frame / bad window size         */
#define LZR_REJECT_DISTANCE 9  /* fixed width: position moved > 2w from the start (Localization.cpp:349) */

/* GSL-equivalent status that ended the LM loop (gsl_errno.h:50-84). */
#define LZR_LM_SUCCESS 0       /* test_delta satisfied                                         */
#define LZR_LM_CONTINUE (-2)   /* 200-iteration cap reached while still GSL_CONTINUE           */
#define LZR_LM_FAILURE (-1)    /* model evaluation failed (zero width)                         */
#define LZR_LM_ENOPROG 27
#define LZR_LM_ETOLF 29
#define LZR_LM_ETOLX 30
#define LZR_LM_ETOLG 31

/* One candidate: the reference Particle (source/Storage.h:90-101) plus the frame it
 * belongs to and an optional per-spot window half-width and model. */
typedef struct LzrCandidate {
    double x;          /* start x (px, absolute; x is the fast image axis)        */
    double y;          /* start y                                                 */
    double intensity;  /* start amplitude A0                                      */
    double background; /* start offset b0                                         */
    int32_t frame;     /* index of the frame in the stack                         */
    int16_t half_width;/* window half-width h; <= 0 selects ceil(4*initialPSFWidth) */
    int16_t model;     /* LZR_MODEL_*; < 0 selects the context's model            */
} LzrCandidate;

/* One fit result.  params/deviations are in the model's parameter order:
 *   symmetric   (A, r, x0, y0, b)
 *   ellipsoidal (A, sigma_x, sigma_y, x0, y0, rho, b)
 * The record fields reproduce LocalizedPosition_2DGauss (Storage.h:117-138) and
 * LocalizedPosition_Ellipsoidal2DGauss (Storage.h:161-186). */
typedef struct LzrFit {
    double params[LZR_MAX_PARAMS];
    double deviations[LZR_MAX_PARAMS]; /* c*sqrt(C_kk), c = max(1, chi/sqrt(dof)) */
    double integral;
    double integral_deviation;
    double sd1, sd2, theta;            /* ellipsoidal principal widths; symmetric: sd1=sd2=r */
    double sd1_deviation, sd2_deviation;
    double chi;                        /* ||f||_2 at the final point              */
    int32_t status;                    /* LZR_ACCEPTED or LZR_REJECT_*            */
    int32_t lm_status;                 /* LZR_LM_*                                */
    int16_t iterations;                /* outer LM iterations (reference counter) */
    int16_t trials;                    /* residual evaluations at trial points    */
    int16_t accepted_steps;            /* accepted LM steps                       */
    int16_t model;                     /* model actually fitted                   */
} LzrFit;

typedef struct LzrContext LzrContext;

/* Options for lzr_set_option. */
#define LZR_OPT_MAX_ITERATIONS 1   /* outer-iteration cap, default 200 (Localization.cpp:155)        */
#define LZR_OPT_EPSABS 2           /* test_delta absolute tolerance, default 1e-4 (Localization.cpp:154) */
#define LZR_OPT_EPSREL 3           /* test_delta relative tolerance, default 1e-4                      */
#define LZR_OPT_DIAG_UPDATE_ROWS 4 /* rows used by the lmsder scale update: 0 = all, default p (GSL 1.16) */
#define LZR_OPT_WAVES_PER_CU 5     /* persistent-kernel occupancy (autotuned when 0)                   */
#define LZR_OPT_MAX_HALF_WIDTH 6   /* largest per-candidate window half-width accepted (default: the context's) */
#define LZR_OPT_OCCUPANCY 7        /* register budget of the fitting kernel: 1 or 2 waves per SIMD (default 1) */
#define LZR_OPT_ROUNDS 8           /* fitting-kernel launches per model class; unfinished fits are parked between them (default 1) */
#define LZR_OPT_PARK_THRESHOLD 9   /* a wave parks its fits once the queue is drained and at most this many lanes are busy (default 16) */

/* Creates a fitting context on the current HIP device.
 * Replaces the constructors FitPositions_SymmetricGaussian(initialPSFWidth, sigma)
 * (source/Localization.h:68), FitPositions_EllipsoidalGaussian (:97) and
 * FitPositions_EllipsoidalGaussian_SymmetricPSF (:110-113).  default_half_width <= 0
 * keeps the reference window cutoff_radius = ceil(4*initialPSFWidth) (:68).
 * Returns 0, or -1 (no context is created; lzr_last_error(NULL, ...) explains). */
LZR_EXPORT int lzr_create(LzrContext **ctx, int model, double initial_psf_width, double sigma,
                          int default_half_width);
/* (also FitPositions_FixedWidthGaussian(initialPSFWidth, sigma), source/Localization.h:83-95) */

/* Releases a context and its device memory.  Returns 0 or -1.
 * Replaces the fitter's destructor and LocalizerFree (source/LocalizerDLL.cpp:131-133). */
LZR_EXPORT int lzr_destroy(LzrContext *ctx);

/* Copies the last error message (NUL-terminated, truncated to len) and returns its length.
 * ctx may be NULL for errors raised by lzr_create.  The reference surfaces errors as
 * std::runtime_error text (source/Errors.h:36-127); the C++ drop-in rethrows this text. */
LZR_EXPORT int lzr_last_error(const LzrContext *ctx, char *buf, int len);

/* Sets an LZR_OPT_* option.  Returns 0 or -1. (No reference counterpart: the
 * reference hard-codes these constants at source/Localization.cpp:154-155.) */
LZR_EXPORT int lzr_set_option(LzrContext *ctx, int option, double value);

/* Fits every candidate against a host-resident uint16 frame stack (nframes frames of
 * nx*ny pixels, pixel (x,y) of frame f at frames[(f*ny + y)*nx + x]).  Writes one
 * LzrFit per candidate, in candidate order.  Returns the number of accepted fits, or -1.
 * Replaces FitPositions_*::fit_positions (source/Localization.h:55-61) batched over
 * frames, and the numeric core of LocalizerFitEmitters (source/LocalizerDLL.cpp:89-129). */
LZR_EXPORT int lzr_fit_batch(LzrContext *ctx, const uint16_t *frames, int nx, int ny, int nframes,
                             const LzrCandidate *candidates, int ncandidates, LzrFit *results);

/* As lzr_fit_batch, but every pointer is device memory and the work is enqueued on
 * `stream` (a hipStream_t; NULL = the null stream).  Returns 0 (the accepted count is
 * not known until the stream completes; read LzrFit.status), or -1.
 * Same reference interface as lzr_fit_batch; this is the HBM-resident form. */
LZR_EXPORT int lzr_fit_batch_device(LzrContext *ctx, const uint16_t *d_frames, int nx, int ny,
                                    int nframes, const LzrCandidate *d_candidates, int ncandidates,
                                    LzrFit *d_results, void *stream);

/* ABI version (LZR_ABI_VERSION). */
LZR_EXPORT int lzr_abi_version(void);

#endif /* LOCALIZER_HIP_H */
