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
#define LZR_OPT_OCCUPANCY 7        /* register budget of the fitting kernel: 1 or 2 waves per SIMD; 0 (default) = 2 when the window tiles of 6 waves fit a CU's LDS */
#define LZR_OPT_ROUNDS 8           /* fitting-kernel launches per model class; unfinished fits are parked between them (default 1, which runs a build of the fitting kernel without the parking paths: no register spills, ~20 % faster; records identical) */
#define LZR_OPT_PARK_THRESHOLD 9   /* a wave parks its fits once the queue is drained and at most this many lanes are busy (default 16) */
#define LZR_OPT_STRAGGLER_TRIALS 10 /* a fit still running after this many trials continues in the straggler kernel, on 1..16 lanes by the count of such fits (default 48, 0 = off; until set, a call of at most 1024 candidates hands every fit over after its first trial, which shortens small calls and leaves the records unchanged) */
#define LZR_OPT_DEFAULT_HALF_WIDTH 11 /* window half-width of candidates with half_width 0 (read-only: set at lzr_create) */
#define LZR_OPT_STRAGGLER_GROUP 12  /* lanes per fit in the straggler kernel (64/lanes ROIs per wavefront): 1, 2, 4, 8, 16, or 0 (default) = by the number of handed-off fits */
#define LZR_OPT_NET_MAX_INTENSITY 13 /* lzr_fit_batch_device_netinit starts from the network only the candidates whose intensity is below this value, the others from the reference start (default +inf: every candidate) */

/* Creates a fitting context on the current HIP device.
 * Replaces the constructors FitPositions_SymmetricGaussian(initialPSFWidth, sigma)
 * (source/Localization.h:68), FitPositions_EllipsoidalGaussian (:97) and
 * FitPositions_EllipsoidalGaussian_SymmetricPSF (:110-113).  default_half_width <= 0
 * keeps the reference window cutoff_radius = ceil(4*initialPSFWidth) (:68).
 * Returns 0, or -1 (no context is created; lzr_last_error(NULL, ...) explains). */
LZR_EXPORT int lzr_create(LzrContext **ctx, int model, double initial_psf_width, double sigma,
                          int default_half_width);
/* (also FitPositions_FixedWidthGaussian(initialPSFWidth, sigma), source/Localization.h:83-95) */

/* A second context with this one's model, widths, sigma and options (its own work memory
 * and streams): fits on the two are identical, and the two may fit at the same time.
 * (No reference counterpart: a FitPositions instance is stateless between calls.) */
LZR_EXPORT int lzr_clone(LzrContext *ctx, LzrContext **out);

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

/* lzr_fit_batch / lzr_fit_batch_device on float frames: pixel values that are not camera
 * counts (the reference fits any FP64 Image, source/Localization.cpp:126-130; FP64 values
 * are stored as FP32 here, every sum is FP64).  Integer frames give the same records as the
 * uint16 entry points.  The pixel tile is twice as large in LDS: windows up to 23x23
 * (half-width 11).  Same return values. */
LZR_EXPORT int lzr_fit_batch_f32(LzrContext *ctx, const float *frames, int nx, int ny, int nframes,
                                 const LzrCandidate *candidates, int ncandidates, LzrFit *results);
LZR_EXPORT int lzr_fit_batch_device_f32(LzrContext *ctx, const float *d_frames, int nx, int ny, int nframes,
                                        const LzrCandidate *d_candidates, int ncandidates, LzrFit *d_results,
                                        void *stream);

/* As lzr_fit_batch_device, but allocating nothing: no device allocation, no allocating
 * library call.  No counterpart in the reference; it is for callers that cannot allocate
 * (code under stream capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * work memory is a fixed 64 MiB device array per GPU, held until the process ends: at most
 * LZR_NOALLOC_MAX_CANDIDATES candidates per call (larger batches return -1); fits run in
 * one round (LZR_OPT_ROUNDS is ignored); long fits that find the array's hand-off entries
 * taken finish in the first kernel, with the same results.  Every call on one GPU shares
 * that array: calls that overlap on one GPU (any context, any stream) race -- order them.
 * Results are bit-identical to lzr_fit_batch_device's. */
#define LZR_NOALLOC_MAX_CANDIDATES (2 << 20)
LZR_EXPORT int lzr_fit_batch_device_noalloc(LzrContext *ctx, const uint16_t *d_frames, int nx, int ny,
                                            int nframes, const LzrCandidate *d_candidates, int ncandidates,
                                            LzrFit *d_results, void *stream);

/* ABI version (LZR_ABI_VERSION). */
LZR_EXPORT int lzr_abi_version(void);

/* Reads an LZR_OPT_* option (e.g. the configuration lzr_autotune_device chose).
 * Returns 0 or -1.  (No reference counterpart.) */
LZR_EXPORT int lzr_get_option(LzrContext *ctx, int option, double *value);

/* ---- frame-block graphs ----------------------------------------------------------------
 * No reference counterpart (the reference fits frame by frame on host threads,
 * source/PALMAnalysis.cpp:196-280).  A graph is one lzr_fit_batch_device call captured
 * once on fixed device buffers (frames, candidates, results) and replayed per frame
 * block: the caller refills the buffers and *d_count (device int, the block's candidate
 * count, at most max_candidates), then launches the graph.  Options set on the context
 * after capture do not affect the graph.  A graph owns its work buffers (sized at creation):
 * eager calls on the context, of any size, never touch them.  Replays of one graph share
 * them: enqueue a graph's replays on one stream (or order them) -- overlapping replays race. */
typedef struct LzrGraph LzrGraph;
LZR_EXPORT int lzr_graph_create(LzrContext *ctx, LzrGraph **graph, const uint16_t *d_frames, int nx, int ny,
                                int nframes, const LzrCandidate *d_candidates, const int *d_count,
                                int max_candidates, LzrFit *d_results);
LZR_EXPORT int lzr_graph_launch(LzrGraph *graph, void *stream);
LZR_EXPORT int lzr_graph_destroy(LzrGraph *graph);

/* Autotuner: times launch configurations on this batch -- each fits the whole batch once
 * into d_results -- and keeps the fastest in the context: first the register budget
 * (LZR_OPT_OCCUPANCY), persistent grid (LZR_OPT_WAVES_PER_CU) and continuation rounds
 * (LZR_OPT_ROUNDS), then with the best of those the hand-off threshold
 * (LZR_OPT_STRAGGLER_TRIALS 32/48/96) and the straggler kernel's lanes per fit, i.e. its
 * ROIs per wavefront (LZR_OPT_STRAGGLER_GROUP auto/1/2/4/8/16).  Every configuration returns
 * the same records.  *best_ms (may be NULL) receives the best time.  Synchronous.  Returns 0
 * or -1.  (No reference counterpart.) */
LZR_EXPORT int lzr_autotune_device(LzrContext *ctx, const uint16_t *d_frames, int nx, int ny, int nframes,
                                   const LzrCandidate *d_candidates, int ncandidates, LzrFit *d_results,
                                   void *stream, double *best_ms);

/* ---- candidate generation ---------------------------------------------------------------
 * The reference's segmentation and particle finding on the device:
 *   ThresholdImage_GLRT_FFT(PFA, width)  (source/Segmentation.cpp:241-430)
 *   ParticleFinder_adjacent8             (source/ParticleFinding.cpp:283-427)
 * A finder holds the GLRT parameters (window side ceil(4 w), made odd, at most 15) and
 * work buffers; max_segmented_per_frame (<= 16384; 0: 16384) bounds the segmented pixels
 * of one frame. */
typedef struct LzrFinder LzrFinder;
LZR_EXPORT int lzr_finder_create(LzrFinder **finder, double psf_width, double pfa, int max_segmented_per_frame);
LZR_EXPORT int lzr_finder_destroy(LzrFinder *finder);
LZR_EXPORT int lzr_finder_last_error(const LzrFinder *finder, char *buf, int len);

/* Segments nframes device frames and writes their particles as candidates (frame-major,
 * each frame in the reference's particle order; x, y = mean pixel coordinates of the
 * 8-connected component, intensity = max - background, background = mean of the
 * unsegmented pixels; half_width 0, model -1) to d_candidates; *d_count (device int)
 * receives their number, or -1 when a frame had more segmented pixels than the finder
 * holds or the particles exceed max_candidates (then nothing is written).  Enqueued on
 * `stream`; the count can feed a frame-block graph directly.  Returns 0 or -1. */
LZR_EXPORT int lzr_find_candidates_device(LzrFinder *finder, const uint16_t *d_frames, int nx, int ny, int nframes,
                                          LzrCandidate *d_candidates, int max_candidates, int *d_count,
                                          void *stream);

/* ---- whole pipeline ---------------------------------------------------------------------
 * Segmentation + particle finding + fit of a frame stack in host memory, results as the
 * reference's positions matrix (column-major, n_emitters rows, sorted by frame; within a
 * frame in the particle finder's order), allocated by the callee and released with
 * lzr_free.  Return 0 or -1 (lzr_pipeline_last_error explains). */

/* LocalizerFitEmitters (source/LocalizerDLL.h:51-52, source/LocalizerDLL.cpp:89-129):
 * ellipsoidal fit, 16 columns [frame, integral, sd1, sd2, x, y, theta, background,
 * integralDev, sd1Dev, sd2Dev, xDev, yDev, thetaDev, backgroundDev, nFramesPresent]
 * (source/Storage.cpp:614-640).  n_rows = x extent, n_cols = y extent. */
LZR_EXPORT int lzr_fit_emitters(const uint16_t *image_data, int n_rows, int n_cols, int n_images, double psf_width,
                                double glrt_insensitivity, int *n_emitters, double **emitter_coordinates);
/* The same with any LZR_MODEL_*: 12 columns (symmetric, source/Storage.cpp:268-290),
 * 10 (fixed width, :405-425) or 16 (ellipsoidal, symmetric-PSF filter). */
LZR_EXPORT int lzr_localize(const uint16_t *image_data, int n_rows, int n_cols, int n_images, double psf_width,
                            double glrt_insensitivity, int model, int *n_emitters, int *n_columns, double **matrix);
/* LocalizerFree (source/LocalizerDLL.cpp:131-133). */
LZR_EXPORT void lzr_free(double *ptr);
LZR_EXPORT int lzr_pipeline_last_error(char *buf, int len);

/* ---- LM start points --------------------------------------------------------------
 * No reference counterpart: the reference always starts the LM at the particle
 * (A0 = intensity, w, x, y, b0; source/Localization.cpp:138-142).  A start point
 * replaces those values for one candidate; the window still follows the candidate's
 * x, y, so the data fitted and every post-check are the reference's.
 *   width >  0  start at (amplitude, width, x, y, background), absolute pixels
 *               (ellipsoidal: sigma_x = sigma_y = width, rho = 0; fixed width: width unused)
 *   width == 0  the reference start
 *   width <  0  skip the candidate: its result record is left untouched */
typedef struct LzrStart {
    double amplitude;
    double width;
    double x;
    double y;
    double background;
} LzrStart;

/* lzr_fit_batch_device with one start point per candidate (d_starts may be NULL). */
LZR_EXPORT int lzr_fit_batch_device_start(LzrContext *ctx, const uint16_t *d_frames, int nx, int ny,
                                          int nframes, const LzrCandidate *d_candidates,
                                          const LzrStart *d_starts, int ncandidates, LzrFit *d_results,
                                          void *stream);

/* ---- learned initializer --------------------------------------------------------------
 * No reference counterpart.  A small CNN predicts the LM start of the symmetric model
 * from the 11 x 11 patch around each candidate (the window of half-width 5 placed by
 * the reference rule x_offset = floor(x - 5)), normalized as (patch - median) / scale,
 * scale = max(patch) - median (1 if smaller).  Network: conv 3x3 1->32 + ReLU (11->9),
 * conv 3x3 32->32 + ReLU (9->7), flatten 1568 (index (y*7 + x)*32 + channel), FC 64 +
 * ReLU, FC 5.  BF16 weights and activations, FP32 accumulation (MFMA), FP32 master
 * weights and Adam state.  Outputs t = (A/scale, log(r/w), x0 - xc, y0 - yc,
 * (b - median)/scale), (xc, yc) = window centre.
 * Parameters (FP32, this order): conv1 W[32][3][3] b[32]; conv2 W[32 out][32 in][3][3]
 * b[32]; fc1 W[64][1568] b[64]; fc2 W[5][64] b[5]. */
typedef struct LzrNet LzrNet;
#define LZR_NET_HALF_WIDTH 5
#define LZR_NET_OUTPUTS 5
#define LZR_NET_OFF_CONV1_W 0
#define LZR_NET_OFF_CONV1_B 288
#define LZR_NET_OFF_CONV2_W 320
#define LZR_NET_OFF_CONV2_B 9536
#define LZR_NET_OFF_FC1_W 9568
#define LZR_NET_OFF_FC1_B 109920
#define LZR_NET_OFF_FC2_W 109984
#define LZR_NET_OFF_FC2_B 110304
#define LZR_NET_NPARAMS 110309

/* Creates a network on the current device, He-initialized from `seed`.  Returns 0 or -1. */
LZR_EXPORT int lzr_net_create(LzrNet **net, double initial_psf_width, unsigned long long seed);
LZR_EXPORT int lzr_net_destroy(LzrNet *net);
LZR_EXPORT int lzr_net_last_error(const LzrNet *net, char *buf, int len);
/* Host FP32 parameters in the order above; set_params also resets the Adam state. */
LZR_EXPORT int lzr_net_set_params(LzrNet *net, const float *params, int count);
LZR_EXPORT int lzr_net_get_params(LzrNet *net, float *params, int count);

/* Raw network outputs t (ncandidates x 5 FP32, device) for the candidates' patches;
 * rows of candidates whose patch leaves the frame are NaN. */
LZR_EXPORT int lzr_net_forward_device(LzrNet *net, const uint16_t *d_frames, int nx, int ny, int nframes,
                                      const LzrCandidate *d_candidates, int ncandidates, float *d_outputs,
                                      void *stream);

/* Start points from the network (width = 0, the reference start, where the patch
 * leaves the frame or the prediction is not finite). */
LZR_EXPORT int lzr_net_predict_device(LzrNet *net, const uint16_t *d_frames, int nx, int ny, int nframes,
                                      const LzrCandidate *d_candidates, int ncandidates, LzrStart *d_starts,
                                      void *stream);

/* One training step on a batch: labels are LM fits of the same candidates (accepted
 * symmetric fits are used, the others are ignored); loss = mean over used candidates of
 * sum_k (t_k - target_k)^2.  apply = 1: Adam step with `learning_rate`; apply = 0: only
 * gradients.  *loss (host, may be NULL) receives the batch loss and grads (host, may be
 * NULL, LZR_NET_NPARAMS floats) the gradient; either makes the call synchronous. */
LZR_EXPORT int lzr_net_train_step_device(LzrNet *net, const uint16_t *d_frames, int nx, int ny, int nframes,
                                         const LzrCandidate *d_candidates, const LzrFit *d_labels,
                                         int ncandidates, float learning_rate, int apply, float *loss,
                                         float *grads, void *stream);

/* Net-initialized LM with fallback: starts from lzr_net_predict_device, then refits
 * from the reference start (source/Localization.cpp:138-142, :485-491) every candidate
 * whose net-start fit
 *   - was rejected or did not converge (LZR_LM_SUCCESS), when refit_rejected != 0, or
 *   - was accepted with a residual |f| above residual_factor times the residual at the
 *     reference start point (the normalized residuals chi/sqrt(dof) share the dof).
 * A refitted candidate's record is the plain fit's record bit for bit.  The network starts
 * only the candidates whose intensity is below LZR_OPT_NET_MAX_INTENSITY (default: all); the
 * others start from the reference start and get the plain record.  d_work:
 * ncandidates LzrStart of scratch; afterwards its width is 0 for refitted candidates and
 * -1 for the others, and its amplitude holds |f| at the reference start for the accepted
 * net-start fits whose test evaluated it (NaN for the others: a fit whose window is not the
 * network's 11 x 11 patch and whose model lies within |f| (1 - 1/residual_factor) of the
 * start point's model passes without reading its pixels).  d_nrefit (may be null):
 * one device unsigned, set to the number of refitted candidates.  No reference counterpart (the reference has no network start);
 * the records it returns are those of lzr_fit_batch_device for the refitted candidates.
 * Returns 0 or -1. */
LZR_EXPORT int lzr_fit_batch_device_netinit(LzrContext *ctx, LzrNet *net, const uint16_t *d_frames, int nx,
                                            int ny, int nframes, const LzrCandidate *d_candidates,
                                            int ncandidates, LzrFit *d_results, LzrStart *d_work,
                                            double residual_factor, int refit_rejected, unsigned *d_nrefit,
                                            void *stream);

/* ---- Localization bitmap ------------------------------------------------------
 * PALMBitmapImageCalculator::CalculateImage(positions, xSize, ySize, outputScaleFactor)
 * (source/PALMImages.cpp:68-157): every localization adds a Gaussian, integrated over
 * each output pixel with the reference's erf lookup table (source/PALMImages.cpp:34-66),
 * to a (int)(xSize*scale) x (int)(ySize*scale) FP64 image (x fast, the Eigen column-major
 * Image of the reference), pixels accumulated in position order (the reference's sums,
 * bit for bit).  The positions are the container getters the calculator reads
 * (getXPosition, getYPosition, getIntegral, getXPositionDeviation, getBackground). */
typedef struct LzrPosition {
    double x, y;         /* px                                            */
    double integral;     /* fitted integral                               */
    double x_deviation;  /* fit uncertainty of x                          */
    double background;   /* fitted offset                                 */
} LzrPosition;

/* Deviation of each Gaussian (PALMBITMAP_DEVIATION_*, source/Defines.h:113-115). */
#define LZR_DEVIATION_CONSTANT 0        /* _Constant(p[0] = deviation), PALMImages.h:122-133          */
#define LZR_DEVIATION_FIT_UNCERTAINTY 1 /* _FitUncertainty(p[0] = scale, p[1] = upperLimit), :106-120  */
#define LZR_DEVIATION_GAUSSIAN_MASK 2   /* _GaussianMask(p[0] = PSFWidth, p[1] = cameraOffset,
                                           p[2] = cameraMultiplicationFactor), :138-157, PALMImages.cpp:159-177 */
/* Amplitude of each Gaussian (PALMBITMAP_EMITTERWEIGHING_*, source/Defines.h:110-111). */
#define LZR_WEIGHING_SAME 0
#define LZR_WEIGHING_INTEGRAL 1 /* the fitted integral; positions with integral 0 are skipped */

typedef struct LzrRenderer LzrRenderer;

/* Renderer with the erf table on the current device (PALMBitmapImageCalculator
 * constructor, source/PALMImages.h:70-74).  Returns 0 or -1. */
LZR_EXPORT int lzr_renderer_create(LzrRenderer **renderer);
LZR_EXPORT int lzr_renderer_destroy(LzrRenderer *renderer);
/* Error text of the last failing call (renderer NULL: of lzr_renderer_create). */
LZR_EXPORT int lzr_renderer_last_error(const LzrRenderer *renderer, char *buf, int len);

/* Output size (int)(x_size*output_scale) x (int)(y_size*output_scale); -1 when
 * output_scale <= 0 (the reference throws) or the image would be smaller than 2x2. */
LZR_EXPORT int lzr_render_image_size(int x_size, int y_size, double output_scale, int *width, int *height);

/* CalculateImage on the device: d_image receives width*height doubles (x fast), every
 * pixel written.  deviation_params holds 3 doubles (unused ones ignored).  Returns 0, or
 * -1 with the reference's error text (unknown weighing, positions without integral /
 * background for the Gaussian mask, output_scale <= 0).  Synchronizes the stream once
 * (the number of (tile, position) pairs sizes the sort). */
LZR_EXPORT int lzr_render_device(LzrRenderer *renderer, const LzrPosition *d_positions, int n_positions,
                                 int x_size, int y_size, double output_scale, int deviation_method,
                                 const double *deviation_params, int weighing, double *d_image, void *stream);

/* The same on host buffers: positions in host memory, image (width*height doubles)
 * written in host memory; blocks until done. */
LZR_EXPORT int lzr_render(LzrRenderer *renderer, const LzrPosition *positions, int n_positions, int x_size,
                          int y_size, double output_scale, int deviation_method, const double *deviation_params,
                          int weighing, double *image);

/* The accepted fits of a result array as positions, in order (the container the
 * reference fitters return holds exactly these, source/Localization.cpp:190-212).
 * *d_count (device) receives the number written. */
LZR_EXPORT int lzr_positions_from_fits_device(LzrRenderer *renderer, const LzrFit *d_fits, int n_fits,
                                              LzrPosition *d_positions, int *d_count, void *stream);

/* ---- host batcher: the reference's per-frame call pattern ------------------------ */

/* The reference's analysis controller shares one fitter between hardware_concurrency()
 * worker threads, each calling fit_positions(image, particles) once per frame
 * (source/PALMAnalysis.cpp:96-130, :204-252, call at :234; source/Localization.h:61).
 * A batcher gathers such concurrent per-frame calls into one device launch: each caller
 * deposits its candidates' fitting windows (only those pixels travel, through pinned
 * memory) and blocks until its batch is fitted.  Every record equals the one
 * lzr_fit_batch returns for that frame alone (nframes = 1), bit for bit.
 * nx, ny: the frame size of every call; max_frames / max_rois: one batch's capacity
 * (frames, candidates; window pixels for max_rois candidates at the default window);
 * max_wait_us: how long an open batch waits for more callers after its first one (it
 * is launched earlier when full or when every caller inside the batcher has joined it);
 * inflight (1..64): batches fitted at the same time, each on its own clone of ctx
 * (lzr_clone at creation: options set on ctx later do not reach the batcher). */
typedef struct LzrBatcher LzrBatcher;
LZR_EXPORT int lzr_batcher_create(LzrBatcher **batcher, LzrContext *ctx, int nx, int ny, int max_frames,
                                  int max_rois, int max_wait_us, int inflight);

/* One frame (nx*ny, x fast) and its candidates (frame field 0; any other value is an
 * invalid request, as in a one-frame lzr_fit_batch call).  Safe to call from any number
 * of threads.  Returns the number of accepted fits, or -1 (frame size differs from the
 * batcher's, the frame exceeds one batch's capacity, or a device error). */
LZR_EXPORT int lzr_batcher_fit_frame(LzrBatcher *batcher, const uint16_t *frame, int nx, int ny,
                                     const LzrCandidate *candidates, int ncandidates, LzrFit *results);

/* The same for an FP64 frame (the reference's Image, column-major with rows = x, i.e. x
 * fast): the pixels inside fitting windows must be integer counts in [0, 65535] (-1
 * otherwise, for this call only). */
LZR_EXPORT int lzr_batcher_fit_frame_f64(LzrBatcher *batcher, const double *frame, int nx, int ny,
                                         const LzrCandidate *candidates, int ncandidates, LzrFit *results);

/* Batches launched and frames fitted so far. */
LZR_EXPORT int lzr_batcher_stats(LzrBatcher *batcher, unsigned long long *batches, unsigned long long *frames);
/* Error text of the last failing call (batcher NULL: of lzr_batcher_create). */
LZR_EXPORT int lzr_batcher_last_error(const LzrBatcher *batcher, char *buf, int len);
/* Stops the worker and frees the buffers; no call may be in progress. */
LZR_EXPORT int lzr_batcher_destroy(LzrBatcher *batcher);

#endif /* LOCALIZER_HIP_H */
