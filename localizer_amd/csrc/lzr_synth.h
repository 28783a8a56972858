/*
 * lzr_synth.h -- synthetic single-molecule frame stacks for tests and benchmarks.
 *
 * Frames follow the reference's point-sampled forward model
 *     lambda(x, y) = b + A exp(-(x-x0)^2/(2 sx^2) - (y-y0)^2/(2 sy^2)),  A = N/(2 pi sx sy)
 * (the symmetric case matches integral = 2 pi r^2 A, source/Localization.cpp:202),
 * with Poisson noise at gain 1, stored as uint16 with x as the fast axis
 * (source/FileIO.h:96-116).  Candidates mimic the reference finder's output
 * (source/ParticleFinding.cpp:170-196): position = true position rounded to the
 * nearest pixel, intensity = pixel value there minus the background estimate, and
 * background = mean of the frame pixels outside every candidate window.
 *
 * Random numbers: Philox-4x32-10 (Salmon et al., SC'11), key = (seed, config id),
 * counter = (frame, emitter or pixel, draw, 0), so every value is reproducible
 * independently of threading.
 */
#ifndef LZR_SYNTH_H
#define LZR_SYNTH_H

#include <stdint.h>
#include "../../include/localizer_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct LzrSynthSpec {
    uint32_t seed;           /* master seed (0x10CA112E for the standard workloads) */
    uint32_t config_id;      /* sub-stream key                                        */
    int nx, ny, nframes;
    int emitters_per_frame;
    double sigma_min, sigma_max;     /* symmetric PSF sigma, uniform                  */
    double photons_min, photons_max; /* photon count, log-uniform                     */
    double bg_min, bg_max;           /* per-frame background, uniform                 */
    double bg_gradient;              /* extra background rising linearly 0..g along x */
    int half_width_min, half_width_max; /* candidate window half-width, uniform int   */
    double ellipsoidal_fraction;     /* emitters drawn with an elliptical PSF         */
    double sigma_xy_min, sigma_xy_max; /* their sigma_x, sigma_y, uniform             */
    double pair_fraction;            /* emitters placed near the previous one         */
    double pair_distance_sigmas;     /* max distance of such a pair, in PSF sigmas    */
    int margin;                      /* min distance to the frame edge (0: h_max + 1) */
    int candidate_model_from_truth;  /* 1: candidate.model = emitter's PSF model      */
} LzrSynthSpec;

typedef struct LzrSynthEmitter {
    double x, y;          /* true position (px)       */
    double photons;
    double sigma_x, sigma_y;
    double amplitude;     /* photons / (2 pi sx sy)   */
    double background;    /* frame background at (x,y) */
    int32_t frame;
    int16_t model;        /* LZR_MODEL_SYMMETRIC or LZR_MODEL_ELLIPSOIDAL */
    int16_t half_width;
} LzrSynthEmitter;

/* Philox-4x32-10 block: out = Philox(ctr, key). */
void lzr_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Renders nframes frames (nx*ny uint16 each), writes emitters_per_frame truths and
 * candidates per frame (frame-major).  nthreads <= 0: all cores.  Returns 0 or -1. */
int lzr_synth_generate(const LzrSynthSpec *spec, uint16_t *frames,
                       LzrSynthEmitter *truth, LzrCandidate *candidates, int nthreads);

/* Simulates independent n x n ROIs (n = 2h+1) centred near the window centre with the
 * sub-pixel offset drawn uniformly in [-max_offset, max_offset]; frame index of
 * candidate i is i (each ROI is its own 1-frame image of n x n).  Returns 0 or -1. */
int lzr_synth_rois(const LzrSynthSpec *spec, int nrois, int half_width,
                   double max_offset, uint16_t *rois, LzrSynthEmitter *truth,
                   LzrCandidate *candidates, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
