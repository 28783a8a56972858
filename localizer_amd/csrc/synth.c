/*
 * synth.c -- synthetic frame stacks and ROIs (see lzr_synth.h).
 */
#define _GNU_SOURCE
#include "lzr_synth.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define SYN_PI 3.14159265358979323846

void lzr_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        const uint32_t n3 = (uint32_t)p0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* (0,1) doubles: 53 bits from two words, 24 bits from one. */
static double u53(uint32_t a, uint32_t b) {
    const uint64_t m = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
    return ((double)m + 0.5) * (1.0 / 9007199254740992.0);
}
static double u24(uint32_t a) { return ((double)(a >> 8) + 0.5) * (1.0 / 16777216.0); }

typedef struct Stream {
    uint32_t ctr[4], key[2], buf[4];
    int used;
} Stream;

static void stream_init(Stream *s, const uint32_t key[2], uint32_t c0, uint32_t c1) {
    s->key[0] = key[0];
    s->key[1] = key[1];
    s->ctr[0] = c0;
    s->ctr[1] = c1;
    s->ctr[2] = 0;
    s->ctr[3] = 0;
    s->used = 4;
}

static double stream_uniform(Stream *s) {
    if (s->used >= 4) {
        lzr_philox4x32_10(s->ctr, s->key, s->buf);
        s->ctr[2]++;
        s->used = 0;
    }
    return u24(s->buf[s->used++]);
}

/* Poisson sampler: multiplication method below 10, else PTRS (Hormann 1993). */
static int poisson(Stream *s, double lam) {
    if (!(lam > 0.0)) return 0;
    if (lam < 10.0) {
        const double lim = exp(-lam);
        double prod = stream_uniform(s);
        int k = 0;
        while (prod > lim) {
            prod *= stream_uniform(s);
            ++k;
        }
        return k;
    }
    const double slam = sqrt(lam), loglam = log(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (;;) {
        const double U = stream_uniform(s) - 0.5;
        const double V = stream_uniform(s);
        const double us = 0.5 - fabs(U);
        const double k = floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return (int)k;
        if (k < 0.0 || (us < 0.013 && V > us)) continue;
        if (log(V) + log(inv_alpha) - log(a / (us * us) + b) <= -lam + k * loglam - lgamma(k + 1.0))
            return (int)k;
    }
}

static uint16_t clip16(int v) { return (uint16_t)(v < 0 ? 0 : (v > 65535 ? 65535 : v)); }

typedef struct Job {
    const LzrSynthSpec *spec;
    uint16_t *frames;
    LzrSynthEmitter *truth;
    LzrCandidate *cands;
    int next;
    int total;
    int half_width; /* ROI mode */
    double max_offset;
    int roi_mode;
} Job;

static void draw_emitter(const LzrSynthSpec *sp, const uint32_t key[2], int f, int e,
                         const LzrSynthEmitter *prev, int margin, LzrSynthEmitter *em) {
    uint32_t ctr[4] = {(uint32_t)f, (uint32_t)e, 0, 0}, r0[4], r1[4], r2[4], r3[4];
    lzr_philox4x32_10(ctr, key, r0);
    ctr[2] = 1;
    lzr_philox4x32_10(ctr, key, r1);
    ctr[2] = 2;
    lzr_philox4x32_10(ctr, key, r2);
    ctr[2] = 3;
    lzr_philox4x32_10(ctr, key, r3);
    const double lo_x = margin, hi_x = sp->nx - 1 - margin;
    const double lo_y = margin, hi_y = sp->ny - 1 - margin;
    em->frame = f;
    em->x = lo_x + u53(r0[0], r0[1]) * (hi_x - lo_x);
    em->y = lo_y + u53(r0[2], r0[3]) * (hi_y - lo_y);
    const double lp0 = log(sp->photons_min), lp1 = log(sp->photons_max);
    em->photons = exp(lp0 + u53(r1[0], r1[1]) * (lp1 - lp0));
    const int ell = u24(r2[0]) < sp->ellipsoidal_fraction;
    if (ell) {
        em->model = LZR_MODEL_ELLIPSOIDAL;
        em->sigma_x = sp->sigma_xy_min + u53(r1[2], r1[3]) * (sp->sigma_xy_max - sp->sigma_xy_min);
        em->sigma_y = sp->sigma_xy_min + u24(r2[1]) * (sp->sigma_xy_max - sp->sigma_xy_min);
    } else {
        em->model = LZR_MODEL_SYMMETRIC;
        em->sigma_x = sp->sigma_min + u53(r1[2], r1[3]) * (sp->sigma_max - sp->sigma_min);
        em->sigma_y = em->sigma_x;
    }
    const int hspan = sp->half_width_max - sp->half_width_min + 1;
    int hh = sp->half_width_min + (int)(u24(r2[2]) * hspan);
    if (hh > sp->half_width_max) hh = sp->half_width_max;
    em->half_width = (int16_t)hh;
    if (prev && u24(r2[3]) < sp->pair_fraction) {
        const double ang = 2.0 * SYN_PI * u53(r3[0], r3[1]);
        const double d = u53(r3[2], r3[3]) * sp->pair_distance_sigmas * em->sigma_x;
        double x = prev->x + d * cos(ang), y = prev->y + d * sin(ang);
        em->x = x < lo_x ? lo_x : (x > hi_x ? hi_x : x);
        em->y = y < lo_y ? lo_y : (y > hi_y ? hi_y : y);
    }
    em->amplitude = em->photons / (2.0 * SYN_PI * em->sigma_x * em->sigma_y);
}

static void render_add(double *lam, int nx, int ny, const LzrSynthEmitter *em) {
    const double smax = em->sigma_x > em->sigma_y ? em->sigma_x : em->sigma_y;
    const int R = (int)ceil(6.0 * smax) + 1;
    const int cx = (int)floor(em->x + 0.5), cy = (int)floor(em->y + 0.5);
    const double ix = 1.0 / (2.0 * em->sigma_x * em->sigma_x);
    const double iy = 1.0 / (2.0 * em->sigma_y * em->sigma_y);
    for (int y = cy - R; y <= cy + R; ++y) {
        if (y < 0 || y >= ny) continue;
        const double dy = y - em->y;
        for (int x = cx - R; x <= cx + R; ++x) {
            if (x < 0 || x >= nx) continue;
            const double dx = x - em->x;
            lam[(long)y * nx + x] += em->amplitude * exp(-dx * dx * ix - dy * dy * iy);
        }
    }
}

static void make_frame(Job *J, int f, double *lam, unsigned char *mask) {
    const LzrSynthSpec *sp = J->spec;
    const uint32_t key[2] = {sp->seed, sp->config_id};
    const int nx = sp->nx, ny = sp->ny, E = sp->emitters_per_frame;
    const long npx = (long)nx * ny;
    int margin = sp->margin > 0 ? sp->margin : sp->half_width_max + 1;
    uint32_t bctr[4] = {(uint32_t)f, 0xFFFFFFFFu, 0, 0}, br[4];
    lzr_philox4x32_10(bctr, key, br);
    const double bg = sp->bg_min + u24(br[0]) * (sp->bg_max - sp->bg_min);
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x)
            lam[(long)y * nx + x] = bg + (nx > 1 ? sp->bg_gradient * (double)x / (double)(nx - 1) : 0.0);
    LzrSynthEmitter *tr = J->truth + (long)f * E;
    for (int e = 0; e < E; ++e) {
        draw_emitter(sp, key, f, e, e > 0 ? &tr[e - 1] : NULL, margin, &tr[e]);
        tr[e].background = bg + (nx > 1 ? sp->bg_gradient * tr[e].x / (double)(nx - 1) : 0.0);
        render_add(lam, nx, ny, &tr[e]);
    }
    uint16_t *img = J->frames + (long)f * npx;
    for (long k = 0; k < npx; ++k) {
        Stream s;
        stream_init(&s, key, (uint32_t)f, 0x80000000u | (uint32_t)k);
        img[k] = clip16(poisson(&s, lam[k]));
    }
    /* background estimate: mean of pixels outside every candidate window */
    memset(mask, 0, (size_t)npx);
    for (int e = 0; e < E; ++e) {
        const int cx = (int)round(tr[e].x), cy = (int)round(tr[e].y), h = tr[e].half_width;
        for (int y = cy - h; y <= cy + h; ++y)
            for (int x = cx - h; x <= cx + h; ++x)
                if (x >= 0 && x < nx && y >= 0 && y < ny) mask[(long)y * nx + x] = 1;
    }
    double sum = 0.0;
    long cnt = 0;
    for (long k = 0; k < npx; ++k)
        if (!mask[k]) {
            sum += img[k];
            cnt++;
        }
    const double bhat = cnt ? sum / (double)cnt : 0.0;
    LzrCandidate *cd = J->cands + (long)f * E;
    for (int e = 0; e < E; ++e) {
        const double cx = round(tr[e].x), cy = round(tr[e].y);
        cd[e].x = cx;
        cd[e].y = cy;
        cd[e].intensity = (double)img[(long)cy * nx + (long)cx] - bhat;
        cd[e].background = bhat;
        cd[e].frame = f;
        cd[e].half_width = sp->half_width_min > 0 ? tr[e].half_width : 0;
        cd[e].model = sp->candidate_model_from_truth ? tr[e].model : -1;
    }
}

static void make_roi(Job *J, int i) {
    const LzrSynthSpec *sp = J->spec;
    const uint32_t key[2] = {sp->seed, sp->config_id};
    const int h = J->half_width, n = 2 * h + 1;
    uint32_t ctr[4] = {(uint32_t)i, 0x40000000u, 0, 0}, r0[4], r1[4];
    lzr_philox4x32_10(ctr, key, r0);
    ctr[2] = 1;
    lzr_philox4x32_10(ctr, key, r1);
    LzrSynthEmitter *em = &J->truth[i];
    em->frame = i;
    em->model = LZR_MODEL_SYMMETRIC;
    em->half_width = (int16_t)h;
    em->x = h + (2.0 * u53(r0[0], r0[1]) - 1.0) * J->max_offset;
    em->y = h + (2.0 * u53(r0[2], r0[3]) - 1.0) * J->max_offset;
    const double lp0 = log(sp->photons_min), lp1 = log(sp->photons_max);
    em->photons = exp(lp0 + u24(r1[0]) * (lp1 - lp0));
    em->sigma_x = em->sigma_y = sp->sigma_min + u24(r1[1]) * (sp->sigma_max - sp->sigma_min);
    em->background = sp->bg_min + u24(r1[2]) * (sp->bg_max - sp->bg_min);
    em->amplitude = em->photons / (2.0 * SYN_PI * em->sigma_x * em->sigma_y);
    uint16_t *img = J->frames + (long)i * n * n;
    const double iw = 1.0 / (2.0 * em->sigma_x * em->sigma_x);
    double dsum = 0.0;
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            const double dx = x - em->x, dy = y - em->y;
            const double lam = em->background + em->amplitude * exp(-(dx * dx + dy * dy) * iw);
            Stream s;
            stream_init(&s, key, (uint32_t)i, 0x80000000u | (uint32_t)(x + n * y));
            img[x + n * y] = clip16(poisson(&s, lam));
            if (x == 0 || y == 0 || x == n - 1 || y == n - 1) dsum += img[x + n * y];
        }
    /* background estimate: mean of the ROI border; start at the centre pixel */
    const double bhat = dsum / (double)(4 * (n - 1));
    LzrCandidate *c = &J->cands[i];
    c->x = h;
    c->y = h;
    c->intensity = (double)img[h + n * h] - bhat;
    c->background = bhat;
    c->frame = i;
    c->half_width = (int16_t)h;
    c->model = LZR_MODEL_SYMMETRIC;
}

static void *worker(void *arg) {
    Job *J = (Job *)arg;
    double *lam = NULL;
    unsigned char *mask = NULL;
    if (!J->roi_mode) {
        const long npx = (long)J->spec->nx * J->spec->ny;
        lam = (double *)malloc(sizeof(double) * npx);
        mask = (unsigned char *)malloc((size_t)npx);
    }
    for (;;) {
        const int f = __atomic_fetch_add(&J->next, 1, __ATOMIC_RELAXED);
        if (f >= J->total) break;
        if (J->roi_mode)
            make_roi(J, f);
        else
            make_frame(J, f, lam, mask);
    }
    free(lam);
    free(mask);
    return NULL;
}

static void run_jobs(Job *J, int nthreads) {
    if (nthreads <= 0) nthreads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > J->total) nthreads = J->total > 0 ? J->total : 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
    for (int t = 1; t < nthreads; ++t) pthread_create(&th[t], NULL, worker, J);
    worker(J);
    for (int t = 1; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th);
}

int lzr_synth_generate(const LzrSynthSpec *spec, uint16_t *frames, LzrSynthEmitter *truth,
                       LzrCandidate *candidates, int nthreads) {
    if (!spec || spec->nx < 1 || spec->ny < 1 || spec->nframes < 0 || spec->emitters_per_frame < 0)
        return -1;
    int margin = spec->margin > 0 ? spec->margin : spec->half_width_max + 1;
    if (2 * margin >= spec->nx || 2 * margin >= spec->ny) return -1;
    Job J;
    memset(&J, 0, sizeof J);
    J.spec = spec;
    J.frames = frames;
    J.truth = truth;
    J.cands = candidates;
    J.total = spec->nframes;
    run_jobs(&J, nthreads);
    return 0;
}

int lzr_synth_rois(const LzrSynthSpec *spec, int nrois, int half_width, double max_offset,
                   uint16_t *rois, LzrSynthEmitter *truth, LzrCandidate *candidates, int nthreads) {
    if (!spec || nrois < 0 || half_width < 1 || half_width > LZR_MAX_HALF_WIDTH) return -1;
    Job J;
    memset(&J, 0, sizeof J);
    J.spec = spec;
    J.frames = rois;
    J.truth = truth;
    J.cands = candidates;
    J.total = nrois;
    J.roi_mode = 1;
    J.half_width = half_width;
    J.max_offset = max_offset;
    run_jobs(&J, nthreads);
    return 0;
}
