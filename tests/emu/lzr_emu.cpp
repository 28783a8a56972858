// Host build of the GPU lane logic (localizer_amd/csrc/lzr_lane.h) for the CPU tests:
// every candidate runs the same state machine and the same pixel pass as one GPU lane,
// sequentially.  TEST INFRASTRUCTURE ONLY -- the product never calls it.
#include <cmath>
#include <cstring>

#include "../../localizer_amd/csrc/lzr_lane.h"

using namespace lzr;

template <int P>
static void fit_one(const LaneConfig &cfg, const uint16_t *frames, int nx, int ny, int nframes,
                    const LzrCandidate &c, int idx, int model, LzrFit &out) {
    LaneState<P> s;
    memset(&out, 0, sizeof out);
    if (!lane_start<P>(s, cfg, c, idx, nx, ny, nframes, model, out)) return;
    const int n = s.n;
    float D[(2 * LZR_MAX_HALF_WIDTH + 1) * (2 * LZR_MAX_HALF_WIDTH + 1)];
    const uint16_t *src = frames + ((size_t)s.frame * ny + s.yoff) * nx + s.xoff;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) D[i + n * j] = (float)src[(size_t)j * nx + i];
    for (;;) {
        PassAcc<P, double> a;
        a.zero();
        struct {
            const float *D;
            int n;
            float operator()(int i, int j) const { return D[i + n * j]; }
            float raw(int i, int j) const { return D[i + n * j]; }
        } pix{D, n};
        struct {
            double v[2 * LZR_MAX_HALF_WIDTH + 1 + kXbEySlots];
            void set(int i, double x) { v[i] = x; }
            double get(int i) const { return v[i]; }
        } xb;
        if constexpr (P == 7) {
            EllIn q;
            lane_ell_in(s, cfg, q);
            double v[kEllSums];
            ell_pass(q, n, n, false, pix, xb, v);
            ell_finish<double>(q, n * n, v, s.phase == kInit, s.fnorm * s.fnorm, a);
        } else {
            SepIn q;
            lane_sep_in<P>(s, cfg, q);
            double v[kSepSums];
            sep_pass(q, n, n, false, pix, xb, v, xb_ey_slot(cfg));
            sep_finish<P, double>(q, n * n, v, s.phase == kInit, s.fnorm * s.fnorm, a);
        }
        if (!lane_after_pass<P, double>(s, cfg, a, xb, model, reinterpret_cast<LzrRawFit *>(&out))) {
            finish_any(out, cfg, c);
            return;
        }
    }
}

static LaneConfig make_cfg(int model, double w, double sigma, int default_h, int max_h, int diag_rows, int max_iter,
                           double epsabs, double epsrel) {
    LaneConfig cfg;
    cfg.initial_psf_width = w;
    cfg.sigma = sigma;
    cfg.inv_sigma = 1.0 / sigma;
    cfg.epsabs = epsabs;
    cfg.epsrel = epsrel;
    cfg.max_iterations = max_iter;
    cfg.default_half_width = default_h;
    cfg.max_half_width = max_h;
    cfg.diag_update_rows = diag_rows;
    cfg.model = model;
    return cfg;
}

extern "C" int emu_fit_batch(int model, double w, double sigma, int default_h, int max_h, int diag_rows,
                             int max_iter, double epsabs, double epsrel, const uint16_t *frames, int nx, int ny,
                             int nframes, const LzrCandidate *cands, int ncands, LzrFit *out) {
    const LaneConfig cfg = make_cfg(model, w, sigma, default_h, max_h, diag_rows, max_iter, epsabs, epsrel);
    int accepted = 0;
    for (int i = 0; i < ncands; ++i) {
        const int m = cands[i].model >= 0 ? cands[i].model : model;
        if (m == LZR_MODEL_SYMMETRIC)
            fit_one<5>(cfg, frames, nx, ny, nframes, cands[i], i, m, out[i]);
        else if (m == LZR_MODEL_FIXED_WIDTH)
            fit_one<4>(cfg, frames, nx, ny, nframes, cands[i], i, m, out[i]);
        else
            fit_one<7>(cfg, frames, nx, ny, nframes, cands[i], i, m, out[i]);
        accepted += out[i].status == LZR_ACCEPTED;
    }
    return accepted;
}

// |f| at each candidate's reference start point (start_residual: the net-start fallback's
// yardstick); NaN for a window outside the frame.
extern "C" void emu_start_residual(int model, double w, double sigma, int default_h, int max_h, const uint16_t *frames,
                                   int nx, int ny, int nframes, const LzrCandidate *cands, int ncands, double *out) {
    const LaneConfig cfg = make_cfg(model, w, sigma, default_h, max_h, -1, 200, 1e-4, 1e-4);
    for (int i = 0; i < ncands; ++i) out[i] = start_residual(cfg, cands[i], frames, nx, ny, nframes);
}

// One separable pass (symmetric P = 5, fixed width P = 4) at theta + dx over an n x n
// window D (x fast), looping to nloop >= n with the pixels beyond n masked as in a wave
// of mixed window sizes.  out = [hq (packed lower, P-1 columns), sj, gq, sf, |f1|^2].
template <int P>
static void sep_one(const double *theta, const double *dx, double sigma, double w, int n, int nloop,
                    const uint16_t *D, double *out) {
    double t0[P], st[P];
    for (int i = 0; i < P; ++i) {
        t0[i] = theta[i];
        st[i] = dx[i];
    }
    SepIn q;
    make_sep_in(t0, st, 1.0 / sigma, w, q);
    struct {
        const uint16_t *D;
        int n;
        uint16_t raw(int i, int j) const { return i < n && j < n ? D[i + n * j] : (uint16_t)0; }
    } pix{D, n};
    struct {
        double v[64];
        void set(int i, double x) { v[i] = x; }
        double get(int i) const { return v[i]; }
    } xb;
    double v[kSepSums];
    sep_pass(q, n, nloop, nloop != n, pix, xb, v, 62);
    PassAcc<P, double> a;
    a.zero();
    sep_finish<P, double>(q, n * n, v, true, 0.0, a);
    int k = 0;
    for (int i = 0; i < NP<P>::SQ; ++i) out[k++] = a.hq[i];
    for (int i = 0; i < P - 1; ++i) out[k++] = a.sj[i];
    for (int i = 0; i < P - 1; ++i) out[k++] = a.gq[i];
    out[k++] = a.sf;
    out[k++] = a.sse0;
}

// One ellipsoidal moment pass at theta + dx (A, sx, sy, x0, y0, rho, b) over an n x n
// window, loop bound nloop >= n.  out = [hq (21, packed lower), sj (6), gq (6), sf, |f1|^2].
extern "C" int emu_ell_pass(const double *theta, const double *dx, double sigma, int n, int nloop,
                            const uint16_t *D, double *out) {
    if (n < 1 || nloop < n || nloop > 64) return -1;
    double t0[7], st[7];
    for (int i = 0; i < 7; ++i) {
        t0[i] = theta[i];
        st[i] = dx[i];
    }
    EllIn q;
    make_ell_in(t0, st, 1.0 / sigma, q);
    struct {
        const uint16_t *D;
        int n;
        uint16_t raw(int i, int j) const { return i < n && j < n ? D[i + n * j] : (uint16_t)0; }
    } pix{D, n};
    struct {
        double v[64];
        void set(int i, double x) { v[i] = x; }
        double get(int i) const { return v[i]; }
    } xb;
    double v[kEllSums];
    ell_pass(q, n, nloop, nloop != n, pix, xb, v);
    PassAcc<7, double> a;
    a.zero();
    ell_finish<double>(q, n * n, v, true, 0.0, a);
    int k = 0;
    for (int i = 0; i < NP<7>::SQ; ++i) out[k++] = a.hq[i];
    for (int i = 0; i < 6; ++i) out[k++] = a.sj[i];
    for (int i = 0; i < 6; ++i) out[k++] = a.gq[i];
    out[k++] = a.sf;
    out[k++] = a.sse0;
    return 0;
}

extern "C" int emu_sep_pass(int P, const double *theta, const double *dx, double sigma, double w, int n, int nloop,
                            const uint16_t *D, double *out) {
    if (n < 1 || nloop < n || nloop > 64) return -1;
    if (P == 5)
        sep_one<5>(theta, dx, sigma, w, n, nloop, D, out);
    else if (P == 4)
        sep_one<4>(theta, dx, sigma, w, n, nloop, D, out);
    else
        return -1;
    return 0;
}

// The pass of a straggler group of k lanes, emulated lane by lane on the host: every lane
// evaluates the column factors and the row recurrence itself, lane r sums the rows r, r + k,
// ..., and every lane folds the rows' results in row order after the exchange.  The result
// must equal the one-lane pass bit for bit for every k (the kernel's claim).
// out as emu_sep_pass / emu_ell_pass.
template <int P>
static int sep_group_one(const double *theta, const double *dx, double sigma, double w, int n, int k,
                          const uint16_t *D, double *out) {
    double t0[P], st[P];
    for (int i = 0; i < P; ++i) {
        t0[i] = theta[i];
        st[i] = dx[i];
    }
    SepIn q;
    make_sep_in(t0, st, 1.0 / sigma, w, q);
    struct XB {
        double v[64];
        void set(int i, double x) { v[i] = x; }
        double get(int i) const { return v[i]; }
    };
    struct Row {
        const uint16_t *row;
        uint16_t raw(int i, int) const { return row[i]; }
    };
    const int rounds = (n + k - 1) / k;
    double lane_v[16][kSepSums];
    XB lane_xb[16];
    GaussRec lane_gy[16];
    for (int r = 0; r < k; ++r) sep_columns(q, n, n, lane_xb[r], lane_v[r], lane_gy[r]);
    for (int m = 0; m < rounds; ++m) {
        double xe[16];
        SepRow xr[16];
        for (int r = 0; r < k; ++r) {
            double eo = 0.0;
            for (int rr = 0; rr < k; ++rr) {
                const int j = m * k + rr;
                if (j < n) {
                    const double ey = sep_row_factor(q, lane_gy[r], j, n, lane_v[r]);
                    if (rr == r) eo = ey;
                }
            }
            const int jo = m * k + r;
            SepRow ro{0.0, 0.0, 0.0, 0.0, 0.0};
            if (jo < n) ro = sep_row<false>(q, n, n, true, Row{D + n * jo}, lane_xb[r], jo);
            xe[r] = eo;
            xr[r] = ro;
        }
        for (int r = 0; r < k; ++r)
            for (int rr = 0; rr < k; ++rr) {
                const int j = m * k + rr;
                if (j < n) row_sums(xe[rr], (double)j - q.y1, xr[rr].r1, xr[rr].ru, xr[rr].ru2, xr[rr].rd, xr[rr].rdd,
                                    lane_v[r]);
            }
    }
    for (int r = 1; r < k; ++r)
        if (memcmp(lane_v[r], lane_v[0], sizeof lane_v[0]) != 0) return -2; /* lanes disagree */
    PassAcc<P, double> a;
    a.zero();
    sep_finish<P, double>(q, n * n, lane_v[0], true, 0.0, a);
    int o = 0;
    for (int i = 0; i < NP<P>::SQ; ++i) out[o++] = a.hq[i];
    for (int i = 0; i < P - 1; ++i) out[o++] = a.sj[i];
    for (int i = 0; i < P - 1; ++i) out[o++] = a.gq[i];
    out[o++] = a.sf;
    out[o++] = a.sse0;
    return 0;
}

extern "C" int emu_sep_pass_group(int P, const double *theta, const double *dx, double sigma, double w, int n, int k,
                                  const uint16_t *D, double *out) {
    if (n < 1 || n > 63 || k < 1 || k > 16) return -1;
    if (P == 5) return sep_group_one<5>(theta, dx, sigma, w, n, k, D, out);
    if (P == 4) return sep_group_one<4>(theta, dx, sigma, w, n, k, D, out);
    return -1;
}

extern "C" int emu_ell_pass_group(const double *theta, const double *dx, double sigma, int n, int k,
                                  const uint16_t *D, double *out) {
    if (n < 1 || n > 63 || k < 1 || k > 16) return -1;
    double t0[7], st[7];
    for (int i = 0; i < 7; ++i) {
        t0[i] = theta[i];
        st[i] = dx[i];
    }
    EllIn q;
    make_ell_in(t0, st, 1.0 / sigma, q);
    struct XB {
        double v[64];
        void set(int i, double x) { v[i] = x; }
        double get(int i) const { return v[i]; }
    };
    struct Row {
        const uint16_t *row;
        uint16_t raw(int i, int) const { return row[i]; }
    };
    const int rounds = (n + k - 1) / k;
    double lane_s[16][kEllSums];
    XB lane_xb[16];
    EllRows lane_rw[16];
    for (int r = 0; r < k; ++r) ell_columns(q, n, n, lane_xb[r], lane_s[r], lane_rw[r]);
    for (int m = 0; m < rounds; ++m) {
        double xw[16];
        EllRow xr[16];
        for (int r = 0; r < k; ++r) {
            double wo = 0.0, go = 1.0, co = 1.0;
            for (int rr = 0; rr < k; ++rr) {
                const int j = m * k + rr;
                if (j < n) {
                    double w, g, c;
                    ell_row_factor(q, lane_rw[r], j, w, g, c);
                    if (rr == r) {
                        wo = w;
                        go = g;
                        co = c;
                    }
                }
            }
            const int jo = m * k + r;
            EllRow ro{{0.0, 0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, 0.0, 0.0, 0.0};
            if (jo < n) ro = ell_row_sums<false>(q, n, n, lane_rw[r].fac, true, go, co, Row{D + n * jo}, lane_xb[r], jo);
            xw[r] = wo;
            xr[r] = ro;
        }
        for (int r = 0; r < k; ++r)
            for (int rr = 0; rr < k; ++rr) {
                const int j = m * k + rr;
                if (j < n)
                    ell_row(xw[rr], (double)j - q.y1, xr[rr].P, xr[rr].R, xr[rr].W, xr[rr].rd, xr[rr].rdd, xr[rr].bad,
                            lane_s[r]);
            }
    }
    for (int r = 1; r < k; ++r)
        if (memcmp(lane_s[r], lane_s[0], sizeof lane_s[0]) != 0) return -2; /* lanes disagree */
    PassAcc<7, double> a;
    a.zero();
    ell_finish<double>(q, n * n, lane_s[0], true, 0.0, a);
    int o = 0;
    for (int i = 0; i < NP<7>::SQ; ++i) out[o++] = a.hq[i];
    for (int i = 0; i < 6; ++i) out[o++] = a.sj[i];
    for (int i = 0; i < 6; ++i) out[o++] = a.gq[i];
    out[o++] = a.sf;
    out[o++] = a.sse0;
    return 0;
}
