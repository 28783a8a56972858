/*
 * lzr_lane.h -- the per-spot Levenberg-Marquardt state machine of the MI355X engine.
 *
 * One spot (ROI) is fitted by one GPU lane.  The lane runs the same algorithm as
 * GSL 1.16 gsl_multifit_fdfsolver_lmsder (source/Localization.cpp:75, :145-155), the
 * MINPACK lmder trust-region method with automatic column scaling, restated on the
 * normal equations instead of a QR factorisation of J:
 *
 *   MINPACK quantity               here (H = J^T J, g = J^T f, D = diag)
 *   Gauss-Newton step R P^T x=Q^T f  H x = g                      (Cholesky)
 *   qrsolv([R; sqrt(par) D])        (H + par D^2) x = g            (Cholesky)
 *   |R^-T P^T u|^2 (parl, parc)     u^T (H + par D^2)^-1 u
 *   |D^-1 R^T Q^T f| (paru)         |D^-1 g|
 *   |R P^T dx| (predicted red.)     sqrt(dx^T H dx)
 *   rank deficiency (nsing < p)     Cholesky breakdown
 * so the trial points, trust radius, Levenberg-Marquardt parameter, acceptance
 * ratio, status codes and the reference driver's test_delta / 200-iteration loop
 * are those of the reference.  The p x p solves run in FP64.
 *
 * Pixel passes.  Symmetric and fixed-width models (separable exponential): the pass
 * at the trial point is a set of FP64 1-D moments over the window's columns and rows
 * plus five FP64 data sums over the pixels, from which J^T J, J^T f and |f|^2 follow
 * in closed form (sep_pass / sep_finish: seven FP64 operations per pixel plus the pixel's
 * conversion and column offset, pixels read from LDS three at a time).
 * Ellipsoidal model: its Jacobian columns are the exponential times polynomials of
 * degree <= 2 in the pixel offsets, so the pass accumulates FP64 moments of e1 and e1^2
 * up to degree 4 row by row (the cross factor exp(-2 qc u v) by recurrence along a row)
 * and forms the normal equations from them (ell_pass / ell_finish).
 *
 * The file compiles as HIP device code (kernel) and as plain C++ (the CPU
 * emulation used by the tests to check the logic against the FP64 oracle).
 */
#ifndef LZR_LANE_H
#define LZR_LANE_H

#include <stdint.h>

#include "../../include/localizer_hip.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define LZR_HD __host__ __device__
#define LZR_INLINE __forceinline__
#else
#include <cmath>
#define LZR_HD
#define LZR_INLINE inline
#endif

namespace lzr {

constexpr double kPi = 3.1415926535897932384626433; /* source/Defines.h:97 */
constexpr double kSqrt2 = 1.4142135623730950488;    /* source/Defines.h:98 */
constexpr double kDblEpsilon = 2.2204460492503131e-16;
constexpr double kDblMin = 2.2250738585072014e-308;

enum Phase : int { kIdle = 0, kInit = 1, kTrial = 2, kFinished = 3 };

/* Fitter-wide settings (the reference constructor arguments plus the loop constants
 * source/Localization.cpp:154-155). */
struct LaneConfig {
    double initial_psf_width; /* w */
    double sigma;             /* residual weight */
    double inv_sigma;         /* 1.0 / sigma, rounded once on the host */
    double epsabs, epsrel;    /* test_delta tolerances */
    int max_iterations;       /* 200 */
    int default_half_width;   /* window half-width when a candidate gives none */
    int max_half_width;       /* largest window this launch holds */
    int diag_update_rows;     /* -1: p rows (GSL 1.16), 0: all rows (MINPACK) */
    int model;                /* default model */
};

constexpr LZR_HD int tri(int i, int j) { return i >= j ? (i * (i + 1)) / 2 + j : (j * (j + 1)) / 2 + i; }

template <int P>
struct NP {
    static constexpr int S = P * (P + 1) / 2; /* packed symmetric P x P */
    static constexpr int Q = P - 1;           /* columns other than the offset b */
    static constexpr int SQ = Q * (Q + 1) / 2;
};

/* Raw end state of one fit, written by the fitting kernel into the caller's LzrFit
 * slot (same size and layout of the trailing fields) and turned into the reference
 * record by the epilogue (finish_record).  th is in window coordinates (x0, y0 relative to
 * the window's corner); the epilogue rebuilds J^T J at th from the window. */
struct LzrRawFit {
    double th[LZR_MAX_PARAMS];
    double unused[2 * LZR_MAX_PARAMS];
    double chi;
    int32_t status; /* kPending */
    int32_t lm_status;
    int16_t iterations, trials, accepted_steps, model;
};
static_assert(sizeof(LzrRawFit) == sizeof(LzrFit), "raw record must overlay LzrFit");
static_assert(__builtin_offsetof(LzrRawFit, status) == __builtin_offsetof(LzrFit, status), "status overlays status");
constexpr int32_t kPending = -1;
/* A record whose fit passed the post-checks and whose covariance waits for the QR factor
 * (finish_deferred): its raw parameters are kept in the deviations' slots. */
constexpr int32_t kDeferredQr = -2;

/* Per-lane LM state.  x0/y0 are kept window-local (exact integer offsets xoff/yoff);
 * every quantity the reference computes from absolute coordinates (test_delta,
 * |D x|) uses xoff + local. */
template <int P>
struct LaneState {
    double th[P];   /* current parameters                           */
    double dx[P];   /* last step (GSL s->dx)                        */
    double diag[P]; /* column scaling D                             */
    double H[NP<P>::S]; /* J^T J at th (packed)                     */
    double g[P];        /* J^T f at th                              */
    double delta, par, fnorm, xnorm;
    double pnorm;   /* |D dx| of the proposed step (lmpar's dxnorm) */
    int xoff, yoff, n, frame;
    int roi;
    int phase;
    int iter_gsl;   /* GSL state->iter: 1 + accepted steps          */
    int iterations; /* reference loop counter                        */
    int inner;      /* trials inside the current iterate()          */
    int trials, accepted;
};

/* Accumulators of one pass.  The offset column of J is the constant 1/sigma, so its
 * normal-equation entries come from plain sums. */
template <int P, typename Acc>
struct PassAcc {
    Acc hq[NP<P>::SQ]; /* sum J_i J_j, i,j < P-1 */
    Acc sj[P - 1];     /* sum J_i                */
    Acc gq[P - 1];     /* sum J_i f1             */
    Acc sf;            /* sum f1                 */
    Acc sse0;          /* sum f0^2               */
    Acc dsse;          /* sum (f1-f0)(f1+f0)     */
    int bad;           /* residual vector at the pass point: 0 finite, 1 one +-inf entry, 2 NaN norm */
    LZR_HD LZR_INLINE void zero() {
        bad = 0;
#pragma unroll
        for (int i = 0; i < NP<P>::SQ; ++i) hq[i] = Acc(0);
#pragma unroll
        for (int i = 0; i < P - 1; ++i) {
            sj[i] = Acc(0);
            gq[i] = Acc(0);
        }
        sf = sse0 = dsse = Acc(0);
    }
};

/* Per-pixel arithmetic is written with explicit fused multiply-adds and compiled with
 * contraction off: every inlined copy of a pixel function (the loop variants of a pass,
 * the three kernels, the host emulation) then evaluates exactly the same operations, so a
 * fit's result never depends on which copy ran it. */
#if defined(__clang__)
#define LZR_NO_CONTRACT _Pragma("clang fp contract(off)")
#else
#define LZR_NO_CONTRACT
#endif

LZR_HD LZR_INLINE double dfma(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_fma(a, b, c);
#else
    return std::fma(a, b, c);
#endif
}

LZR_HD LZR_INLINE float fexp(float a) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __expf(a);
#else
    return std::exp(a);
#endif
}

LZR_HD LZR_INLINE double dexp(double a) {
#if defined(__HIP_DEVICE_COMPILE__)
    return exp(a);
#else
    return std::exp(a);
#endif
}

/* ---- FP64 helpers ---------------------------------------------------------- */

LZR_HD LZR_INLINE double dsqrt(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_sqrt(v);
#else
    return std::sqrt(v);
#endif
}
/* 1/sqrt(v) for v > 0: the hardware estimate refined by two Newton steps
 * (r <- r + r (1/2 - v r^2 / 2)), within ~1 ulp; one quarter-rate instruction and six
 * FMAs instead of a correctly rounded square root followed by a division. */
LZR_HD LZR_INLINE double drsqrt(double v) {
    LZR_NO_CONTRACT
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rsq(v);
    const double h = 0.5 * v;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const double e = __builtin_fma(-(h * r), r, 0.5);
        r = __builtin_fma(r, e, r);
    }
    return r;
#else
    return 1.0 / std::sqrt(v);
#endif
}
/* 1/v: the hardware estimate refined by two Newton steps (e = 1 - v r, r <- r + r e),
 * within ~1 ulp; 0, +-inf and NaN take the estimate as is (inf, 0, NaN).  Used for the
 * LM logic's divisions, where a last-bit difference from a correctly rounded quotient
 * is far below every tolerance the reference tests. */
LZR_HD LZR_INLINE double drcp(double v) {
    LZR_NO_CONTRACT
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(v);
    if (__builtin_isfinite(r) && r != 0.0) {
        double e = __builtin_fma(-v, r, 1.0);
        r = __builtin_fma(r, e, r);
        e = __builtin_fma(-v, r, 1.0);
        r = __builtin_fma(r, e, r);
    }
    return r;
#else
    return 1.0 / v;
#endif
}
LZR_HD LZR_INLINE double ddiv(double a, double b) {
    LZR_NO_CONTRACT
    return a * drcp(b);
}
LZR_HD LZR_INLINE double dabs(double v) { return v < 0 ? -v : v; }
LZR_HD LZR_INLINE double dmax(double a, double b) { return a > b ? a : b; }
LZR_HD LZR_INLINE double dmin(double a, double b) { return a < b ? a : b; }

/* Cholesky M = L L^T (packed); false on a non-positive pivot.  M and L may be the same array
 * (column j reads M's column j before writing L's). */
template <int P>
LZR_HD LZR_INLINE bool cholesky(const double (&M)[NP<P>::S], double (&L)[NP<P>::S], double (&inv_diag)[P]) {
    LZR_NO_CONTRACT
#pragma unroll
    for (int j = 0; j < P; ++j) {
        double s = M[tri(j, j)];
#pragma unroll
        for (int k = 0; k < j; ++k) s = dfma(-L[tri(j, k)], L[tri(j, k)], s);
        if (!(s > 0.0)) return false;
        const double inv = drsqrt(s);
        L[tri(j, j)] = s * inv;
        inv_diag[j] = inv;
#pragma unroll
        for (int i = j + 1; i < P; ++i) {
            double t = M[tri(i, j)];
#pragma unroll
            for (int k = 0; k < j; ++k) t = dfma(-L[tri(i, k)], L[tri(j, k)], t);
            L[tri(i, j)] = t * inv;
        }
    }
    return true;
}

/* y = L^-1 b */
template <int P>
LZR_HD LZR_INLINE void forward(const double (&L)[NP<P>::S], const double (&inv_diag)[P], const double (&b)[P],
                               double (&y)[P]) {
    LZR_NO_CONTRACT
#pragma unroll
    for (int i = 0; i < P; ++i) {
        double s = b[i];
#pragma unroll
        for (int k = 0; k < i; ++k) s = dfma(-L[tri(i, k)], y[k], s);
        y[i] = s * inv_diag[i];
    }
}

/* x = L^-T y */
template <int P>
LZR_HD LZR_INLINE void backward(const double (&L)[NP<P>::S], const double (&inv_diag)[P], const double (&y)[P],
                                double (&x)[P]) {
    LZR_NO_CONTRACT
#pragma unroll
    for (int i = P - 1; i >= 0; --i) {
        double s = y[i];
#pragma unroll
        for (int k = i + 1; k < P; ++k) s = dfma(-L[tri(k, i)], x[k], s);
        x[i] = s * inv_diag[i];
    }
}

template <int P>
LZR_HD LZR_INLINE double scaled_norm(const double (&d)[P], const double (&v)[P]) {
    LZR_NO_CONTRACT
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < P; ++i) {
        const double t = d[i] * v[i];
        s = dfma(t, t, s);
    }
    return dsqrt(s);
}

template <int P>
LZR_HD LZR_INLINE double plain_norm2(const double (&v)[P]) {
    LZR_NO_CONTRACT
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < P; ++i) s = dfma(v[i], v[i], s);
    return s;
}

template <int P>
LZR_HD LZR_INLINE double plain_norm(const double (&v)[P]) {
    LZR_NO_CONTRACT
    return dsqrt(plain_norm2<P>(v));
}

/* Levenberg-Marquardt parameter (MINPACK lmpar on the normal equations).
 * Returns x with (H + par D^2) x = g (the LM step is -x) and |D x| in dxnorm_out. */
template <int P>
LZR_HD LZR_INLINE void lmpar(const double (&H)[NP<P>::S], const double (&g)[P], const double (&diag)[P],
                             double delta, double &par_io, double (&x)[P], double &dxnorm_out) {
    LZR_NO_CONTRACT
    constexpr int S = NP<P>::S;
    double L[S], id[P], y[P], u[P];
    /* Gauss-Newton step.  A Cholesky breakdown is MINPACK's rank-deficient case: the
     * undamped step is not usable, the lower bound parl is 0 and the damped iteration
     * below (H + par D^2 is positive definite for par > 0) determines the step. */
    const bool full_rank = cholesky<P>(H, L, id);
    double dxnorm, fp;
    if (full_rank) {
        forward<P>(L, id, g, y);
        backward<P>(L, id, y, x);
        dxnorm = scaled_norm<P>(diag, x);
        fp = dxnorm - delta;
        if (fp <= 0.1 * delta) {
            par_io = 0.0;
            dxnorm_out = dxnorm;
            return;
        }
    } else {
#pragma unroll
        for (int i = 0; i < P; ++i) x[i] = 0.0;
        dxnorm = 0.0;
        fp = 1.0e300;
    }
    double parl = 0.0;
    if (full_rank) {
        const double inv_dxnorm = drcp(dxnorm);
#pragma unroll
        for (int i = 0; i < P; ++i) u[i] = diag[i] * (diag[i] * x[i] * inv_dxnorm);
        forward<P>(L, id, u, y);
        const double phider = plain_norm2<P>(y);
        parl = phider > 0.0 ? ddiv(fp, delta * phider) : 0.0;
    }
#pragma unroll
    for (int i = 0; i < P; ++i) u[i] = ddiv(g[i], diag[i]);
    const double gnorm = plain_norm<P>(u);
    const double inv_delta = drcp(delta);
    double paru = gnorm * inv_delta;
    if (paru == 0.0) paru = kDblMin / dmin(delta, 0.1);
    double par = par_io;
    if (par > paru)
        par = paru;
    else if (par < parl)
        par = parl;
    if (par == 0.0 && full_rank) par = ddiv(gnorm, dxnorm);
    for (int iter = 1;; ++iter) {
        if (par == 0.0) par = dmax(0.001 * paru, kDblMin);
        /* H + par D^2, factored in place (the factor overwrites the matrix column by column) */
#pragma unroll
        for (int i = 0; i < S; ++i) L[i] = H[i];
#pragma unroll
        for (int i = 0; i < P; ++i) L[tri(i, i)] = dfma(par * diag[i], diag[i], L[tri(i, i)]);
        if (!cholesky<P>(L, L, id)) {
            /* (H + par D^2) lost definiteness in rounding: keep the last x */
            break;
        }
        forward<P>(L, id, g, y);
        backward<P>(L, id, y, x);
        dxnorm = scaled_norm<P>(diag, x);
        const double fp_old = fp;
        fp = dxnorm - delta;
        if (dabs(fp) <= 0.1 * delta) break;
        if (parl == 0.0 && fp <= fp_old && fp_old < 0.0) break;
        if (iter == 10) break;
        const double inv_dxnorm = drcp(dxnorm);
#pragma unroll
        for (int i = 0; i < P; ++i) u[i] = diag[i] * (diag[i] * x[i] * inv_dxnorm);
        forward<P>(L, id, u, y);
        const double parc = ddiv(fp, delta * plain_norm2<P>(y));
        if (fp > 0.0) {
            if (par > parl) parl = par;
        } else if (fp < 0.0) {
            if (par < paru) paru = par;
        }
        par = dmax(parl, par + parc);
    }
    par_io = par;
    dxnorm_out = dxnorm;
}

/* ---- model-specific pieces -------------------------------------------------- */

/* Index of x0 in the parameter vector: (A, x0, y0, b), (A, r, x0, y0, b),
 * (A, sx, sy, x0, y0, rho, b). */
template <int P>
LZR_HD LZR_INLINE int x_index() { return P == 4 ? 1 : (P == 5 ? 2 : 3); }

/* Absolute-coordinate parameter vector (for test_delta and |D x|). */
template <int P>
LZR_HD LZR_INLINE void absolute(const LaneState<P> &s, const double (&t)[P], double (&a)[P]) {
    LZR_NO_CONTRACT
#pragma unroll
    for (int i = 0; i < P; ++i) a[i] = t[i];
    a[x_index<P>()] += (double)s.xoff;
    a[x_index<P>() + 1] += (double)s.yoff;
}

LZR_HD LZR_INLINE void quad_coeffs(double sx, double sy, double rho, double &qa, double &qc, double &qb) {
    LZR_NO_CONTRACT
    const double pre = 0.5 * drcp(1.0 - rho * rho);
    const double isx = drcp(sx), isy = drcp(sy);
    qa = pre * isx * isx;
    qb = pre * isy * isy;
    qc = -pre * rho * isx * isy;
}

/* ---- separable-moment pass (symmetric and fixed-width models) -----------------------
 * At the trial point the model's exponential is separable, e1(i, j) = Ex(i) Ey(j) with
 * Ex(i) = exp(-c1 u^2), u = i - x1 (and Ey over rows, v = j - y1).  Every sum of a pass
 * that does not involve the data is therefore a product of 1-D moments over the window's
 * columns and rows,
 *   M_a = sum Ex^2 u^a (a = 0..4),   N_a = sum Ex u^a (a = 0..2),
 * and only five data sums need the 2-D loop over pixels:
 *   U1 = sum e1 D, Uu = sum e1 u D, Uv = sum e1 v D, Ud = sum e1 (u^2 + v^2) D,
 *   SD = sum D, SDD = sum D^2.
 * J^T J, J^T 1, J^T f1, sum f1 and |f1|^2 then follow in closed form, all in FP64
 * (sep_finish); |f0|^2 is the current point's |f|^2, known from the step that moved there.
 * Slots of the sums array: */
enum SepSum : int {
    kMx0 = 0,  /* M_0..M_4 over columns, then N_0..N_2 */
    kMy0 = 8,  /* the same over rows */
    kU1 = 16,
    kUu = 17,
    kUv = 18,
    kUd = 19,
    kSD = 20,
    kSDD = 21,
    kSepSums = 22
};

/* FP64 constants of a separable pass. */
struct SepIn {
    double x1, y1, c1; /* trial centre (window coordinates), c1 = 1/(2 r1^2) */
    double A1, b1, is; /* trial amplitude and offset, 1/sigma */
    double kr, kx;     /* J columns: d f/d r = kr A1 e1 d^2, d f/d x0 = kx A1 e1 u */
};

/* Trial point t0 + dx (dx = 0 for the start point).  Symmetric: (A, r, x0, y0, b),
 * d f/d r = A e d^2 / (sqrt(2) r^3 sigma), d f/d x0 = A e (x - x0) / (r^2 sigma)
 * (source/Localization.cpp:1151-1166); fixed width: (A, x0, y0, b) with r = w. */
LZR_HD LZR_INLINE void make_sep_in(const double (&t0)[5], const double (&dx)[5], double is, double, SepIn &q) {
    LZR_NO_CONTRACT
    const double ir1 = drcp(t0[1] + dx[1]);
    q.A1 = t0[0] + dx[0];
    q.x1 = t0[2] + dx[2];
    q.y1 = t0[3] + dx[3];
    q.b1 = t0[4] + dx[4];
    q.c1 = 0.5 * ir1 * ir1;
    q.is = is;
    q.kr = ir1 * ir1 * ir1 * is * (1.0 / kSqrt2);
    q.kx = ir1 * ir1 * is;
}
LZR_HD LZR_INLINE void make_sep_in(const double (&t0)[4], const double (&dx)[4], double is, double w, SepIn &q) {
    LZR_NO_CONTRACT
    const double iw = drcp(w);
    q.A1 = t0[0] + dx[0];
    q.x1 = t0[1] + dx[1];
    q.y1 = t0[2] + dx[2];
    q.b1 = t0[3] + dx[3];
    q.c1 = 0.5 * iw * iw;
    q.is = is;
    q.kr = 0.0;
    q.kx = iw * iw * is;
}

/* Adds one column (or row) with factor e at offset u to the moments m[0..7]. */
LZR_HD LZR_INLINE void axis_moments(double e, double u, double *m) {
    LZR_NO_CONTRACT
    const double e2 = e * e;
    const double e2u = e2 * u, e2u2 = e2u * u, e2u3 = e2u2 * u;
    const double eu = e * u;
    m[0] += e2;
    m[1] += e2u;
    m[2] += e2u2;
    m[3] += e2u3;
    m[4] = dfma(e2u3, u, m[4]);
    m[5] += e;
    m[6] += eu;
    m[7] = dfma(eu, u, m[7]);
}

/* Data sums of one pixel row j (ey = Ey(j), w = j - y1) from the row's inner sums
 * r1 = sum_i Ex D, ru = sum_i Ex u D, ru2 = sum_i Ex u^2 D. */
LZR_HD LZR_INLINE void row_sums(double ey, double w, double r1, double ru, double ru2, double rd, double rdd,
                                double (&v)[kSepSums]) {
    LZR_NO_CONTRACT
    v[kU1] = dfma(ey, r1, v[kU1]);
    v[kUu] = dfma(ey, ru, v[kUu]);
    v[kUv] = dfma(ey * w, r1, v[kUv]);
    v[kUd] = dfma(ey, dfma(w * w, r1, ru2), v[kUd]);
    v[kSD] += rd;
    v[kSDD] += rdd;
}

/* Gaussian factors E(t) = exp(-c (t - x)^2), t = 0, 1, ..., by the recurrence
 * E(t+1) = E(t) R(t), R(t+1) = R(t) Q with R(0) = exp(-c (1 - 2x)), Q = exp(-2c): two
 * exponentials per axis (Q is shared) instead of one per column, ~t ulps of rounding.
 * `ok` is false when an exponent at the ends of the window leaves +-690 (intermediate
 * overflow or underflow, e.g. a collapsed width); the caller then evaluates each factor
 * directly. */
struct GaussRec {
    double e, r, q;
    bool ok;
    LZR_HD LZR_INLINE void start(double c, double x, int n, double q_exp) {
        LZR_NO_CONTRACT
        const double xe = (double)(n - 1) - x;
        const double a0 = (c * x) * x, a1 = (c * xe) * xe;
        const double b0 = c * (1.0 - 2.0 * x), b1 = c * (2.0 * (double)n - 1.0 - 2.0 * x);
        ok = c >= 0.0 && 2.0 * c < 690.0 && a0 < 690.0 && a1 < 690.0 && dabs(b0) < 690.0 && dabs(b1) < 690.0;
        e = dexp(-a0);
        r = dexp(-b0);
        q = q_exp;
    }
    LZR_HD LZR_INLINE double next() {
        LZR_NO_CONTRACT
        const double v = e;
        e = e * r;
        r = r * q;
        return v;
    }
};

/* A lane's column buffer holds the column factors Ex(0..nmax-1) of its last separable pass,
 * then Ey(0) and Ey(1) at slots ey_slot, ey_slot + 1 (ey_slot = nmax, the launch's largest
 * window side): the scale update of an accepted step reads the first rows of J from them
 * (sep_diag_norms). */
constexpr int kXbEySlots = 2;
LZR_HD LZR_INLINE int xb_ey_slot(const LaneConfig &cfg) { return 2 * cfg.max_half_width + 1; }

/* Pixels per step of the separable pass's row loop (odd windows of 9 and 15 split evenly). */
constexpr int kSepChunk = 3;

/* Adds one pixel (value D, column factor ex at offset u) to a row's inner sums. */
LZR_HD LZR_INLINE void sep_row_pixel(double D, double ex, double u, double &r1, double &ru, double &ru2,
                                     double &rd, double &rdd) {
    LZR_NO_CONTRACT
    const double t = ex * D;
    const double tu = t * u;
    r1 += t;
    ru += tu;
    ru2 = dfma(tu, u, ru2);
    rd += D;
    rdd = dfma(D, D, rdd);
}

/* Inner sums of one pixel row of a separable pass: sum_i Ex D, sum_i Ex u D, sum_i Ex u^2 D,
 * sum_i D, sum_i D^2. */
struct SepRow {
    double r1, ru, ru2, rd, rdd;
};

/* kSepChunk pixels of row j from column i0: all their reads are issued before their
 * arithmetic, so one LDS latency covers the chunk (full chunks read at fixed offsets from the
 * chunk's first pixel).  Tail: reads past nloop are clamped in bounds and skipped (nloop is
 * uniform over the wave).  Guard: pixels beyond the lane's own n x n window are masked. */
template <bool Tail, bool Guard, typename Pix, typename XBuf>
LZR_HD LZR_INLINE void sep_chunk(const SepIn &q, int n, int nloop, bool row_in, const Pix &pix,
                                 const XBuf &xb, int i0, int j, SepRow &r) {
    LZR_NO_CONTRACT
    decltype(pix.raw(0, 0)) Dc[kSepChunk];
    double Xc[kSepChunk];
#pragma unroll
    for (int k = 0; k < kSepChunk; ++k) {
        const int i = Tail && i0 + k >= nloop ? nloop - 1 : i0 + k;
        Dc[k] = pix.raw(i, j);
        Xc[k] = xb.get(i);
    }
#pragma unroll
    for (int k = 0; k < kSepChunk; ++k) {
        const int i = i0 + k;
        if (Tail && i >= nloop) break;
        const double D = !Guard || (row_in && i < n) ? (double)Dc[k] : 0.0;
        sep_row_pixel(D, Xc[k], (double)i - q.x1, r.r1, r.ru, r.ru2, r.rd, r.rdd);
    }
}

/* Column factors Ex(i) (into xb) and their moments, the first step of a separable pass over
 * columns 0..nloop-1 (factors beyond n are 0); gy is set up for the rows. */
template <typename XBuf>
LZR_HD LZR_INLINE void sep_columns(const SepIn &q, int n, int nloop, XBuf &xb, double (&v)[kSepSums], GaussRec &gy) {
    LZR_NO_CONTRACT
#pragma unroll
    for (int k = 0; k < kSepSums; ++k) v[k] = 0.0;
    const double qexp = dexp(-2.0 * q.c1);
    GaussRec gx;
    gx.start(q.c1, q.x1, n, qexp);
    gy.start(q.c1, q.y1, n, qexp);
    /* the lane's branch is taken once, outside the column loop: the common recurrence loop
     * has no branch in its body (a lane that needs direct exponentials runs the second loop) */
    if (gx.ok) {
        for (int i = 0; i < nloop; ++i) {
            const double u = (double)i - q.x1;
            double e = gx.next();
            e = i < n ? e : 0.0;
            xb.set(i, e);
            axis_moments(e, u, v + kMx0);
        }
    } else {
        for (int i = 0; i < nloop; ++i) {
            const double u = (double)i - q.x1;
            double e = dexp(-(q.c1 * u) * u);
            e = i < n ? e : 0.0;
            xb.set(i, e);
            axis_moments(e, u, v + kMx0);
        }
    }
}

/* Row factor Ey(j) (the next step of the row recurrence; 0 beyond n), added to the row
 * moments.  Rows must be taken in order j = 0, 1, ... */
LZR_HD LZR_INLINE double sep_row_factor(const SepIn &q, GaussRec &gy, int j, int n, double (&v)[kSepSums]) {
    LZR_NO_CONTRACT
    const double w = (double)j - q.y1;
    double ey = gy.next();
    if (!gy.ok) ey = dexp(-(q.c1 * w) * w);
    ey = j < n ? ey : 0.0;
    axis_moments(ey, w, v + kMy0);
    return ey;
}

/* K consecutive row factors Ey(j0..j0+K-1) (0 at and beyond n), added to the row moments in
 * row order: the same operations as K calls of sep_row_factor, with the lane's branch between
 * the recurrence and direct exponentials taken once.  Rows at or beyond n add zeros, which
 * leaves every moment unchanged, so K may run past the window. */
template <int K>
LZR_HD LZR_INLINE void sep_row_factors(const SepIn &q, GaussRec &gy, int j0, int n, double (&v)[kSepSums],
                                       double (&ey)[K]) {
    LZR_NO_CONTRACT
    if (gy.ok) {
#pragma unroll
        for (int rr = 0; rr < K; ++rr) {
            const int j = j0 + rr;
            const double w = (double)j - q.y1;
            double e = gy.next();
            e = j < n ? e : 0.0;
            axis_moments(e, w, v + kMy0);
            ey[rr] = e;
        }
    } else {
#pragma unroll
        for (int rr = 0; rr < K; ++rr) {
            const int j = j0 + rr;
            const double w = (double)j - q.y1;
            double e = dexp(-(q.c1 * w) * w);
            e = j < n ? e : 0.0;
            axis_moments(e, w, v + kMy0);
            ey[rr] = e;
        }
    }
}

/* Inner sums of row j over columns 0..nloop-1 (Guard: pixels outside the lane's own window
 * masked; row_in false masks the whole row). */
template <bool Guard, typename Pix, typename XBuf>
LZR_HD LZR_INLINE SepRow sep_row(const SepIn &q, int n, int nloop, bool row_in, const Pix &pix, const XBuf &xb,
                                 int j) {
    SepRow r{0.0, 0.0, 0.0, 0.0, 0.0};
    int i0 = 0;
    for (; i0 + kSepChunk <= nloop; i0 += kSepChunk) sep_chunk<false, Guard>(q, n, nloop, row_in, pix, xb, i0, j, r);
    if (i0 < nloop) sep_chunk<true, Guard>(q, n, nloop, row_in, pix, xb, i0, j, r);
    return r;
}

/* The separable pass of one lane over its n x n window (nloop >= n: the loop bound shared
 * by the wave; with `guard`, pixels beyond the lane's own n are masked).  pix.raw(i, j)
 * returns pixel (i, j) as stored, xb holds one FP64 value per column.  The pieces (columns,
 * row factor, row inner sums, row fold) are also what a group of lanes sharing one fit runs:
 * the same operations in the same order, so a fit's sums do not depend on who computes them. */
template <typename Pix, typename XBuf>
LZR_HD LZR_INLINE void sep_pass(const SepIn &q, int n, int nloop, bool guard, const Pix &pix, XBuf &xb,
                                double (&v)[kSepSums], int ey_slot) {
    LZR_NO_CONTRACT
    GaussRec gy;
    sep_columns(q, n, nloop, xb, v, gy);
    for (int j = 0; j < nloop; ++j) {
        const double ey = sep_row_factor(q, gy, j, n, v);
        if (j < kXbEySlots) xb.set(ey_slot + j, ey);
        const bool row_in = !guard || j < n;
        const SepRow r = guard ? sep_row<true>(q, n, nloop, row_in, pix, xb, j)
                               : sep_row<false>(q, n, nloop, row_in, pix, xb, j);
        row_sums(ey, (double)j - q.y1, r.r1, r.ru, r.ru2, r.rd, r.rdd, v);
    }
}

/* Normal equations and residual sums of a pass from its moments and data sums (npix
 * pixels).  J columns in the reference's order without the constant offset column:
 * (A, r, x0, y0) symmetric, (A, x0, y0) fixed width; f1 = (A1 e1 + b1 - D)/sigma.
 * init: the pass is at the start point (|f0|^2 = |f1|^2); otherwise fn2 = |f0|^2. */
template <int P, typename Acc>
LZR_HD LZR_INLINE void sep_finish(const SepIn &q, int npix, const double (&v)[kSepSums], bool init, double fn2,
                                  PassAcc<P, Acc> &a) {
    LZR_NO_CONTRACT
    static_assert(P == 4 || P == 5, "separable models only");
    const double *mx = v + kMx0, *my = v + kMy0;
    const double S1 = mx[0] * my[0], Su = mx[1] * my[0], Sv = mx[0] * my[1];
    const double Suu = mx[2] * my[0], Suv = mx[1] * my[1], Svv = mx[0] * my[2];
    const double T1 = mx[5] * my[5], Tu = mx[6] * my[5], Tv = mx[5] * my[6];
    const double is = q.is, A1 = q.A1, b1 = q.b1, N = (double)npix;
    const double bx = q.kx * A1;
    constexpr int X = P == 5 ? 2 : 1; /* column of x0 */
    a.hq[tri(0, 0)] = Acc(is * is * S1);
    a.hq[tri(X, 0)] = Acc(is * bx * Su);
    a.hq[tri(X, X)] = Acc(bx * bx * Suu);
    a.hq[tri(X + 1, 0)] = Acc(is * bx * Sv);
    a.hq[tri(X + 1, X)] = Acc(bx * bx * Suv);
    a.hq[tri(X + 1, X + 1)] = Acc(bx * bx * Svv);
    a.sj[0] = Acc(is * T1);
    a.sj[X] = Acc(bx * Tu);
    a.sj[X + 1] = Acc(bx * Tv);
    a.gq[0] = Acc(is * is * ((A1 * S1 + b1 * T1) - v[kU1]));
    a.gq[X] = Acc(is * bx * ((A1 * Su + b1 * Tu) - v[kUu]));
    a.gq[X + 1] = Acc(is * bx * ((A1 * Sv + b1 * Tv) - v[kUv]));
    if constexpr (P == 5) {
        const double Sd = mx[2] * my[0] + mx[0] * my[2];
        const double Sdu = mx[3] * my[0] + mx[1] * my[2];
        const double Sdv = mx[2] * my[1] + mx[0] * my[3];
        const double Sdd = (mx[4] * my[0] + 2.0 * (mx[2] * my[2])) + mx[0] * my[4];
        const double Td = mx[7] * my[5] + mx[5] * my[7];
        const double ar = q.kr * A1;
        a.hq[tri(1, 0)] = Acc(is * ar * Sd);
        a.hq[tri(1, 1)] = Acc(ar * ar * Sdd);
        a.hq[tri(2, 1)] = Acc(ar * bx * Sdu);
        a.hq[tri(3, 1)] = Acc(ar * bx * Sdv);
        a.sj[1] = Acc(ar * Td);
        a.gq[1] = Acc(is * ar * ((A1 * Sd + b1 * Td) - v[kUd]));
    }
    a.sf = Acc(is * ((A1 * T1 + N * b1) - v[kSD]));
    /* |f1|^2 sigma^2 = A1 (A1 S1 + 2 b1 T1 - 2 U1) + b1 (N b1 - 2 SD) + SDD */
    const double sse1 =
        is * is * ((A1 * ((A1 * S1 + 2.0 * (b1 * T1)) - 2.0 * v[kU1]) + b1 * (N * b1 - 2.0 * v[kSD])) + v[kSDD]);
    a.sse0 = Acc(init ? sse1 : 0.0);
    a.dsse = Acc(init ? 0.0 : sse1 - fn2);
}

/* ---- moment pass of the ellipsoidal model -------------------------------------------
 * e1 = exp(-(qa u^2 + 2 qc u v + qb v^2)) (u = i - x1, v = j - y1) and every column of the
 * reference Jacobian (source/Localization.cpp:1275-1280) is e1 times a polynomial of degree
 * <= 2 in (u, v) over the monomials 1, u, v, u^2, uv, v^2.  J^T J, J^T 1 and J^T f1 are
 * therefore fixed combinations of the moments
 *   S_ab = sum e1^2 u^a v^b (a + b <= 4), T_ab = sum e1 u^a v^b, U_ab = sum e1 D u^a v^b
 * (a + b <= 2), plus SD = sum D and SDD = sum D^2.  With e1 = Ex(i) Ey(j) C(i, j),
 * Ex = exp(-qa u^2), Ey = exp(-qb v^2), the cross factor C = exp(-2 qc u v) follows the
 * recurrence C(i+1, j) = C(i, j) exp(-2 qc v) along a row, so a row needs 13 inner sums
 * over its pixels and the rows are combined with the powers of v (ell_pass).  When a
 * factor could overflow or underflow on its own (a collapsed width, |rho| near 1) the lane
 * evaluates e1 per pixel instead and weights its rows by 1. */
enum EllSum : int {
    kES = 0,     /* S_ab, a + b <= 4: 15 slots, index ell_s(a, b) */
    kET = 15,    /* T_ab, a + b <= 2: 6 slots, index ell_t(a, b) */
    kEU = 21,    /* U_ab: 6 slots */
    kESD = 27,
    kESDD = 28,
    kEBad = 29, /* pixels whose residual is +-inf (1 each) or NaN (2 each), see ell_chunk */
    kEllSums = 30
};
LZR_HD constexpr int ell_s(int a, int b) { return (a + b) * (a + b + 1) / 2 + b; }
LZR_HD constexpr int ell_t(int a, int b) { return (a + b) * (a + b + 1) / 2 + b; }

/* FP64 constants of an ellipsoidal moment pass at the trial point. */
struct EllIn {
    double x1, y1, qa, qb, qc; /* centre (window coordinates), exponent coefficients */
    double A1, b1, is, rho;
    double k1, k2, k3, k4, k5, k6, k7, k8, k9; /* reference column factors */
};

/* Trial point t0 + dx of (A, sx, sy, x0, y0, rho, b); columns with den = 2(1-rho^2) sigma,
 * k = 2/(den ...) = 1/(omr sigma ...) as in source/Localization.cpp:1275-1280. */
LZR_HD LZR_INLINE void make_ell_in(const double (&t0)[7], const double (&dx)[7], double is, EllIn &q) {
    LZR_NO_CONTRACT
    const double sx = t0[1] + dx[1], sy = t0[2] + dx[2], rho = t0[5] + dx[5];
    quad_coeffs(sx, sy, rho, q.qa, q.qc, q.qb);
    q.A1 = t0[0] + dx[0];
    q.x1 = t0[3] + dx[3];
    q.y1 = t0[4] + dx[4];
    q.b1 = t0[6] + dx[6];
    q.is = is;
    q.rho = rho;
    const double iomr = drcp(1.0 - rho * rho);
    const double isx = drcp(sx), isy = drcp(sy);
    const double t = iomr * is;
    q.k1 = t * isx * isx * isx;
    q.k2 = t * rho * isx * isx * isy;
    q.k3 = t * isy * isy * isy;
    q.k4 = t * rho * isx * isy * isy;
    q.k5 = t * isx * isx;
    q.k6 = t * rho * isx * isy;
    q.k7 = t * isy * isy;
    q.k8 = t * isx * isy;
    q.k9 = t * iomr * rho * isx * isy;
}

/* True when the three factors Ex, Ey, C stay within exp(+-600) over an n x n window. */
LZR_HD LZR_INLINE bool ell_factored(const EllIn &q, int n) {
    LZR_NO_CONTRACT
    const double um = dmax(dabs(q.x1), dabs((double)(n - 1) - q.x1));
    const double vm = dmax(dabs(q.y1), dabs((double)(n - 1) - q.y1));
    return q.qa >= 0.0 && q.qb >= 0.0 && q.qa * um * um <= 600.0 && q.qb * vm * vm <= 600.0 &&
           2.0 * dabs(q.qc) * um * vm <= 600.0;
}

/* Row-end accumulation: inner sums P_a = sum e'^2 u^a (a <= 4), R_a = sum e' u^a,
 * W_a = sum e' D u^a (a <= 2) of row v with row weight w (e1 = w e'). */
LZR_HD LZR_INLINE void ell_row(double w, double v, const double (&P)[5], const double (&R)[3], const double (&W)[3],
                               double rd, double rdd, double bad, double (&s)[kEllSums]) {
    LZR_NO_CONTRACT
    const double w2 = w * w;
    double vb[5];
    vb[0] = 1.0;
#pragma unroll
    for (int b = 1; b < 5; ++b) vb[b] = vb[b - 1] * v;
#pragma unroll
    for (int a = 0; a < 5; ++a)
#pragma unroll
        for (int b = 0; a + b < 5; ++b) s[kES + ell_s(a, b)] = dfma(w2 * vb[b], P[a], s[kES + ell_s(a, b)]);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; a + b < 3; ++b) {
            s[kET + ell_t(a, b)] = dfma(w * vb[b], R[a], s[kET + ell_t(a, b)]);
            s[kEU + ell_t(a, b)] = dfma(w * vb[b], W[a], s[kEU + ell_t(a, b)]);
        }
    s[kESD] += rd;
    s[kESDD] += rdd;
    s[kEBad] += bad;
}

/* Inner sums of one pixel with factor e (e1 / row weight), offset u and data D. */
LZR_HD LZR_INLINE void ell_pixel(double e, double u, double D, double (&P)[5], double (&R)[3], double (&W)[3],
                                 double &rd, double &rdd) {
    LZR_NO_CONTRACT
    const double e2 = e * e, e2u = e2 * u, e2u2 = e2u * u, e2u3 = e2u2 * u;
    P[0] += e2;
    P[1] += e2u;
    P[2] += e2u2;
    P[3] += e2u3;
    P[4] = dfma(e2u3, u, P[4]);
    const double eu = e * u;
    R[0] += e;
    R[1] += eu;
    R[2] = dfma(eu, u, R[2]);
    const double ed = e * D, edu = ed * u;
    W[0] += ed;
    W[1] += edu;
    W[2] = dfma(edu, u, W[2]);
    rd += D;
    rdd = dfma(D, D, rdd);
}

/* Inner sums of one pixel row of the ellipsoidal pass (factors e' = e1 / row weight):
 * P_a = sum e'^2 u^a, R_a = sum e' u^a, W_a = sum e' D u^a, sum D, sum D^2, and the count of
 * pixels whose residual is not finite. */
struct EllRow {
    double P[5], R[3], W[3], rd, rdd, bad;
};

/* kSepChunk pixels of row j (offset v) from column i0 for the ellipsoidal pass, reads first
 * (Tail and Guard as for sep_chunk); c is the row's cross factor, advanced by g per column.
 * Per-pixel exponentials (fac false) can overflow where |rho| >= 1: the reference's residual
 * vector then holds +-inf (A e overflows) or NaN entries, and GSL's scaled 2-norm is +inf for
 * one infinite entry and NaN for two or more (or any NaN), which the LM step treats
 * differently; `bad` counts such pixels (1 per infinite, 2 per NaN residual).  In the
 * factored path every factor stays within exp(+-600) and the count is 0. */
template <bool Tail, bool Guard, typename Pix, typename XBuf>
LZR_HD LZR_INLINE void ell_chunk(const EllIn &q, int n, int nloop, bool fac, bool row_in, double v, double g,
                                 double &c, const Pix &pix, const XBuf &xb, int i0, int j, EllRow &r) {
    LZR_NO_CONTRACT
    decltype(pix.raw(0, 0)) Dc[kSepChunk];
    double Xc[kSepChunk];
#pragma unroll
    for (int k = 0; k < kSepChunk; ++k) {
        const int i = Tail && i0 + k >= nloop ? nloop - 1 : i0 + k;
        Dc[k] = pix.raw(i, j);
        Xc[k] = xb.get(i);
    }
#pragma unroll
    for (int k = 0; k < kSepChunk; ++k) {
        const int i = i0 + k;
        if (Tail && i >= nloop) break;
        const bool in = !Guard || (row_in && i < n);
        const double D = in ? (double)Dc[k] : 0.0;
        const double u = (double)i - q.x1;
        /* outside the lane's window the factors may overflow (the range check covers the
         * window only): masked by a select, never by a product with 0 */
        double e;
        if (fac) {
            const double ef = Xc[k] * c;
            c = c * g;
            e = in ? ef : 0.0;
        } else {
            e = in ? dexp(-((q.qa * u) * u + (2.0 * q.qc * u) * v + (q.qb * v) * v)) : 0.0;
            const double t = q.A1 * e;
            r.bad += t - t != 0.0 ? (t == t ? 1.0 : 2.0) : 0.0; /* +-inf: inf - inf is NaN */
        }
        ell_pixel(e, u, D, r.P, r.R, r.W, r.rd, r.rdd);
    }
}

/* Row recurrences of the ellipsoidal pass: Ey(j) and, for the factored form, the cross
 * factor's row start C(0, j) = exp(2 qc x1 v_j) and step g_j = exp(-2 qc v_j). */
struct EllRows {
    GaussRec gy;
    double gj, cj, gr, cr;
    bool fac;
};

/* Column factors Ex(i) (into xb) of the ellipsoidal pass over columns 0..nloop-1 and the set-up
 * of the row recurrences. */
template <typename XBuf>
LZR_HD LZR_INLINE void ell_columns(const EllIn &q, int n, int nloop, XBuf &xb, double (&s)[kEllSums], EllRows &rw) {
    LZR_NO_CONTRACT
#pragma unroll
    for (int k = 0; k < kEllSums; ++k) s[k] = 0.0;
    rw.fac = ell_factored(q, n);
    GaussRec gx;
    gx.start(q.qa, q.x1, n, dexp(-2.0 * q.qa));
    rw.gy.start(q.qb, q.y1, n, dexp(-2.0 * q.qb));
    for (int i = 0; i < nloop; ++i) {
        const double u = (double)i - q.x1;
        double e = gx.next();
        if (!gx.ok) e = dexp(-(q.qa * u) * u);
        xb.set(i, i < n ? e : 0.0);
    }
    /* row quantities by recurrence over rows: g_j = C(i + 1, j) / C(i, j) = exp(-2 qc v_j),
     * c_j = C(0, j) = exp(2 qc x1 v_j); within the window both stay inside exp(+-600) */
    rw.gj = rw.cj = rw.gr = rw.cr = 1.0;
    if (rw.fac) {
        rw.gj = dexp(2.0 * q.qc * q.y1);
        rw.cj = dexp(-2.0 * q.qc * q.x1 * q.y1);
        rw.gr = dexp(-2.0 * q.qc);
        rw.cr = dexp(2.0 * q.qc * q.x1);
    }
}

/* The next row's weight w (= Ey(j) in the factored form, else 1), cross-factor step g and start
 * c.  Rows must be taken in order j = 0, 1, ... */
LZR_HD LZR_INLINE void ell_row_factor(const EllIn &q, EllRows &rw, int j, double &w, double &g, double &c) {
    LZR_NO_CONTRACT
    const double v = (double)j - q.y1;
    double ey = rw.gy.next();
    if (!rw.gy.ok) ey = dexp(-(q.qb * v) * v);
    w = 1.0;
    g = 1.0;
    c = 1.0;
    if (rw.fac) {
        w = ey;
        g = rw.gj;
        c = rw.cj;
        rw.gj = rw.gj * rw.gr;
        rw.cj = rw.cj * rw.cr;
    }
}

/* Inner sums of row j over columns 0..nloop-1 (Guard and row_in as for sep_row). */
template <bool Guard, typename Pix, typename XBuf>
LZR_HD LZR_INLINE EllRow ell_row_sums(const EllIn &q, int n, int nloop, bool fac, bool row_in, double g, double c,
                                      const Pix &pix, const XBuf &xb, int j) {
    EllRow r{{0.0, 0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, 0.0, 0.0, 0.0};
    const double v = (double)j - q.y1;
    int i0 = 0;
    for (; i0 + kSepChunk <= nloop; i0 += kSepChunk)
        ell_chunk<false, Guard>(q, n, nloop, fac, row_in, v, g, c, pix, xb, i0, j, r);
    if (i0 < nloop) ell_chunk<true, Guard>(q, n, nloop, fac, row_in, v, g, c, pix, xb, i0, j, r);
    return r;
}

/* The ellipsoidal moment pass of one lane over its n x n window (nloop >= n: the wave's
 * loop bound; with `guard`, pixels beyond the lane's own n are masked).  pix.raw(i, j)
 * returns pixel (i, j) as stored, xb holds one FP64 value per column.  Built from the same
 * pieces a group of lanes sharing one fit runs (see sep_pass). */
template <typename Pix, typename XBuf>
LZR_HD LZR_INLINE void ell_pass(const EllIn &q, int n, int nloop, bool guard, const Pix &pix, XBuf &xb,
                                double (&s)[kEllSums]) {
    LZR_NO_CONTRACT
    EllRows rw;
    ell_columns(q, n, nloop, xb, s, rw);
    for (int j = 0; j < nloop; ++j) {
        double w, g, c;
        ell_row_factor(q, rw, j, w, g, c);
        const bool row_in = j < n;
        const EllRow r = guard ? ell_row_sums<true>(q, n, nloop, rw.fac, row_in, g, c, pix, xb, j)
                               : ell_row_sums<false>(q, n, nloop, rw.fac, row_in, g, c, pix, xb, j);
        ell_row(row_in ? w : 0.0, (double)j - q.y1, r.P, r.R, r.W, r.rd, r.rdd, r.bad, s);
    }
}

/* Normal equations and residual sums from the moments (npix pixels); J columns
 * (A, sx, sy, x0, y0, rho) = e1 times the polynomials of C below.  init: the pass is at
 * the start point; otherwise fn2 = |f0|^2. */
template <typename Acc>
LZR_HD LZR_INLINE void ell_finish(const EllIn &q, int npix, const double (&s)[kEllSums], bool init, double fn2,
                                  PassAcc<7, Acc> &a) {
    LZR_NO_CONTRACT
    constexpr int ma[6] = {0, 1, 0, 2, 1, 0}, mb[6] = {0, 0, 1, 0, 1, 2}; /* 1, u, v, uu, uv, vv */
    const double is = q.is, A1 = q.A1, b1 = q.b1;
    double C[6][6]; /* [monomial][column] */
#pragma unroll
    for (int m = 0; m < 6; ++m)
#pragma unroll
        for (int k = 0; k < 6; ++k) C[m][k] = 0.0;
    C[0][0] = is;
    C[3][1] = A1 * q.k1;
    C[4][1] = -(A1 * q.k2);
    C[5][2] = A1 * q.k3;
    C[4][2] = -(A1 * q.k4);
    C[1][3] = A1 * q.k5;
    C[2][3] = -(A1 * q.k6);
    C[2][4] = A1 * q.k7;
    C[1][4] = -(A1 * q.k6);
    C[3][5] = -(A1 * q.k9);
    C[4][5] = A1 * (q.k8 + 2.0 * q.rho * q.k9);
    C[5][5] = -(A1 * q.k9);
    /* MC = M2 C with M2[m][m'] = S(a_m + a_m', b_m + b_m') */
    double MC[6][6];
#pragma unroll
    for (int m = 0; m < 6; ++m)
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            double t = 0.0;
#pragma unroll
            for (int mm = 0; mm < 6; ++mm) t = dfma(s[kES + ell_s(ma[m] + ma[mm], mb[m] + mb[mm])], C[mm][k], t);
            MC[m][k] = t;
        }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
#pragma unroll
        for (int l = 0; l <= k; ++l) {
            double t = 0.0;
#pragma unroll
            for (int m = 0; m < 6; ++m) t = dfma(C[m][k], MC[m][l], t);
            a.hq[tri(k, l)] = Acc(t);
        }
        double tj = 0.0, tg = 0.0;
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            const int it = ell_t(ma[m], mb[m]);
            tj = dfma(C[m][k], s[kET + it], tj);
            const double f = (A1 * s[kES + ell_s(ma[m], mb[m])] + b1 * s[kET + it]) - s[kEU + it];
            tg = dfma(C[m][k], f, tg);
        }
        a.sj[k] = Acc(tj);
        a.gq[k] = Acc(is * tg);
    }
    const double N = (double)npix, S00 = s[kES], T00 = s[kET], U00 = s[kEU], SD = s[kESD];
    a.sf = Acc(is * ((A1 * T00 + N * b1) - SD));
    const double sse1 =
        is * is * ((A1 * ((A1 * S00 + 2.0 * (b1 * T00)) - 2.0 * U00) + b1 * (N * b1 - 2.0 * SD)) + s[kESDD]);
    a.sse0 = Acc(init ? sse1 : 0.0);
    a.dsse = Acc(init ? 0.0 : sse1 - fn2);
    a.bad = s[kEBad] == 0.0 ? 0 : (s[kEBad] == 1.0 ? 1 : 2);
}

/* The reference Jacobian at t, one row (pixel (i, j) of the window, window coordinates) at
 * a time: source/Localization.cpp:1128-1181 (symmetric, width column / sqrt(2)), :1183-1225
 * (fixed width), :1227-1295 (ellipsoidal, the rho column's quirk included).  The reciprocals
 * are formed once (is = 1/sigma, w = the fixed width). */
template <int P>
struct JacRows {
    double t[P];
    double is, r0, r1, r2, r3;
    LZR_HD LZR_INLINE void init(const double (&th)[P], double is_, double w) {
        LZR_NO_CONTRACT
#pragma unroll
        for (int c = 0; c < P; ++c) t[c] = th[c];
        is = is_;
        r0 = r1 = r2 = r3 = 0.0;
        if constexpr (P == 4) {
            r0 = 1.0 / (w * w); /* 1/w^2 */
        } else if constexpr (P == 5) {
            r0 = 1.0 / t[1]; /* 1/r */
            r1 = r0 * r0;    /* 1/r^2 */
        } else {
            r0 = 1.0 / t[1];                /* 1/sx */
            r1 = 1.0 / t[2];                /* 1/sy */
            r2 = 1.0 / (1.0 - t[5] * t[5]); /* 1/(1 - rho^2) */
            r3 = 0.5 * r2 * is;             /* 1/den */
        }
    }
    LZR_HD LZR_INLINE void row(int i, int j, double (&o)[P]) const {
        LZR_NO_CONTRACT
        if constexpr (P == 4) {
            const double A = t[0];
            const double ux = i - t[1], uy = j - t[2];
            const double e = dexp(-(ux * ux + uy * uy) * (0.5 * r0));
            const double ae = A * e * r0 * is;
            o[0] = e * is;
            o[1] = ae * ux;
            o[2] = ae * uy;
            o[3] = is;
        } else if constexpr (P == 5) {
            const double A = t[0];
            const double ux = i - t[2], uy = j - t[3];
            const double d2 = ux * ux + uy * uy;
            const double e = dexp(-d2 * (0.5 * r1));
            const double ae = A * e * r1 * is;
            o[0] = e * is;
            o[1] = ae * d2 * r0 * (1.0 / kSqrt2);
            o[2] = ae * ux;
            o[3] = ae * uy;
            o[4] = is;
        } else {
            const double A = t[0], rho = t[5];
            const double ux = i - t[3], uy = j - t[4];
            const double vx = ux * r0, vy = uy * r1; /* ux/sx, uy/sy */
            const double q = vx * vx + vy * vy - 2.0 * rho * vx * vy;
            const double e = dexp(-q * (0.5 * r2));
            const double aed = A * e * r3;
            o[0] = e * is;
            o[1] = -aed * (2.0 * rho * vx * vy - 2.0 * vx * vx) * r0;
            o[2] = -aed * (2.0 * rho * vx * vy - 2.0 * vy * vy) * r1;
            o[3] = -aed * (2.0 * rho * vy - 2.0 * vx) * r0;
            o[4] = -aed * (2.0 * rho * vx - 2.0 * vy) * r1;
            /* the reference rho column divides the unscaled quadratic form by sx sy
             * (exact only for sx = sy, source/Localization.cpp:1280) */
            const double qr = (ux * ux + uy * uy) * r0 * r1 - 2.0 * rho * vx * vy;
            o[5] = A * e * is * r2 * (vx * vy - rho * qr * r2);
            o[6] = is;
        }
    }
};

/* Column norms of the first `rows` Jacobian rows at t (GSL 1.16 update_diag reads the
 * first p rows of J).  Pixel k = i + n j, window coordinates; w is the fixed width. */
template <int P>
LZR_HD LZR_INLINE void partial_column_norms(const double (&t)[P], int n, double is, double w, int rows,
                                            double (&cn)[P]) {
    LZR_NO_CONTRACT
#pragma unroll
    for (int c = 0; c < P; ++c) cn[c] = 0.0;
    JacRows<P> jr;
    jr.init(t, is, w);
    for (int k = 0, i = 0, j = 0; k < rows; ++k) {
        double row[P];
        jr.row(i, j, row);
#pragma unroll
        for (int c = 0; c < P; ++c) cn[c] = dfma(row[c], row[c], cn[c]);
        /* pixel k + 1 = i + n j, without a division */
        if (++i == n) {
            i = 0;
            ++j;
        }
    }
#pragma unroll
    for (int c = 0; c < P; ++c) cn[c] = dsqrt(cn[c] == 0.0 ? 1.0 : cn[c]);
}

/* The same column norms for the separable models (P = 5, 4) and rows = P, from the Gaussian
 * factors the pass at t left in the lane's column buffer: J row k (pixel i = k mod n,
 * j = k div n; n >= 3, so j <= 1) has e = Ex(i) Ey(j) and the pass's own column constants, so
 * no exponential is evaluated here.  The factors are those of the pass's recurrence, a few
 * ulps from a direct exponential. */
template <int P, typename XBuf>
LZR_HD LZR_INLINE void sep_diag_norms(const double (&t)[P], int n, double is, double w, const XBuf &xb, int ey_slot,
                                      double (&cn)[P]) {
    LZR_NO_CONTRACT
    static_assert(P == 4 || P == 5, "separable models only");
    constexpr int X = P == 5 ? 2 : 1; /* column of x0 */
    double z[P];
#pragma unroll
    for (int c = 0; c < P; ++c) z[c] = 0.0;
    SepIn q;
    make_sep_in(t, z, is, w, q);
    const double ey0 = xb.get(ey_slot), ey1 = xb.get(ey_slot + 1);
    double ss[P];
#pragma unroll
    for (int c = 0; c < P; ++c) ss[c] = 0.0;
#pragma unroll
    for (int k = 0; k < P; ++k) {
        const int j = k >= n ? 1 : 0, i = k - j * n;
        const double e = xb.get(i) * (j == 0 ? ey0 : ey1);
        const double u = (double)i - q.x1, v = (double)j - q.y1;
        const double ae = q.A1 * e;
        double row[P];
        row[0] = e * is;
        if constexpr (P == 5) row[1] = q.kr * ae * dfma(u, u, v * v);
        row[X] = q.kx * ae * u;
        row[X + 1] = q.kx * ae * v;
        row[P - 1] = is;
#pragma unroll
        for (int c = 0; c < P; ++c) ss[c] = dfma(row[c], row[c], ss[c]);
    }
#pragma unroll
    for (int c = 0; c < P; ++c) cn[c] = dsqrt(ss[c] == 0.0 ? 1.0 : ss[c]);
}

/* Full normal equations from the pass accumulators. */
template <int P, typename Acc>
LZR_HD LZR_INLINE void assemble(const PassAcc<P, Acc> &a, int npix, double is, double (&H)[NP<P>::S], double (&g)[P]) {
    LZR_NO_CONTRACT
#pragma unroll
    for (int i = 0; i < P - 1; ++i) {
#pragma unroll
        for (int j = 0; j <= i; ++j) H[tri(i, j)] = (double)a.hq[tri(i, j)];
        H[tri(P - 1, i)] = (double)a.sj[i] * is;
        g[i] = (double)a.gq[i];
    }
    H[tri(P - 1, P - 1)] = (double)npix * is * is;
    g[P - 1] = (double)a.sf * is;
}

template <int P>
LZR_HD LZR_INLINE void to_double(const double (&Hf)[NP<P>::S], const double (&gf)[P], double (&H)[NP<P>::S], double (&g)[P]) {
    LZR_NO_CONTRACT
#pragma unroll
    for (int i = 0; i < NP<P>::S; ++i) H[i] = (double)Hf[i];
#pragma unroll
    for (int i = 0; i < P; ++i) g[i] = (double)gf[i];
}

/* Zeroes an output record field by field (the record may live in device memory). */
LZR_HD LZR_INLINE void clear_record(LzrFit &o) {
#pragma unroll
    for (int i = 0; i < LZR_MAX_PARAMS; ++i) {
        o.params[i] = 0.0;
        o.deviations[i] = 0.0;
    }
    o.integral = o.integral_deviation = 0.0;
    o.sd1 = o.sd2 = o.theta = o.sd1_deviation = o.sd2_deviation = 0.0;
    o.chi = 0.0;
    o.status = 0;
    o.lm_status = 0;
    o.iterations = o.trials = o.accepted_steps = o.model = 0;
}

/* ---- the state machine -------------------------------------------------------- */

/* Prepares a lane for a new candidate.  Returns false when the candidate is rejected
 * before fitting (window outside the frame, invalid request); `out` then holds the
 * final record. */
/* st (may be null): an LM start point replacing the reference start when st->width > 0. */
template <int P>
LZR_HD LZR_INLINE bool lane_start(LaneState<P> &s, const LaneConfig &cfg, const LzrCandidate &c, int roi, int nx,
                                  int ny, int nframes, int model, LzrFit &out, const LzrStart *st = nullptr) {
    LZR_NO_CONTRACT
    s.roi = roi;
    const int h = c.half_width > 0 ? c.half_width : cfg.default_half_width;
    if (h < 1 || h > cfg.max_half_width || c.frame < 0 || c.frame >= nframes) {
        clear_record(out);
        out.model = (int16_t)model;
        out.status = LZR_REJECT_INVALID;
        return false;
    }
    const int n = 2 * h + 1;
    /* x_offset = floor(x - cutoff_radius) (source/Localization.cpp:112-124) */
    /* floor(fx) + n - 1 <= nx - 1  <=>  fx < nx - n + 1 for fx >= 0, compared in FP64 so that no
     * candidate coordinate can overflow an int (NaN coordinates fail every test) */
    const double fx = c.x - (double)h, fy = c.y - (double)h;
    const bool inside = (fx >= 0.0) && (fy >= 0.0) && (fx < (double)(nx - n + 1)) && (fy < (double)(ny - n + 1));
    if (!inside) {
        clear_record(out);
        out.model = (int16_t)model;
        out.status = LZR_REJECT_EDGE;
        return false;
    }
    const int xoff = (int)fx, yoff = (int)fy; /* floor of a non-negative value */
    s.xoff = xoff;
    s.yoff = yoff;
    s.n = n;
    s.frame = c.frame;
    const double w = cfg.initial_psf_width;
    if constexpr (P == 4) {
        s.th[0] = c.intensity;
        s.th[1] = c.x - xoff;
        s.th[2] = c.y - yoff;
        s.th[3] = c.background;
    } else if constexpr (P == 5) {
        s.th[0] = c.intensity;
        s.th[1] = w;
        s.th[2] = c.x - xoff;
        s.th[3] = c.y - yoff;
        s.th[4] = c.background;
    } else {
        s.th[0] = c.intensity;
        s.th[1] = w;
        s.th[2] = w;
        s.th[3] = c.x - xoff;
        s.th[4] = c.y - yoff;
        s.th[5] = 0.0;
        s.th[6] = c.background;
    }
    if (st && st->width > 0.0) {
        const double sx = st->x - xoff, sy = st->y - yoff;
        if constexpr (P == 4) {
            s.th[0] = st->amplitude;
            s.th[1] = sx;
            s.th[2] = sy;
            s.th[3] = st->background;
        } else if constexpr (P == 5) {
            s.th[0] = st->amplitude;
            s.th[1] = st->width;
            s.th[2] = sx;
            s.th[3] = sy;
            s.th[4] = st->background;
        } else {
            s.th[0] = st->amplitude;
            s.th[1] = st->width;
            s.th[2] = st->width;
            s.th[3] = sx;
            s.th[4] = sy;
            s.th[5] = 0.0;
            s.th[6] = st->background;
        }
    }
#pragma unroll
    for (int i = 0; i < P; ++i) s.dx[i] = 0.0;
    s.phase = kInit;
    s.iterations = 0;
    s.trials = 0;
    s.accepted = 0;
    return true;
}

/* FP32 pass constants for the lane's next evaluation (trial point = th + dx for a
 * trial, th itself for the initial evaluation). */
/* Constants of a separable pass at the lane's trial point (the start point at kInit). */
template <int P>
LZR_HD LZR_INLINE void lane_sep_in(const LaneState<P> &s, const LaneConfig &cfg, SepIn &q) {
    LZR_NO_CONTRACT
    double step[P];
#pragma unroll
    for (int i = 0; i < P; ++i) step[i] = s.phase == kTrial ? s.dx[i] : 0.0;
    make_sep_in(s.th, step, cfg.inv_sigma, cfg.initial_psf_width, q);
}

/* Constants of an ellipsoidal moment pass at the lane's trial point. */
LZR_HD LZR_INLINE void lane_ell_in(const LaneState<7> &s, const LaneConfig &cfg, EllIn &q) {
    LZR_NO_CONTRACT
    double step[7];
#pragma unroll
    for (int i = 0; i < 7; ++i) step[i] = s.phase == kTrial ? s.dx[i] : 0.0;
    make_ell_in(s.th, step, cfg.inv_sigma, q);
}

/* One data sum term of pixel (u, w) with e = e1 and data D (the group-split pass). */
LZR_HD LZR_INLINE void sep_pixel(double e, double u, double w, double D, double (&v)[kSepSums]) {
    LZR_NO_CONTRACT
    const double t = e * D;
    v[kU1] += t;
    v[kUu] = dfma(t, u, v[kUu]);
    v[kUv] = dfma(t, w, v[kUv]);
    v[kUd] = dfma(t, dfma(u, u, w * w), v[kUd]);
    v[kSD] += D;
    v[kSDD] = dfma(D, D, v[kSDD]);
}

template <int P>
LZR_HD LZR_INLINE int test_delta(const LaneState<P> &s, const LaneConfig &cfg) {
    LZR_NO_CONTRACT
    double a[P];
    absolute<P>(s, s.th, a);
#pragma unroll
    for (int i = 0; i < P; ++i) {
        const double tol = cfg.epsabs + cfg.epsrel * dabs(a[i]);
        if (!(dabs(s.dx[i]) < tol)) return LZR_LM_CONTINUE;
    }
    return LZR_LM_SUCCESS;
}

/* Proposes the next trial (one lmpar + the first-iteration radius clamp). */
template <int P>
LZR_HD LZR_INLINE void propose(LaneState<P> &s) {
    LZR_NO_CONTRACT
    double H[NP<P>::S], g[P], x[P];
    to_double<P>(s.H, s.g, H, g);
    s.inner++;
    double pnorm; /* |D dx| = |D x|, bit for bit */
    lmpar<P>(H, g, s.diag, s.delta, s.par, x, pnorm);
#pragma unroll
    for (int i = 0; i < P; ++i) s.dx[i] = -x[i];
    s.pnorm = pnorm;
    if (s.iter_gsl == 1 && pnorm < s.delta) s.delta = pnorm;
    s.phase = kTrial;
}

/* The scaled gradient norm max_i |g_i / d_i| / |f| that iterate() computes at its
 * start; g, d and |f| change only when a step is accepted, so evaluating it where the
 * ETOLG test reads it gives the same value. */
template <int P>
LZR_HD LZR_INLINE double gradient_norm(const LaneState<P> &s) {
    LZR_NO_CONTRACT
    double gmax = 0.0;
#pragma unroll
    for (int i = 0; i < P; ++i) gmax = dmax(gmax, dabs((double)s.g[i] / s.diag[i]));
    return gmax / s.fnorm;
}

/* Start of one gsl_multifit_fdfsolver_iterate call of the reference loop
 * (source/Localization.cpp:148-155).  Returns true when a trial is to be proposed;
 * false when the loop ended (status in *st). */
template <int P>
LZR_HD LZR_INLINE bool begin_iterate(LaneState<P> &s, const LaneConfig &cfg, int *st) {
    LZR_NO_CONTRACT
    for (;;) {
        s.iterations++;
        if (s.fnorm != 0.0) break;
        /* |f| == 0: iterate() returns success with dx unchanged */
        const int td = test_delta<P>(s, cfg);
        if (td != LZR_LM_CONTINUE || s.iterations >= cfg.max_iterations) {
            *st = td;
            return false;
        }
    }
    s.inner = 0;
    return true;
}

/* Stores the raw end state of the lane's fit (the reference's x, |f| and loop counters) for
 * the epilogue. */
template <int P>
LZR_HD LZR_INLINE void store_final(LaneState<P> &s, int lm_status, int model, LzrRawFit *o) {
    if (o) { /* null: another lane holding the same fit writes the record */
#pragma unroll
        for (int i = 0; i < P; ++i) o->th[i] = s.th[i];
        o->chi = s.fnorm;
        o->status = kPending;
        o->lm_status = lm_status;
        o->iterations = (int16_t)s.iterations;
        o->trials = (int16_t)s.trials;
        o->accepted_steps = (int16_t)s.accepted;
        o->model = (int16_t)model;
    }
    s.phase = kIdle;
}

/* The bookkeeping of one pass; true when a new trial is to be proposed, otherwise false
 * with the fit's final status in *st_out. */
template <int P, typename Acc, typename XBuf>
LZR_HD LZR_INLINE bool lane_step(LaneState<P> &s, const LaneConfig &cfg, const PassAcc<P, Acc> &a, const XBuf &xb,
                                 int *st_out) {
    LZR_NO_CONTRACT
    const int npix = s.n * s.n;
    if (s.phase == kInit) {
        /* lmsder set(): f, J at the start point */
        assemble<P, Acc>(a, npix, cfg.inv_sigma, s.H, s.g);
        s.fnorm = dsqrt((double)a.sse0);
#pragma unroll
        for (int i = 0; i < P; ++i) {
            const double hii = (double)s.H[tri(i, i)];
            s.diag[i] = dsqrt(hii == 0.0 ? 1.0 : hii);
        }
        double absx[P];
        absolute<P>(s, s.th, absx);
        s.xnorm = scaled_norm<P>(s.diag, absx);
        s.delta = s.xnorm > 0.0 ? 100.0 * s.xnorm : 100.0;
        s.par = 0.0;
        s.iter_gsl = 1;
        int st = 0;
        if (begin_iterate<P>(s, cfg, &st)) return true;
        *st_out = st;
        return false;
    }
    /* trial evaluated at th + dx */
    s.trials++;
    /* a zero width makes the reference's model function fail (source/Localization.cpp:1029-1031,
     * :1098-1100): iterate() returns that status and the driver leaves its loop */
    bool zero_width = false;
    if constexpr (P == 5) zero_width = s.th[1] + s.dx[1] == 0.0;
    if constexpr (P == 7) zero_width = s.th[1] + s.dx[1] == 0.0 || s.th[2] + s.dx[2] == 0.0;
    if (zero_width) {
        *st_out = LZR_LM_FAILURE;
        return false;
    }
    const double fn2 = s.fnorm * s.fnorm;
    /* |f1| as GSL's scaled 2-norm sees the trial residual vector: NaN when it holds a NaN or
     * two infinite entries, +inf for exactly one (a step to |rho| >= 1 overflows the model).
     * Sums that overflow while every entry is finite stand for an astronomically large but
     * finite norm, which the tests below treat like +inf. */
    const double dsse = (double)a.dsse;
    const bool finite = dsse == dsse && dabs(dsse) <= 1.0e300;
    double fnorm1 = finite ? dsqrt(dmax(fn2 + dsse, 0.0)) : 1.0e300;
    if (a.bad == 1) fnorm1 = __builtin_inf();
    if (a.bad == 2) fnorm1 = __builtin_nan("");
    const double inv_fn2 = drcp(fn2);
    double actred = -1.0;
    if (0.1 * fnorm1 < s.fnorm) actred = -dsse * inv_fn2; /* = 1 - (fnorm1/fnorm)^2 */
    double H[NP<P>::S];
    double gd[P];
    to_double<P>(s.H, s.g, H, gd);
    /* |J dx|^2 = dx^T H dx */
    double jdx2 = 0.0;
#pragma unroll
    for (int i = 0; i < P; ++i) {
        double row = 0.0;
#pragma unroll
        for (int j = 0; j < P; ++j) row = dfma(H[tri(i, j)], s.dx[j], row);
        jdx2 = dfma(s.dx[i], row, jdx2);
    }
    const double pnorm = s.pnorm; /* D is unchanged since the step was proposed */
    const double t1sq = dmax(jdx2, 0.0) * inv_fn2;
    const double t2sq = s.par * pnorm * pnorm * inv_fn2; /* (sqrt(par) |D dx| / |f|)^2 */
    const double prered = t1sq + t2sq / 0.5;
    const double dirder = -(t1sq + t2sq);
    const double ratio = prered > 0.0 ? ddiv(actred, prered) : 0.0;
    if (ratio > 0.25) {
        if (s.par == 0.0 || ratio >= 0.75) {
            s.delta = pnorm / 0.5;
            s.par *= 0.5;
        }
    } else {
        double temp = actred >= 0.0 ? 0.5 : 0.5 * dirder / (dirder + 0.5 * actred);
        if (0.1 * fnorm1 >= s.fnorm || temp < 0.1) temp = 0.1;
        s.delta = temp * dmin(s.delta, pnorm / 0.1);
        s.par /= temp;
    }
    int st = LZR_LM_CONTINUE;
    if (ratio >= 1e-4) {
        /* accepted: x, f, J move to the trial point */
#pragma unroll
        for (int i = 0; i < P; ++i) s.th[i] += s.dx[i];
        s.accepted++;
        assemble<P, Acc>(a, npix, cfg.inv_sigma, s.H, s.g);
        double absx[P];
        absolute<P>(s, s.th, absx);
        s.xnorm = scaled_norm<P>(s.diag, absx);
        s.fnorm = fnorm1;
        s.iter_gsl++;
        double cn[P];
        const int rows = cfg.diag_update_rows < 0 ? P : (cfg.diag_update_rows == 0 ? npix : cfg.diag_update_rows);
        if (rows >= npix) {
#pragma unroll
            for (int i = 0; i < P; ++i) {
                const double hii = (double)s.H[tri(i, i)];
                cn[i] = dsqrt(hii == 0.0 ? 1.0 : hii);
            }
        } else {
            bool done = false;
            if constexpr (P != 7) {
                if (rows == P) {
                    sep_diag_norms<P>(s.th, s.n, cfg.inv_sigma, cfg.initial_psf_width, xb, xb_ey_slot(cfg), cn);
                    done = true;
                }
            }
            if (!done) partial_column_norms<P>(s.th, s.n, cfg.inv_sigma, cfg.initial_psf_width, rows, cn);
        }
#pragma unroll
        for (int i = 0; i < P; ++i)
            if (cn[i] > s.diag[i]) s.diag[i] = cn[i];
        /* iterate() returned GSL_SUCCESS: the reference tests the step */
        st = test_delta<P>(s, cfg);
        if (st == LZR_LM_CONTINUE && s.iterations < cfg.max_iterations) {
            if (begin_iterate<P>(s, cfg, &st)) return true;
        }
        *st_out = st;
        return false;
    }
    if (dabs(actred) <= kDblEpsilon && prered <= kDblEpsilon && 0.5 * ratio <= 1.0) {
        st = LZR_LM_ETOLF;
    } else if (s.delta <= kDblEpsilon * s.xnorm) {
        st = LZR_LM_ETOLX;
    } else if (gradient_norm<P>(s) <= kDblEpsilon) {
        st = LZR_LM_ETOLG;
    } else if (s.inner >= 10) {
        st = LZR_LM_ENOPROG;
    } else {
        return true;
    }
    *st_out = st;
    return false;
}

/* Consumes one pass.  Returns true if the lane needs another pass (a new trial), false
 * when the fit is complete (*out written unless out is null).  Every path that continues ends in the one
 * call of propose (lmpar) at the bottom, so the lanes of a wave run lmpar together
 * whichever branch each of them took. */
template <int P, typename Acc, typename XBuf>
LZR_HD LZR_INLINE bool lane_after_pass(LaneState<P> &s, const LaneConfig &cfg, const PassAcc<P, Acc> &a,
                                       const XBuf &xb, int model, LzrRawFit *out) {
    int st = LZR_LM_CONTINUE;
    if (lane_step<P, Acc>(s, cfg, a, xb, &st)) {
        propose<P>(s);
        return true;
    }
    store_final<P>(s, st, model, out);
    return false;
}

/* 2x2 eigen-decomposition of the ellipsoidal covariance (source/Localization.cpp:525-547):
 * ascending eigenvalues, theta = atan(v0/v1) of the first eigenvector (0 if v1 == 0). */
LZR_HD LZR_INLINE bool eigen2(double sx, double sy, double rho, double &sd1, double &sd2, double &theta) {
    LZR_NO_CONTRACT
    const double a = sx * sx, c = sy * sy, b = rho * sx * sy;
    const double lim = 1.79769313486231570815e+308;
    if (!(dabs(a) <= lim) || !(dabs(b) <= lim) || !(dabs(c) <= lim)) return false;
    const double hd = 0.5 * (a - c), mean = 0.5 * (a + c);
    const double rad = dsqrt(hd * hd + b * b);
    sd1 = dsqrt(mean - rad);
    sd2 = dsqrt(mean + rad);
    if (b == 0.0) {
        theta = 0.0;
    } else {
        const double ratio = hd >= 0.0 ? -b / (hd + rad) : (hd - rad) / b;
#if defined(__HIP_DEVICE_COMPILE__)
        theta = atan(ratio);
#else
        theta = std::atan(ratio);
#endif
    }
    return true;
}

LZR_HD LZR_INLINE double dcos(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return cos(v);
#else
    return std::cos(v);
#endif
}
LZR_HD LZR_INLINE double dsin(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return sin(v);
#else
    return std::sin(v);
#endif
}

/* J^T J at the window-coordinate point t of an n x n window from the moment passes the lane
 * runs (no pixel data is involved: a window of zeros).  The separable models need only the
 * 1-D moments of the column and row factors (O(n)). */
template <int P>
LZR_HD LZR_INLINE void moment_normal_matrix(const double (&t)[P], int n, double is, double w, double (&H)[NP<P>::S]) {
    LZR_NO_CONTRACT
    double zero[P], g[P];
#pragma unroll
    for (int i = 0; i < P; ++i) zero[i] = 0.0;
    PassAcc<P, double> a;
    a.zero();
    if constexpr (P == 7) {
        struct {
            LZR_HD unsigned raw(int, int) const { return 0u; }
        } pix;
        struct {
            double v[2 * LZR_MAX_HALF_WIDTH + 1];
            LZR_HD void set(int i, double x) { v[i] = x; }
            LZR_HD double get(int i) const { return v[i]; }
        } xb;
        EllIn q;
        make_ell_in(t, zero, is, q);
        double v[kEllSums];
        ell_pass(q, n, n, false, pix, xb, v);
        ell_finish<double>(q, n * n, v, true, 0.0, a);
    } else {
        struct {
            LZR_HD void set(int, double) {}
        } xb;
        SepIn q;
        make_sep_in(t, zero, is, w, q);
        double v[kSepSums];
        GaussRec gy;
        sep_columns(q, n, n, xb, v, gy);
        for (int j = 0; j < n; ++j) sep_row_factor(q, gy, j, n, v);
        sep_finish<P, double>(q, n * n, v, true, 0.0, a);
    }
    assemble<P, double>(a, n * n, is, H, g);
}

/* Diagonal of (J^T J)^-1 by Cholesky in FP64, and the estimate max_k H_kk C_kk of the scaled
 * condition number (within a factor P^2 of cond(D H D), D = diag(H)^-1/2; the relative error
 * of C_kk is about that times the unit roundoff).  Returns false if the factorisation breaks
 * down. */
template <int P>
LZR_HD LZR_INLINE bool covariance_cholesky(const double (&H)[NP<P>::S], double (&cdiag)[P], double &kappa) {
    LZR_NO_CONTRACT
    double L[NP<P>::S], id[P];
    kappa = __builtin_inf();
    if (!cholesky<P>(H, L, id)) return false;
    kappa = 0.0;
    /* C = L^-T L^-1: C_kk is the squared norm of column k of L^-1 */
#pragma unroll
    for (int k = 0; k < P; ++k) {
        double e[P], y[P];
#pragma unroll
        for (int i = 0; i < P; ++i) e[i] = i == k ? 1.0 : 0.0;
        forward<P>(L, id, e, y);
        double s2 = 0.0;
#pragma unroll
        for (int i = k; i < P; ++i) s2 = dfma(y[i], y[i], s2);
        cdiag[k] = s2;
        kappa = dmax(kappa, H[tri(k, k)] * s2);
    }
    return true;
}

/* Diagonal of the covariance gsl_multifit_covar(J, 0.0, C) (source/Localization.cpp:185, :556)
 * of the reference Jacobian at the window-coordinate point t of an n x n window: J is reduced
 * row by row (pixel k = i + n j) to its triangular QR factor R by Givens rotations, which is
 * backward stable like GSL's QRPT, so the covariance keeps GSL's accuracy of about
 * cond(J) ulps instead of the cond(J)^2 of inverting J^T J.  As GSL with epsrel = 0, a
 * column whose pivot is exactly 0 (an exactly-zero Jacobian column, e.g. an amplitude
 * factor that underflows everywhere) is dropped: its covariance entries are 0 and the others
 * come from the remaining columns. */
template <int P>
LZR_HD LZR_INLINE void covariance_qr(const double (&t)[P], int n, double is, double w, double (&cdiag)[P]) {
    LZR_NO_CONTRACT
    constexpr int S = NP<P>::S;
    JacRows<P> jr;
    jr.init(t, is, w);
    double R[S]; /* upper triangle, R_km at tri(k, m) */
#pragma unroll
    for (int k = 0; k < S; ++k) R[k] = 0.0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double row[P];
            jr.row(i, j, row);
#pragma unroll
            for (int k = 0; k < P; ++k) {
                const double b = row[k];
                if (b == 0.0) continue;
                const double a = R[tri(k, k)];
                /* r = hypot(a, b) scaled by the larger magnitude: entries far out in the window
                 * are tiny (exponentials near underflow), their squares would vanish */
                const double aa = dabs(a), ab = dabs(b);
                const double mx = dmax(aa, ab), q = dmin(aa, ab) * drcp(mx);
                const double r = mx * dsqrt(dfma(q, q, 1.0));
                const double ir = drcp(r);
                const double c = a * ir, sn = b * ir;
                R[tri(k, k)] = r;
#pragma unroll
                for (int m = k + 1; m < P; ++m) {
                    const double rm = R[tri(k, m)];
                    R[tri(k, m)] = dfma(c, rm, sn * row[m]);
                    row[m] = dfma(c, row[m], -(sn * rm));
                }
            }
        }
    /* C = R^-1 R^-T over the columns with a non-zero pivot: C_kk = sum_m (R^-1)_km^2 */
    bool on[P];
    double id[P];
#pragma unroll
    for (int k = 0; k < P; ++k) {
        on[k] = R[tri(k, k)] != 0.0;
        id[k] = on[k] ? 1.0 / R[tri(k, k)] : 0.0;
        cdiag[k] = 0.0;
    }
#pragma unroll
    for (int m = 0; m < P; ++m) {
        if (!on[m]) continue;
        double x[P]; /* column m of R^-1 (rows > m are 0) */
#pragma unroll
        for (int k = P - 1; k >= 0; --k) {
            double v = k == m ? 1.0 : 0.0;
            if (k < m && on[k]) {
#pragma unroll
                for (int q = k + 1; q <= m; ++q)
                    if (on[q]) v = dfma(-R[tri(k, q)], x[q], v);
            }
            x[k] = k <= m && on[k] ? v * id[k] : 0.0;
        }
#pragma unroll
        for (int k = 0; k <= m; ++k) cdiag[k] = dfma(x[k], x[k], cdiag[k]);
    }
    /* a numerically singular factor (a pivot near underflow) leaves inf * 0 in the inverse:
     * the variance is unbounded, as GSL reports it there */
#pragma unroll
    for (int k = 0; k < P; ++k)
        if (cdiag[k] != cdiag[k]) cdiag[k] = __builtin_inf();
}

/* The covariance diagonal of gsl_multifit_covar(J, 0.0, C) at t: by Cholesky of J^T J (from the
 * moment pass) where the condition estimate keeps its error below ~1e-9 relative, otherwise
 * by the QR factor of J as GSL computes it. */
constexpr double kCovKappaMax = 1.0e6;
template <int P>
LZR_HD LZR_INLINE bool covariance(const double (&t)[P], int n, double is, double w, double (&cdiag)[P],
                                  bool defer_qr = false) {
    double H[NP<P>::S], kappa;
    moment_normal_matrix<P>(t, n, is, w, H);
    if (covariance_cholesky<P>(H, cdiag, kappa) && kappa <= kCovKappaMax) return true;
    if (defer_qr) return false; /* the caller finishes the record with the QR factor later */
    covariance_qr<P>(t, n, is, w, cdiag);
    return true;
}

/* The fields of an accepted record that follow from the covariance diagonal cd: deviations,
 * integral and its deviation, widths and their deviations, the symmetric-PSF filter
 * (source/Localization.cpp:185-208 symmetric, :371-372 fixed width, :555-596 ellipsoidal).
 * x holds the absolute parameters (out.params); sd1, sd2, theta (ellipsoidal) come from
 * eigen2 of x. */
template <int P>
LZR_HD LZR_INLINE void finish_accepted(LzrFit &out, const double (&x)[P], double chi, int h, double w, int model,
                                       const double (&cd)[P], double sd1, double sd2, double theta) {
    LZR_NO_CONTRACT
    if constexpr (P == 4) {
        const double dof = (double)((2 * h - 1) * (2 * h - 1) - 4);
        const double q = chi / dsqrt(dof);
        const double cs = 1.0 > q ? 1.0 : q;
#pragma unroll
        for (int i = 0; i < P; ++i) out.deviations[i] = cs * dsqrt(cd[i]);
        out.integral = 2 * kPi * w * w * x[0];
        /* the reference scales the amplitude's relative error by 2 pi I w^2 (:371-372) */
        const double relA = cs * dsqrt(cd[0]) / x[0];
        out.integral_deviation = 2 * kPi * out.integral * w * w * dsqrt(relA * relA);
        out.sd1 = out.sd2 = w;
        out.status = LZR_ACCEPTED;
    } else if constexpr (P == 5) {
        const double dof = (double)((2 * h - 1) * (2 * h - 1) - 5);
        const double q = chi / dsqrt(dof);
        const double cs = 1.0 > q ? 1.0 : q;
#pragma unroll
        for (int i = 0; i < P; ++i) out.deviations[i] = cs * dsqrt(cd[i]);
        const double width = x[1];
        out.integral = 2 * kPi * width * width * x[0];
        const double relA = cs * dsqrt(cd[0]) / x[0];
        const double relW = out.deviations[1] / width;
        out.integral_deviation = 2 * kPi * out.integral * dsqrt(relA * relA + 2 * relW * relW);
        out.sd1 = out.sd2 = width;
        out.sd1_deviation = out.sd2_deviation = out.deviations[1];
        out.theta = 0.0;
        out.status = LZR_ACCEPTED;
    } else {
        const double dof = (double)((2 * h - 1) * (2 * h - 1) - 7);
        const double q = chi / dsqrt(dof);
        const double cs = 1.0 > q ? 1.0 : q;
#pragma unroll
        for (int i = 0; i < P; ++i) out.deviations[i] = cs * dsqrt(cd[i]);
        const double rho = x[5];
        const double corr_dev = out.deviations[5];
        const double sxd = out.deviations[1], syd = out.deviations[2];
        out.sd1 = sd1;
        out.sd2 = sd2;
        out.theta = theta;
        out.sd1_deviation = sxd * dcos(theta) + syd * dsin(theta);
        out.sd2_deviation = sxd * dsin(theta) + syd * dcos(theta);
        const double omr = 1 - rho * rho;
        out.integral = 2.0 * kPi * x[0] * dsqrt(omr) * sd1 * sd2;
        const double relA = out.deviations[0] / x[0];
        out.integral_deviation = 2.0 * kPi * x[0] * dsqrt(omr) * sd1 * sd2 *
                                 dsqrt(relA * relA + rho * rho * corr_dev * corr_dev / omr / omr +
                                       out.sd1_deviation * out.sd1_deviation / sd1 / sd1 +
                                       out.sd2_deviation * out.sd2_deviation / sd2 / sd2);
        out.status = LZR_ACCEPTED;
        if (model == LZR_MODEL_ELLIPSOIDAL_SYMMETRIC_PSF) {
            if ((0.75 * sd1 > sd2) || (1.25 * sd1 < sd2))
                out.status = LZR_REJECT_ANISOTROPIC;
            else if (((sd1 < w * 0.7) || (sd1 > w * 1.3)) && ((sd2 < w * 0.7) || (sd2 > w * 1.3)))
                out.status = LZR_REJECT_PSF_WIDTH;
        }
    }
}

/* Marks a record whose covariance needs the QR factor: the raw parameters t go to the
 * deviations' slots (finish_deferred reads them there). */
template <int P>
LZR_HD LZR_INLINE void defer_record(LzrFit &out, const double (&t)[P]) {
#pragma unroll
    for (int i = 0; i < P; ++i) out.deviations[i] = t[i];
    out.status = kDeferredQr;
}

/* Post-checks and output record of the reference drivers
 * (source/Localization.cpp:157-211 symmetric, :336-378 fixed width, :506-596 ellipsoidal,
 * :606-631 filter), from the raw end state r of candidate c. */
template <int P>
LZR_HD LZR_INLINE void finish_record(const LzrRawFit &r, const LaneConfig &cfg, const LzrCandidate &c, LzrFit &out,
                                     bool defer_qr = false) {
    LZR_NO_CONTRACT
    const int model = r.model;
    const int h = c.half_width > 0 ? c.half_width : cfg.default_half_width;
    const int n = 2 * h + 1;
    const double xo = (double)(int)(c.x - (double)h), yo = (double)(int)(c.y - (double)h);
    const double xm = xo + n - 1, ym = yo + n - 1;
    const double w = cfg.initial_psf_width;
    double t[P], x[P];
#pragma unroll
    for (int i = 0; i < P; ++i) t[i] = x[i] = r.th[i];
    /* absolute coordinates as the lane forms them (window offset + local) */
    x[x_index<P>()] += xo;
    x[x_index<P>() + 1] += yo;
    const double chi = r.chi;
    clear_record(out);
    out.iterations = r.iterations;
    out.lm_status = r.lm_status;
    out.trials = r.trials;
    out.accepted_steps = r.accepted_steps;
    out.model = (int16_t)model;
    out.chi = chi;
    if constexpr (P == 4) {
#pragma unroll
        for (int i = 0; i < P; ++i) out.params[i] = x[i];
        if (x[0] <= 0) {
            out.status = LZR_REJECT_AMPLITUDE;
            return;
        }
        if (x[1] < xo || x[1] > xm || x[2] < yo || x[2] > ym) {
            out.status = LZR_REJECT_POSITION;
            return;
        }
        /* fitted position within 2 w of the candidate (:349-353) */
        if (dabs(x[1] - c.x) > 2.0 * w || dabs(x[2] - c.y) > 2.0 * w) {
            out.status = LZR_REJECT_DISTANCE;
            return;
        }
        double cd[P];
        if (!covariance<P>(t, n, cfg.inv_sigma, w, cd, defer_qr)) {
            defer_record<P>(out, t);
            return;
        }
        finish_accepted<P>(out, x, chi, h, w, model, cd, 0.0, 0.0, 0.0);
        return;
    } else if constexpr (P == 5) {
        x[1] = dabs(x[1]);
#pragma unroll
        for (int i = 0; i < P; ++i) out.params[i] = x[i];
        if (x[0] <= 0) {
            out.status = LZR_REJECT_AMPLITUDE;
            return;
        }
        if (x[2] < xo || x[2] > xm || x[3] < yo || x[3] > ym) {
            out.status = LZR_REJECT_POSITION;
            return;
        }
        if (x[1] < 0.4 * w || x[1] > w * 1.5) {
            out.status = LZR_REJECT_WIDTH;
            return;
        }
        double cd[P];
        if (!covariance<P>(t, n, cfg.inv_sigma, w, cd, defer_qr)) {
            defer_record<P>(out, t);
            return;
        }
        finish_accepted<P>(out, x, chi, h, w, model, cd, 0.0, 0.0, 0.0);
        return;
    } else {
        x[1] = dabs(x[1]);
        x[2] = dabs(x[2]);
#pragma unroll
        for (int i = 0; i < P; ++i) out.params[i] = x[i];
        if (x[0] <= 0) {
            out.status = LZR_REJECT_AMPLITUDE;
            return;
        }
        if (x[3] < xo || x[3] > xm || x[4] < yo || x[4] > ym) {
            out.status = LZR_REJECT_POSITION;
            return;
        }
        double sd1, sd2, theta;
        if (!eigen2(x[1], x[2], x[5], sd1, sd2, theta)) {
            out.status = LZR_REJECT_EIGEN;
            return;
        }
        if (!(((sd1 >= 0.4 * w) && (sd1 <= 1.5 * w)) || ((sd2 >= 0.4 * w) && (sd2 <= 1.5 * w)))) {
            out.status = LZR_REJECT_WIDTH;
            return;
        }
        double cd[P];
        if (!covariance<P>(t, n, cfg.inv_sigma, w, cd, defer_qr)) {
            defer_record<P>(out, t);
            return;
        }
        finish_accepted<P>(out, x, chi, h, w, model, cd, sd1, sd2, theta);
    }
}

/* Completes a record finish_record deferred (status kDeferredQr): the covariance from the QR
 * factor of the reference Jacobian at the raw parameters kept in the deviations' slots, then
 * the accepted record's remaining fields, exactly as finish_record computes them inline. */
template <int P>
LZR_HD LZR_INLINE void finish_deferred(LzrFit &out, const LaneConfig &cfg, const LzrCandidate &c) {
    LZR_NO_CONTRACT
    const int h = c.half_width > 0 ? c.half_width : cfg.default_half_width;
    const int n = 2 * h + 1;
    const double w = cfg.initial_psf_width;
    double t[P], x[P], cd[P];
#pragma unroll
    for (int i = 0; i < P; ++i) {
        t[i] = out.deviations[i];
        x[i] = out.params[i];
    }
    covariance_qr<P>(t, n, cfg.inv_sigma, w, cd);
    double sd1 = 0.0, sd2 = 0.0, theta = 0.0;
    if constexpr (P == 7) eigen2(x[1], x[2], x[5], sd1, sd2, theta);
    finish_accepted<P>(out, x, out.chi, h, w, out.model, cd, sd1, sd2, theta);
}

/* Separable sums of start_residual over an n x n window (n <= NMAX): the loops are unrolled
 * to NMAX with the columns beyond n masked, so the column factors stay in registers. */
template <int NMAX, typename FrameT>
LZR_HD LZR_INLINE double start_sse(const FrameT *src, int nx, int n, double x0, double y0, double k, double A,
                                   double b) {
    LZR_NO_CONTRACT
    double ex[NMAX];
    double sx = 0.0, sxx = 0.0;
#pragma unroll
    for (int i = 0; i < NMAX; ++i) {
        const double u = (double)i - x0;
        ex[i] = i < n ? dexp(-(u * u) * k) : 0.0;
        sx += ex[i];
        sxx = dfma(ex[i], ex[i], sxx);
    }
    double sy = 0.0, syy = 0.0, sed = 0.0, sd = 0.0, sdd = 0.0;
    for (int j = 0; j < n; ++j) {
        const double v = (double)j - y0;
        const double ey = dexp(-(v * v) * k);
        sy += ey;
        syy = dfma(ey, ey, syy);
        const FrameT *row = src + (size_t)j * nx;
        double red = 0.0;
#pragma unroll
        for (int i = 0; i < NMAX; ++i) {
            if (i < n) {
                const double d = (double)row[i];
                red = dfma(ex[i], d, red);
                sd += d;
                sdd = dfma(d, d, sdd);
            }
        }
        sed = dfma(ey, red, sed);
    }
    const double N = (double)(n * n);
    return ((A * A) * (sxx * syy) + 2.0 * (A * b) * (sx * sy) - 2.0 * A * sed) + ((N * b) * b - 2.0 * b * sd + sdd);
}

/* |f| at candidate c's reference start point (source/Localization.cpp:138-142, :485-491: the
 * particle's intensity, background and position, the initial width; rho 0) over the window the
 * fit uses (frames in global memory); NaN when the window leaves the frame or the request is
 * invalid.  At that point every model is the same round Gaussian e(i, j) = Ex(i) Ey(j), so
 * sum_ij (A e + b - D)^2 = A^2 Sxx Syy + 2 A b Sx Sy - 2 A sum_j Ey(j) sum_i Ex(i) D(i, j)
 *                          + N b^2 - 2 b sum D + sum D^2
 * costs 2n exponentials and one multiply-add per pixel.  The net-start fallback compares a
 * net-start fit's final |f| with it (the normalized residuals chi/sqrt(dof) share the dof). */
template <typename FrameT>
LZR_HD LZR_INLINE double start_residual(const LaneConfig &cfg, const LzrCandidate &c, const FrameT *frames, int nx,
                                        int ny, int nframes) {
    LZR_NO_CONTRACT
    LaneState<5> s;
    LzrFit unused;
    if (!lane_start<5>(s, cfg, c, 0, nx, ny, nframes, LZR_MODEL_SYMMETRIC, unused)) return __builtin_nan("");
    const int n = s.n;
    const double k = 0.5 / (cfg.initial_psf_width * cfg.initial_psf_width);
    const double x0 = c.x - s.xoff, y0 = c.y - s.yoff; /* window coordinates */
    const FrameT *src = frames + ((size_t)s.frame * ny + s.yoff) * nx + s.xoff;
    const double A = c.intensity, b = c.background;
    static_assert(2 * LZR_MAX_HALF_WIDTH + 1 <= 31, "start_residual window classes");
    double sse;
    if (n <= 11)
        sse = start_sse<11>(src, nx, n, x0, y0, k, A, b);
    else if (n <= 19)
        sse = start_sse<19>(src, nx, n, x0, y0, k, A, b);
    else
        sse = start_sse<31>(src, nx, n, x0, y0, k, A, b);
    return dsqrt(dmax(sse, 0.0)) * cfg.inv_sigma;
}

/* Epilogue entry: finishes a record written by store_final (no-op for records already
 * final, i.e. rejected before fitting). */
LZR_HD LZR_INLINE void finish_any(LzrFit &rec, const LaneConfig &cfg, const LzrCandidate &c, bool defer_qr = false) {
    LzrRawFit r;
    __builtin_memcpy(&r, &rec, sizeof r);
    if (r.status != kPending) return;
    if (r.model == LZR_MODEL_SYMMETRIC)
        finish_record<5>(r, cfg, c, rec, defer_qr);
    else if (r.model == LZR_MODEL_FIXED_WIDTH)
        finish_record<4>(r, cfg, c, rec, defer_qr);
    else
        finish_record<7>(r, cfg, c, rec, defer_qr);
}

/* Completes a deferred record (status kDeferredQr), whatever its model. */
LZR_HD LZR_INLINE void finish_any_deferred(LzrFit &rec, const LaneConfig &cfg, const LzrCandidate &c) {
    if (rec.model == LZR_MODEL_SYMMETRIC)
        finish_deferred<5>(rec, cfg, c);
    else if (rec.model == LZR_MODEL_FIXED_WIDTH)
        finish_deferred<4>(rec, cfg, c);
    else
        finish_deferred<7>(rec, cfg, c);
}

}  // namespace lzr

#endif /* LZR_LANE_H */
