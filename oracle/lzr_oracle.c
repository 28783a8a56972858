/*
 * lzr_oracle.c -- FP64 CPU oracle for the localization-fitting path.
 *
 * TEST INFRASTRUCTURE ONLY (see lzr_oracle.h).  Built with -O2 and without
 * fast-math so that it is deterministic.
 *
 * The LM engine restates GSL 1.16 lmsder, which is MINPACK lmder in mode 1
 * (automatic column scaling), one outer step per iterate() call:
 *   set      evaluate f and J, diag = column norms of J (0 -> 1), delta = 100*|D x|
 *            (100 if zero), par = 0, iter = 1, QR with column pivoting of J.
 *   iterate  qtf = Q^T f, scaled gradient norm, then up to 10 trials:
 *            lmpar (Levenberg-Marquardt parameter by secant iteration on |D p| = delta,
 *            at most 10 steps, 10% tolerance) -> trial x+p -> actual/predicted reduction
 *            -> trust-radius update (0.1/0.25/0.5/0.75) -> accept if ratio >= 1e-4.
 * GSL 1.16 choices frozen here (source absent from the container; parity unpinned):
 *   (G1) update_diag after an accepted step takes column norms over the first p rows
 *        of J only (lmutil.c loops i < diag->size); MINPACK uses all rows.  Option
 *        diag_update_rows = p (default) / 0 (MINPACK).
 *   (G2) an iterate() call gives up with GSL_ENOPROG after 10 unsuccessful trials;
 *        MINPACK retries without a cap.  Option inner_cap = 10 (default) / 0.
 *   (G3) the gradient-norm test (GSL_ETOLG) scales by diag; MINPACK scales by the
 *        current column norms.  Only reachable when the gradient vanishes.
 *   (G4) lmpar clamps the initial par with "if par > paru: paru, else if par < parl:
 *        parl"; MINPACK takes min(max(par, parl), paru).  Identical unless parl > paru.
 *   (G5) iterate() returns GSL_SUCCESS without a step when |f| == 0.
 * The MINPACK parts are checked against SciPy's lmder in tests/test_oracle_minpack.py
 * with (G1, G2) switched to MINPACK mode.
 */
#define _GNU_SOURCE
#include "lzr_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define ORC_PI 3.1415926535897932384626433 /* source/Defines.h:97 */
#define ORC_SQRT2 1.4142135623730950488    /* source/Defines.h:98 */
#define ORC_MAXN ((2 * LZR_MAX_HALF_WIDTH + 1) * (2 * LZR_MAX_HALF_WIDTH + 1))
#define ORC_MAXP 7

enum {
    ST_OK = 0,
    ST_FAILURE = -1,
    ST_CONTINUE = -2,
    ST_ENOPROG = 27,
    ST_ETOLF = 29,
    ST_ETOLX = 30,
    ST_ETOLG = 31
};

void orc_default_options(OrcOptions *opt) {
    opt->max_iterations = 200; /* source/Localization.cpp:155 */
    opt->epsabs = 1e-4;        /* source/Localization.cpp:154 */
    opt->epsrel = 1e-4;
    opt->diag_update_rows = -1; /* -1: p rows (GSL 1.16, G1) */
    opt->inner_cap = 10;        /* G2 */
}

/* ------------------------------------------------------------------------------ */
/* Models and Jacobians (source/Localization.cpp:1007-1048, 1078-1126, 1128-1181,  */
/* 1227-1295).  Residual index k = i + n*j, pixel (xoff+i, yoff+j).                */
/* ------------------------------------------------------------------------------ */

static int model_symmetric(const double *t, const double *data, int n, double xoff, double yoff,
                           double sigma, double *f) {
    const double amp = t[0], r = t[1], x0 = t[2], y0 = t[3], off = t[4];
    if (r == 0) return ST_FAILURE; /* :1029-1031 */
    const double two_r2 = 2.0 * r * r;
    for (int j = 0; j < n; ++j) {
        const double ddy = y0 - (yoff + (double)j);
        for (int i = 0; i < n; ++i) {
            const double ddx = x0 - (xoff + (double)i);
            const double q = -(ddx * ddx + ddy * ddy);
            const int k = i + n * j;
            f[k] = ((exp(q / two_r2) * amp + off) - data[k]) / sigma;
        }
    }
    return ST_OK;
}

static int jacobian_symmetric(const double *t, int n, double xoff, double yoff, double sigma,
                              double *J) {
    const double amp = t[0], r = t[1], x0 = t[2], y0 = t[3];
    if (r == 0) return ST_FAILURE;
    const double two_r2 = 2.0 * r * r;
    /* Width column: the reference divides by 2 r^2 * SQRT2 * r (:1151-1152), i.e. the
     * exact derivative A e d^2 / r^3 scaled by 1/sqrt(2).  Kept as is. */
    const double width_den = two_r2 * ORC_SQRT2 * r;
    for (int j = 0; j < n; ++j) {
        const double y = yoff + (double)j;
        for (int i = 0; i < n; ++i) {
            const double x = xoff + (double)i;
            const double ddx = x0 - x, ddy = y0 - y;
            const double d2 = ddx * ddx + ddy * ddy;
            const double e = exp(-d2 / two_r2);
            double *row = J + 5 * (i + n * j);
            row[0] = e / sigma;
            row[1] = (2.0 * d2 / width_den) * e * amp / sigma;
            row[2] = (2.0 * (x - x0) * e * amp) / (two_r2 * sigma);
            row[3] = (2.0 * (y - y0) * e * amp) / (two_r2 * sigma);
            row[4] = 1.0 / sigma;
        }
    }
    return ST_OK;
}

static int model_ellipsoidal(const double *t, const double *data, int n, double xoff,
                             double yoff, double sigma, double *f) {
    const double amp = t[0], sx = t[1], sy = t[2], x0 = t[3], y0 = t[4], rho = t[5], off = t[6];
    if (sx == 0 || sy == 0) return ST_FAILURE; /* :1098-1100 */
    const double pre = -1.0 / (2.0 * (1.0 - rho * rho));
    const double sx2 = sx * sx, sy2 = sy * sy;
    for (int j = 0; j < n; ++j) {
        const double dy = (yoff + (double)j) - y0;
        for (int i = 0; i < n; ++i) {
            const double dx = (xoff + (double)i) - x0;
            const double q = dx * dx / sx2 + dy * dy / sy2 - 2.0 * rho * dx * dy / sx / sy;
            const int k = i + n * j;
            f[k] = ((exp(pre * q) * amp + off) - data[k]) / sigma;
        }
    }
    return ST_OK;
}

static int jacobian_ellipsoidal(const double *t, int n, double xoff, double yoff, double sigma,
                                double *J) {
    const double amp = t[0], sx = t[1], sy = t[2], x0 = t[3], y0 = t[4], rho = t[5];
    if (sx == 0 || sy == 0) return ST_FAILURE;
    const double omr = 1.0 - rho * rho;
    const double den = 2.0 * omr * sigma;
    const double pre = -1.0 / (2.0 * omr);
    const double sx2 = sx * sx, sy2 = sy * sy, sxy = sx * sy;
    for (int j = 0; j < n; ++j) {
        const double dy = (yoff + (double)j) - y0;
        for (int i = 0; i < n; ++i) {
            const double dx = (xoff + (double)i) - x0;
            const double dx2 = dx * dx, dy2 = dy * dy;
            const double e = exp(pre * (dx2 / sx2 + dy2 / sy2 - 2.0 * rho * dx * dy / sxy));
            double *row = J + 7 * (i + n * j);
            row[0] = e / sigma;
            row[1] = -amp * (2.0 * rho * dx * dy / sx2 / sy - 2.0 * dx2 / (sx2 * sx)) * e / den;
            row[2] = -amp * (2.0 * rho * dx * dy / sx / sy2 - 2.0 * dy2 / (sy2 * sy)) * e / den;
            row[3] = -amp * (2.0 * rho * dy / sxy - 2.0 * dx / sx2) * e / den;
            row[4] = -amp * (2.0 * rho * dx / sxy - 2.0 * dy / sy2) * e / den;
            /* Correlation column as written at :1280: the quadratic form in the second
             * term is (dx^2 + dy^2 - 2 rho dx dy)/(sx sy), exact only when sx == sy. */
            row[5] = amp *
                     (dx * dy / omr / sxy - rho * (dx2 + dy2 - 2.0 * rho * dx * dy) / sxy / omr / omr) *
                     e / sigma;
            row[6] = 1.0 / sigma;
        }
    }
    return ST_OK;
}

int orc_model(int model, const double *theta, const double *data, int n, double xoff,
              double yoff, double sigma, double *f) {
    if (model == LZR_MODEL_SYMMETRIC)
        return model_symmetric(theta, data, n, xoff, yoff, sigma, f);
    return model_ellipsoidal(theta, data, n, xoff, yoff, sigma, f);
}

int orc_jacobian(int model, const double *theta, int n, double xoff, double yoff, double sigma,
                 double *J) {
    if (model == LZR_MODEL_SYMMETRIC) return jacobian_symmetric(theta, n, xoff, yoff, sigma, J);
    return jacobian_ellipsoidal(theta, n, xoff, yoff, sigma, J);
}

/* FitFunction_/Jacobian_FixedWidthGaussian (source/Localization.cpp:1050-1076,
 * :1183-1225): the symmetric model at r = width, Jacobian columns A, x0, y0, b. */
int orc_model_fixed_width(const double *theta, double width, const double *data, int n, double xoff,
                          double yoff, double sigma, double *f) {
    const double t5[5] = {theta[0], width, theta[1], theta[2], theta[3]};
    return model_symmetric(t5, data, n, xoff, yoff, sigma, f);
}

int orc_jacobian_fixed_width(const double *theta, double width, int n, double xoff, double yoff,
                             double sigma, double *J) {
    const double t5[5] = {theta[0], width, theta[1], theta[2], theta[3]};
    double *J5 = (double *)malloc(sizeof(double) * 5 * n * n);
    const int st = jacobian_symmetric(t5, n, xoff, yoff, sigma, J5);
    if (st == ST_OK) {
        for (int k = 0; k < n * n; ++k) {
            J[4 * k + 0] = J5[5 * k + 0];
            J[4 * k + 1] = J5[5 * k + 2];
            J[4 * k + 2] = J5[5 * k + 3];
            J[4 * k + 3] = J5[5 * k + 4];
        }
    }
    free(J5);
    return st;
}

/* ------------------------------------------------------------------------------ */
/* Dense linear algebra                                                            */
/* ------------------------------------------------------------------------------ */

double orc_nrm2(int n, const double *x, int stride) {
    if (n <= 0) return 0.0;
    if (n == 1) return fabs(x[0]);
    double scale = 0.0, ssq = 1.0;
    for (int i = 0; i < n; ++i) {
        const double v = x[(long)i * stride];
        if (v != 0.0) {
            const double a = fabs(v);
            if (scale < a) {
                const double q = scale / a;
                ssq = 1.0 + ssq * q * q;
                scale = a;
            } else {
                const double q = a / scale;
                ssq += q * q;
            }
        }
    }
    return scale * sqrt(ssq);
}

static double scaled_norm(int p, const double *d, const double *v) {
    double s = 0.0;
    for (int i = 0; i < p; ++i) {
        const double u = d[i] * v[i];
        s += u * u;
    }
    return sqrt(s);
}

static double plain_norm(int p, const double *v) { return orc_nrm2(p, v, 1); }

/* Householder vector for column segment a[0..m-1] (stride lda): on return a[0] = beta,
 * a[1..] = v[1..] (v[0] = 1 implied); returns tau.  H = I - tau v v^T maps the segment
 * to beta e1. */
static double householder(double *a, int m, int lda) {
    if (m <= 1) return 0.0;
    const double xnorm = orc_nrm2(m - 1, a + lda, lda);
    if (xnorm == 0.0) return 0.0;
    const double alpha = a[0];
    const double beta = (alpha >= 0.0 ? -1.0 : 1.0) * hypot(alpha, xnorm);
    const double tau = (beta - alpha) / beta;
    const double s = 1.0 / (alpha - beta);
    for (int i = 1; i < m; ++i) a[(long)i * lda] *= s;
    a[0] = beta;
    return tau;
}

/* Applies H = I - tau v v^T (v stored below the pivot, v0 = 1) to the length-m column
 * segment c (stride ldc). */
static void apply_householder(double tau, const double *v, int ldv, double *c, int ldc, int m) {
    if (tau == 0.0) return;
    double w = c[0];
    for (int i = 1; i < m; ++i) w += c[(long)i * ldc] * v[(long)i * ldv];
    c[0] -= tau * w;
    for (int i = 1; i < m; ++i) c[(long)i * ldc] -= tau * v[(long)i * ldv] * w;
}

/* QR factorisation with column pivoting of the N x p row-major matrix A (in place):
 * upper triangle = R, below = Householder vectors, perm[j] = original column of
 * column j.  Pivot = largest remaining column norm, norms downdated and recomputed
 * when cancellation exceeds sqrt(20) sqrt(eps) (GSL linalg/qrpt.c scheme). */
static void qrpt(double *A, int N, int p, double *tau, int *perm) {
    double cn[ORC_MAXP];
    for (int j = 0; j < p; ++j) {
        perm[j] = j;
        cn[j] = orc_nrm2(N, A + j, p);
    }
    const int kmin = N < p ? N : p;
    for (int i = 0; i < kmin; ++i) {
        int kmax = i;
        double best = cn[i];
        for (int j = i + 1; j < p; ++j)
            if (cn[j] > best) {
                best = cn[j];
                kmax = j;
            }
        if (kmax != i) {
            for (int r = 0; r < N; ++r) {
                double t = A[r * p + i];
                A[r * p + i] = A[r * p + kmax];
                A[r * p + kmax] = t;
            }
            int tp = perm[i];
            perm[i] = perm[kmax];
            perm[kmax] = tp;
            double tn = cn[i];
            cn[i] = cn[kmax];
            cn[kmax] = tn;
        }
        double *col = A + (long)i * p + i;
        tau[i] = householder(col, N - i, p);
        for (int j = i + 1; j < p; ++j) apply_householder(tau[i], col, p, A + (long)i * p + j, p, N - i);
        if (i + 1 < N) {
            for (int j = i + 1; j < p; ++j) {
                const double x = cn[j];
                if (x > 0.0) {
                    const double t = A[(long)i * p + j] / x;
                    double y = fabs(t) >= 1.0 ? 0.0 : x * sqrt(1.0 - t * t);
                    if (fabs(y / x) < sqrt(20.0) * sqrt(DBL_EPSILON))
                        y = orc_nrm2(N - i - 1, A + (long)(i + 1) * p + j, p);
                    cn[j] = y;
                }
            }
        }
    }
}

/* qtf = Q^T f from the stored reflections. */
static void apply_qt(const double *QR, int N, int p, const double *tau, double *v) {
    const int kmin = N < p ? N : p;
    for (int i = 0; i < kmin; ++i)
        apply_householder(tau[i], QR + (long)i * p + i, p, v + i, 1, N - i);
}

#define RR(i, j) R[(i) * ldr + (j)]

/* Least-squares solve of [R P^T; lambda D] x ~ [qtb; 0] by Givens rotations (MINPACK
 * qrsolv).  S (p x p) receives the strict lower triangle of the rotated factor
 * (transposed), sdiag its diagonal. */
static void qrsolv(int p, const double *R, int ldr, const int *perm, double lambda,
                   const double *diag, const double *qtb, double *x, double *S, double *sdiag) {
    double rdiag[ORC_MAXP], wa[ORC_MAXP];
    for (int j = 0; j < p; ++j) {
        for (int i = j + 1; i < p; ++i) S[i * p + j] = RR(j, i);
        S[j * p + j] = RR(j, j);
        rdiag[j] = RR(j, j);
        wa[j] = qtb[j];
    }
    for (int j = 0; j < p; ++j) {
        const double dj = lambda * diag[perm[j]];
        if (dj != 0.0) {
            for (int k = j; k < p; ++k) sdiag[k] = 0.0;
            sdiag[j] = dj;
            double extra = 0.0; /* the one element of [qtb, 0] beyond row p */
            for (int k = j; k < p; ++k) {
                if (sdiag[k] == 0.0) continue;
                const double rkk = S[k * p + k];
                double sn, cs;
                if (fabs(rkk) < fabs(sdiag[k])) {
                    const double cot = rkk / sdiag[k];
                    sn = 0.5 / sqrt(0.25 + 0.25 * cot * cot);
                    cs = sn * cot;
                } else {
                    const double tn = sdiag[k] / rkk;
                    cs = 0.5 / sqrt(0.25 + 0.25 * tn * tn);
                    sn = cs * tn;
                }
                S[k * p + k] = cs * rkk + sn * sdiag[k];
                const double w = cs * wa[k] + sn * extra;
                extra = -sn * wa[k] + cs * extra;
                wa[k] = w;
                for (int i = k + 1; i < p; ++i) {
                    const double sik = S[i * p + k];
                    const double t = cs * sik + sn * sdiag[i];
                    sdiag[i] = -sn * sik + cs * sdiag[i];
                    S[i * p + k] = t;
                }
            }
        }
        sdiag[j] = S[j * p + j];
        S[j * p + j] = rdiag[j];
    }
    int nsing = p;
    for (int j = 0; j < p; ++j)
        if (sdiag[j] == 0.0) {
            nsing = j;
            break;
        }
    for (int j = nsing; j < p; ++j) wa[j] = 0.0;
    for (int j = nsing - 1; j >= 0; --j) {
        double sum = 0.0;
        for (int i = j + 1; i < nsing; ++i) sum += S[i * p + j] * wa[i];
        wa[j] = (wa[j] - sum) / sdiag[j];
    }
    for (int j = 0; j < p; ++j) x[perm[j]] = wa[j];
}

/* Levenberg-Marquardt parameter (MINPACK lmpar): returns x with
 * (J^T J + par D^2) x = J^T f and |D x| within 10% of delta, or the Gauss-Newton x
 * with par = 0 when |D x_GN| <= 1.1 delta. */
static void lmpar(int p, const double *R, int ldr, const int *perm, const double *qtf,
                  const double *diag, double delta, double *par_io, double *x) {
    double wa[ORC_MAXP], w[ORC_MAXP], S[ORC_MAXP * ORC_MAXP], sdiag[ORC_MAXP];
    int nsing = p;
    for (int j = 0; j < p; ++j)
        if (RR(j, j) == 0.0) {
            nsing = j;
            break;
        }
    for (int j = 0; j < p; ++j) wa[j] = j < nsing ? qtf[j] : 0.0;
    for (int j = nsing - 1; j >= 0; --j) {
        wa[j] /= RR(j, j);
        for (int i = 0; i < j; ++i) wa[i] -= RR(i, j) * wa[j];
    }
    for (int j = 0; j < p; ++j) x[perm[j]] = wa[j];

    double dxnorm = scaled_norm(p, diag, x);
    double fp = dxnorm - delta;
    if (fp <= 0.1 * delta) {
        *par_io = 0.0;
        return;
    }

    double parl = 0.0;
    if (nsing == p) {
        for (int j = 0; j < p; ++j) {
            const int l = perm[j];
            w[j] = diag[l] * (diag[l] * x[l] / dxnorm);
        }
        for (int j = 0; j < p; ++j) {
            double sum = 0.0;
            for (int i = 0; i < j; ++i) sum += RR(i, j) * w[i];
            w[j] = (w[j] - sum) / RR(j, j);
        }
        const double wn = plain_norm(p, w);
        const double phider = wn * wn;
        parl = phider > 0 ? fp / (delta * phider) : 0.0;
    }

    for (int j = 0; j < p; ++j) {
        double sum = 0.0;
        for (int i = 0; i <= j; ++i) sum += RR(i, j) * qtf[i];
        w[j] = sum / diag[perm[j]];
    }
    const double gnorm = plain_norm(p, w);
    double paru = gnorm / delta;
    if (paru == 0.0) paru = DBL_MIN / (delta < 0.1 ? delta : 0.1);

    double par = *par_io;
    if (par > paru)
        par = paru;
    else if (par < parl)
        par = parl;
    if (par == 0.0) par = gnorm / dxnorm;

    for (int iter = 1;; ++iter) {
        if (par == 0.0) par = (0.001 * paru > DBL_MIN) ? 0.001 * paru : DBL_MIN;
        qrsolv(p, R, ldr, perm, sqrt(par), diag, qtf, x, S, sdiag);
        dxnorm = scaled_norm(p, diag, x);
        const double fp_old = fp;
        fp = dxnorm - delta;
        if (fabs(fp) <= 0.1 * delta) break;
        if (parl == 0.0 && fp <= fp_old && fp_old < 0.0) break;
        if (iter == 10) break;
        for (int j = 0; j < p; ++j) {
            const int l = perm[j];
            w[j] = diag[l] * (diag[l] * x[l] / dxnorm);
        }
        for (int j = 0; j < p; ++j) {
            w[j] /= sdiag[j];
            for (int i = j + 1; i < p; ++i) w[i] -= S[i * p + j] * w[j];
        }
        const double wn = plain_norm(p, w);
        const double parc = fp / (delta * wn * wn);
        if (fp > 0.0) {
            if (par > parl) parl = par;
        } else if (fp < 0.0) {
            if (par < paru) paru = par;
        }
        par = parl > par + parc ? parl : par + parc;
    }
    *par_io = par;
}

/* ------------------------------------------------------------------------------ */
/* lmsder solver state                                                             */
/* ------------------------------------------------------------------------------ */

typedef struct Lm {
    int N, p, model, n;
    double fixed_width; /* LZR_MODEL_FIXED_WIDTH: the width r held at initialPSFWidth */
    const double *data;
    double xoff, yoff, sigma;
    OrcOptions opt;
    double x[ORC_MAXP], dx[ORC_MAXP], diag[ORC_MAXP], tau[ORC_MAXP];
    double x_trial[ORC_MAXP], qtf[ORC_MAXN], f[ORC_MAXN], f_trial[ORC_MAXN];
    double J[ORC_MAXN * ORC_MAXP], R[ORC_MAXN * ORC_MAXP];
    int perm[ORC_MAXP];
    double delta, par, fnorm, xnorm;
    int iter;
    int trials, accepted;
    double *trace;
    double *trial_trace;
    int max_trace;
} Lm;

static int eval_f(Lm *s, const double *t, double *f) {
    if (s->model == LZR_MODEL_FIXED_WIDTH)
        return orc_model_fixed_width(t, s->fixed_width, s->data, s->n, s->xoff, s->yoff, s->sigma, f);
    return orc_model(s->model, t, s->data, s->n, s->xoff, s->yoff, s->sigma, f);
}

static int eval_J(Lm *s, const double *t, double *J) {
    if (s->model == LZR_MODEL_FIXED_WIDTH)
        return orc_jacobian_fixed_width(t, s->fixed_width, s->n, s->xoff, s->yoff, s->sigma, J);
    return orc_jacobian(s->model, t, s->n, s->xoff, s->yoff, s->sigma, J);
}

static void column_norms(const double *J, int rows, int p, double *out) {
    for (int j = 0; j < p; ++j) {
        double sum = 0.0;
        for (int i = 0; i < rows; ++i) sum += J[i * p + j] * J[i * p + j];
        out[j] = sqrt(sum == 0.0 ? 1.0 : sum);
    }
}

static int lm_set(Lm *s) {
    int st = eval_f(s, s->x, s->f);
    if (st) return st;
    st = eval_J(s, s->x, s->J);
    if (st) return st;
    s->par = 0.0;
    s->iter = 1;
    s->fnorm = orc_nrm2(s->N, s->f, 1);
    for (int j = 0; j < s->p; ++j) s->dx[j] = 0.0;
    column_norms(s->J, s->N, s->p, s->diag);
    s->xnorm = scaled_norm(s->p, s->diag, s->x);
    s->delta = s->xnorm > 0.0 ? 100.0 * s->xnorm : 100.0;
    memcpy(s->R, s->J, sizeof(double) * s->N * s->p);
    qrpt(s->R, s->N, s->p, s->tau, s->perm);
    return ST_OK;
}

static void record_trace(Lm *s) {
    if (s->trace && s->trials <= s->max_trace)
        memcpy(s->trace + (long)(s->trials - 1) * s->p, s->x, sizeof(double) * s->p);
}

static int lm_iterate(Lm *s) {
    const int N = s->N, p = s->p;
    if (s->fnorm == 0.0) return ST_OK; /* G5 */
    memcpy(s->qtf, s->f, sizeof(double) * N);
    apply_qt(s->R, N, p, s->tau, s->qtf);

    /* scaled gradient norm (G3) */
    double gnorm = 0.0;
    for (int j = 0; j < p; ++j) {
        double sum = 0.0;
        for (int i = 0; i <= j; ++i) sum += s->R[i * p + j] * s->qtf[i];
        const double g = fabs(sum / s->diag[s->perm[j]]);
        if (g > gnorm) gnorm = g;
    }
    gnorm /= s->fnorm;

    double step[ORC_MAXP], rpdx[ORC_MAXP];
    for (int inner = 1;; ++inner) {
        lmpar(p, s->R, p, s->perm, s->qtf, s->diag, s->delta, &s->par, step);
        for (int j = 0; j < p; ++j) {
            s->dx[j] = -step[j];
            s->x_trial[j] = s->x[j] + s->dx[j];
        }
        const double pnorm = scaled_norm(p, s->diag, s->dx);
        if (s->iter == 1 && pnorm < s->delta) s->delta = pnorm;

        s->trials++;
        int st = eval_f(s, s->x_trial, s->f_trial);
        if (s->trial_trace && s->trials <= s->max_trace) {
            memcpy(s->trial_trace + (long)(s->trials - 1) * p, s->x_trial, sizeof(double) * p);
            s->trial_trace[(long)s->max_trace * p + s->trials - 1] = st ? NAN : orc_nrm2(N, s->f_trial, 1);
        }
        if (st) {
            record_trace(s);
            return st;
        }
        const double fnorm1 = orc_nrm2(N, s->f_trial, 1);
        double actred = -1.0;
        if (0.1 * fnorm1 < s->fnorm) {
            const double u = fnorm1 / s->fnorm;
            actred = 1.0 - u * u;
        }
        for (int i = 0; i < p; ++i) {
            double sum = 0.0;
            for (int j = i; j < p; ++j) sum += s->R[i * p + j] * s->dx[s->perm[j]];
            rpdx[i] = sum;
        }
        const double fnorm1p = plain_norm(p, rpdx);
        const double t1 = fnorm1p / s->fnorm;
        const double t2 = (sqrt(s->par) * pnorm) / s->fnorm;
        const double prered = t1 * t1 + t2 * t2 / 0.5;
        const double dirder = -(t1 * t1 + t2 * t2);
        const double ratio = prered > 0.0 ? actred / prered : 0.0;

        if (ratio > 0.25) {
            if (s->par == 0.0 || ratio >= 0.75) {
                s->delta = pnorm / 0.5;
                s->par *= 0.5;
            }
        } else {
            double temp = actred >= 0.0 ? 0.5 : 0.5 * dirder / (dirder + 0.5 * actred);
            if (0.1 * fnorm1 >= s->fnorm || temp < 0.1) temp = 0.1;
            s->delta = temp * (s->delta < pnorm / 0.1 ? s->delta : pnorm / 0.1);
            s->par /= temp;
        }

        if (ratio >= 1e-4) {
            memcpy(s->x, s->x_trial, sizeof(double) * p);
            memcpy(s->f, s->f_trial, sizeof(double) * N);
            s->accepted++;
            record_trace(s);
            st = eval_J(s, s->x, s->J);
            if (st) return st;
            s->xnorm = scaled_norm(p, s->diag, s->x);
            s->fnorm = fnorm1;
            s->iter++;
            double cn[ORC_MAXP];
            int rows = s->opt.diag_update_rows < 0 ? p : s->opt.diag_update_rows;
            if (rows == 0 || rows > N) rows = N;
            column_norms(s->J, rows, p, cn); /* G1 */
            for (int j = 0; j < p; ++j)
                if (cn[j] > s->diag[j]) s->diag[j] = cn[j];
            memcpy(s->R, s->J, sizeof(double) * N * p);
            qrpt(s->R, N, p, s->tau, s->perm);
            return ST_OK;
        }
        record_trace(s);
        if (fabs(actred) <= DBL_EPSILON && prered <= DBL_EPSILON && 0.5 * ratio <= 1.0)
            return ST_ETOLF;
        if (s->delta <= DBL_EPSILON * s->xnorm) return ST_ETOLX;
        if (gnorm <= DBL_EPSILON) return ST_ETOLG;
        if (s->opt.inner_cap > 0 && inner >= s->opt.inner_cap) return ST_ENOPROG; /* G2 */
    }
}

int orc_test_delta(int p, const double *dx, const double *x, double epsabs, double epsrel) {
    for (int i = 0; i < p; ++i) {
        const double tol = epsabs + epsrel * fabs(x[i]);
        if (!(fabs(dx[i]) < tol)) return ST_CONTINUE;
    }
    return ST_OK;
}

void orc_covar(int N, int p, const double *J, double *C) {
    double *R = (double *)malloc(sizeof(double) * N * p);
    double tau[ORC_MAXP];
    int perm[ORC_MAXP];
    memcpy(R, J, sizeof(double) * N * p);
    qrpt(R, N, p, tau, perm);
    /* inverse of R in place (upper triangle), stopping at the first zero pivot */
    double Ri[ORC_MAXP * ORC_MAXP];
    memset(Ri, 0, sizeof Ri);
    for (int i = 0; i < p; ++i)
        for (int j = i; j < p; ++j) Ri[i * p + j] = R[i * p + j];
    int rank = 0;
    for (int k = 0; k < p; ++k) {
        const double rkk = Ri[k * p + k];
        if (fabs(rkk) <= 0.0) break; /* epsrel = 0: only exact zeros */
        Ri[k * p + k] = 1.0 / rkk;
        for (int j = 0; j < k; ++j) {
            const double t = Ri[j * p + k] / rkk;
            Ri[j * p + k] = 0.0;
            for (int i = 0; i <= j; ++i) Ri[i * p + k] -= t * Ri[i * p + j];
        }
        rank = k + 1;
    }
    /* (R^T R)^-1 = Rinv Rinv^T on the leading rank x rank block */
    double M[ORC_MAXP * ORC_MAXP];
    memset(M, 0, sizeof M);
    for (int i = 0; i < rank; ++i)
        for (int j = 0; j < rank; ++j) {
            double s = 0.0;
            for (int k = (i > j ? i : j); k < rank; ++k) s += Ri[i * p + k] * Ri[j * p + k];
            M[i * p + j] = s;
        }
    for (int i = 0; i < p; ++i)
        for (int j = 0; j < p; ++j) C[perm[i] * p + perm[j]] = M[i * p + j];
    free(R);
}

int orc_lm_run(int model, const double *data, int n, double xoff, double yoff, double sigma,
               const double *theta0, const OrcOptions *opt, double *theta, double *f_out,
               double *J_out, int *iterations, int *lm_status, int *accepted, double *trace,
               int max_trace, double fixed_width, double *trial_trace) {
    Lm *s = (Lm *)calloc(1, sizeof(Lm));
    s->model = (model == LZR_MODEL_SYMMETRIC || model == LZR_MODEL_FIXED_WIDTH) ? model : LZR_MODEL_ELLIPSOIDAL;
    s->p = s->model == LZR_MODEL_SYMMETRIC ? 5 : (s->model == LZR_MODEL_FIXED_WIDTH ? 4 : 7);
    s->fixed_width = fixed_width;
    s->n = n;
    s->N = n * n;
    s->data = data;
    s->xoff = xoff;
    s->yoff = yoff;
    s->sigma = sigma;
    s->opt = *opt;
    s->trace = trace;
    s->trial_trace = trial_trace;
    s->max_trace = max_trace;
    memcpy(s->x, theta0, sizeof(double) * s->p);
    int status = lm_set(s);
    int it = 0;
    if (status == ST_OK) {
        do {
            it++;
            status = lm_iterate(s);
            if (status != ST_OK) break;
            status = orc_test_delta(s->p, s->dx, s->x, opt->epsabs, opt->epsrel);
        } while (status == ST_CONTINUE && it < opt->max_iterations);
    }
    memcpy(theta, s->x, sizeof(double) * s->p);
    if (f_out) memcpy(f_out, s->f, sizeof(double) * s->N);
    if (J_out) memcpy(J_out, s->J, sizeof(double) * s->N * s->p);
    if (iterations) *iterations = it;
    if (lm_status) *lm_status = status;
    if (accepted) *accepted = s->accepted;
    const int trials = s->trials;
    free(s);
    return trials;
}

/* ------------------------------------------------------------------------------ */
/* 2x2 eigen-decomposition (source/Localization.cpp:525-547)                        */
/* ------------------------------------------------------------------------------ */

int orc_eigen2(double sx, double sy, double rho, double *sd1, double *sd2, double *theta) {
    const double a = sx * sx, c = sy * sy, b = rho * sx * sy;
    if (!isfinite(a) || !isfinite(b) || !isfinite(c)) return -1;
    const double half_diff = 0.5 * (a - c);
    const double mean = 0.5 * (a + c);
    const double rad = hypot(half_diff, b);
    *sd1 = sqrt(mean - rad);
    *sd2 = sqrt(mean + rad);
    if (b == 0.0) {
        /* axis-aligned: the first eigenvector is a unit axis, so v0/v1 is 0 or v1 is 0 */
        *theta = 0.0;
    } else {
        /* first eigenvector (b, lambda1 - a) ~ (lambda1 - c, b); v0/v1 in the
         * cancellation-free form for either sign of (a - c) */
        const double ratio = half_diff >= 0.0 ? -b / (half_diff + rad) : (half_diff - rad) / b;
        *theta = atan(ratio);
    }
    return 0;
}

/* ------------------------------------------------------------------------------ */
/* Per-spot drivers (source/Localization.cpp:43-220, 391-631)                       */
/* ------------------------------------------------------------------------------ */

static double gsl_max(double a, double b) { return a > b ? a : b; } /* GSL_MAX_DBL */

/* One candidate's fit; the frame is uint16 counts (frame16) or FP64 values (frame64), the
 * reference's Image (source/Localization.cpp:126-130 reads it as double). */
static void fit_roi(const OrcSettings *s, const uint16_t *frame16, const double *frame64, int nx, int ny,
                    const LzrCandidate *cand, LzrFit *out) {
    memset(out, 0, sizeof *out);
    const int model = cand->model >= 0 ? cand->model : s->model;
    out->model = (int16_t)model;
    const double w = s->initial_psf_width;
    int h = cand->half_width > 0 ? cand->half_width
                                 : (s->default_half_width > 0 ? s->default_half_width
                                                              : (int)ceil(w * 4.0));
    if (h < 1 || h > LZR_MAX_HALF_WIDTH ||
        (model != LZR_MODEL_SYMMETRIC && model != LZR_MODEL_ELLIPSOIDAL &&
         model != LZR_MODEL_ELLIPSOIDAL_SYMMETRIC_PSF && model != LZR_MODEL_FIXED_WIDTH)) {
        out->status = LZR_REJECT_INVALID;
        return;
    }
    const int n = 2 * h + 1;
    const int xoff = (int)floor(cand->x - (double)h);
    const int yoff = (int)floor(cand->y - (double)h);
    const int xmax = xoff + n - 1, ymax = yoff + n - 1;
    if (xoff < 0 || xmax > nx - 1 || yoff < 0 || ymax > ny - 1) {
        out->status = LZR_REJECT_EDGE;
        return;
    }
    const int sym = model == LZR_MODEL_SYMMETRIC;
    const int fixed = model == LZR_MODEL_FIXED_WIDTH;
    const int p = sym ? 5 : (fixed ? 4 : 7);
    double data[ORC_MAXN];
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const long k = (long)(yoff + j) * nx + xoff + i;
            data[i + n * j] = frame64 ? frame64[k] : (double)frame16[k];
        }

    double t0[ORC_MAXP];
    if (fixed) {
        t0[0] = cand->intensity;
        t0[1] = cand->x;
        t0[2] = cand->y;
        t0[3] = cand->background;
    } else if (sym) {
        t0[0] = cand->intensity;
        t0[1] = w;
        t0[2] = cand->x;
        t0[3] = cand->y;
        t0[4] = cand->background;
    } else {
        t0[0] = cand->intensity;
        t0[1] = w;
        t0[2] = w;
        t0[3] = cand->x;
        t0[4] = cand->y;
        t0[5] = 0.0;
        t0[6] = cand->background;
    }
    Lm *lm = (Lm *)calloc(1, sizeof(Lm));
    lm->model = sym ? LZR_MODEL_SYMMETRIC : (fixed ? LZR_MODEL_FIXED_WIDTH : LZR_MODEL_ELLIPSOIDAL);
    lm->fixed_width = w;
    lm->p = p;
    lm->n = n;
    lm->N = n * n;
    lm->data = data;
    lm->xoff = xoff;
    lm->yoff = yoff;
    lm->sigma = s->sigma;
    lm->opt = s->opt;
    memcpy(lm->x, t0, sizeof(double) * p);
    int status = lm_set(lm);
    int it = 0;
    if (status == ST_OK) {
        do {
            it++;
            status = lm_iterate(lm);
            if (status != ST_OK) break;
            status = orc_test_delta(p, lm->dx, lm->x, s->opt.epsabs, s->opt.epsrel);
        } while (status == ST_CONTINUE && it < s->opt.max_iterations);
    }
    out->iterations = (int16_t)it;
    out->lm_status = status;
    out->trials = (int16_t)lm->trials;
    out->accepted_steps = (int16_t)lm->accepted;
    double *x = lm->x;
    out->chi = orc_nrm2(lm->N, lm->f, 1);

    if (fixed) {
        /* source/Localization.cpp:336-378 */
        memcpy(out->params, x, sizeof(double) * p);
        if (x[0] <= 0) {
            out->status = LZR_REJECT_AMPLITUDE;
        } else if (x[1] < xoff || x[1] > xmax || x[2] < yoff || x[2] > ymax) {
            out->status = LZR_REJECT_POSITION;
        } else if ((fabs(x[1] - cand->x) > 2.0 * w) || (fabs(x[2] - cand->y) > 2.0 * w)) {
            out->status = LZR_REJECT_DISTANCE;
        } else {
            double C[ORC_MAXP * ORC_MAXP];
            orc_covar(lm->N, p, lm->J, C);
            const double dof = (double)((2 * h - 1) * (2 * h - 1) - 4);
            const double c = gsl_max(1.0, out->chi / sqrt(dof));
            for (int k = 0; k < p; ++k) out->deviations[k] = c * sqrt(C[k * p + k]);
            out->integral = 2 * ORC_PI * w * w * x[0];
            const double relA = c * sqrt(C[0]) / x[0];
            out->integral_deviation = 2 * ORC_PI * out->integral * w * w * sqrt(relA * relA);
            out->sd1 = out->sd2 = w;
            out->status = LZR_ACCEPTED;
        }
    } else if (sym) {
        x[1] = fabs(x[1]);
        memcpy(out->params, x, sizeof(double) * p);
        if (x[0] <= 0) {
            out->status = LZR_REJECT_AMPLITUDE;
        } else if (x[2] < xoff || x[2] > xmax || x[3] < yoff || x[3] > ymax) {
            out->status = LZR_REJECT_POSITION;
        } else if (x[1] < 0.4 * w || x[1] > w * 1.5) {
            out->status = LZR_REJECT_WIDTH;
        } else {
            double C[ORC_MAXP * ORC_MAXP];
            orc_covar(lm->N, p, lm->J, C);
            const double dof = (double)((2 * h - 1) * (2 * h - 1) - 5);
            const double c = gsl_max(1.0, out->chi / sqrt(dof));
            for (int k = 0; k < p; ++k) out->deviations[k] = c * sqrt(C[k * p + k]);
            const double width = x[1];
            out->integral = 2 * ORC_PI * width * width * x[0];
            const double relA = c * sqrt(C[0]) / x[0];
            const double relW = out->deviations[1] / width;
            out->integral_deviation = 2 * ORC_PI * out->integral * sqrt(relA * relA + 2 * relW * relW);
            out->sd1 = out->sd2 = width;
            out->sd1_deviation = out->sd2_deviation = out->deviations[1];
            out->theta = 0.0;
            out->status = LZR_ACCEPTED;
        }
    } else {
        x[1] = fabs(x[1]);
        x[2] = fabs(x[2]);
        memcpy(out->params, x, sizeof(double) * p);
        double sd1 = 0, sd2 = 0, theta = 0;
        if (x[0] <= 0) {
            out->status = LZR_REJECT_AMPLITUDE;
        } else if (x[3] < xoff || x[3] > xmax || x[4] < yoff || x[4] > ymax) {
            out->status = LZR_REJECT_POSITION;
        } else if (orc_eigen2(x[1], x[2], x[5], &sd1, &sd2, &theta) != 0) {
            out->status = LZR_REJECT_EIGEN;
        } else if (!(((sd1 >= 0.4 * w) && (sd1 <= 1.5 * w)) || ((sd2 >= 0.4 * w) && (sd2 <= 1.5 * w)))) {
            out->status = LZR_REJECT_WIDTH;
        } else {
            double C[ORC_MAXP * ORC_MAXP];
            orc_covar(lm->N, p, lm->J, C);
            const double dof = (double)((2 * h - 1) * (2 * h - 1) - 7);
            const double c = gsl_max(1.0, out->chi / sqrt(dof));
            for (int k = 0; k < p; ++k) out->deviations[k] = c * sqrt(C[k * p + k]);
            const double rho = x[5];
            const double corr_dev = out->deviations[5];
            const double sx_dev = out->deviations[1], sy_dev = out->deviations[2];
            out->sd1 = sd1;
            out->sd2 = sd2;
            out->theta = theta;
            out->sd1_deviation = sx_dev * cos(theta) + sy_dev * sin(theta);
            out->sd2_deviation = sx_dev * sin(theta) + sy_dev * cos(theta);
            const double omr = 1 - rho * rho;
            out->integral = 2.0 * ORC_PI * x[0] * sqrt(omr) * sd1 * sd2;
            const double relA = out->deviations[0] / x[0];
            out->integral_deviation =
                2.0 * ORC_PI * x[0] * sqrt(omr) * sd1 * sd2 *
                sqrt(relA * relA + rho * rho * corr_dev * corr_dev / omr / omr +
                     out->sd1_deviation * out->sd1_deviation / sd1 / sd1 +
                     out->sd2_deviation * out->sd2_deviation / sd2 / sd2);
            out->status = LZR_ACCEPTED;
            if (model == LZR_MODEL_ELLIPSOIDAL_SYMMETRIC_PSF) {
                if ((0.75 * sd1 > sd2) || (1.25 * sd1 < sd2))
                    out->status = LZR_REJECT_ANISOTROPIC;
                else if (((sd1 < w * 0.7) || (sd1 > w * 1.3)) && ((sd2 < w * 0.7) || (sd2 > w * 1.3)))
                    out->status = LZR_REJECT_PSF_WIDTH;
            }
        }
    }
    free(lm);
}

void orc_fit_roi(const OrcSettings *s, const uint16_t *frame, int nx, int ny, const LzrCandidate *cand,
                 LzrFit *out) {
    fit_roi(s, frame, NULL, nx, ny, cand, out);
}

/* ------------------------------------------------------------------------------ */
/* Frame-parallel driver (PALMAnalysisController, source/PALMAnalysis.cpp:79-280)   */
/* ------------------------------------------------------------------------------ */

typedef struct Pool {
    const OrcSettings *s;
    const uint16_t *frames;
    const double *frames64; /* FP64 frames instead of frames (NULL: uint16) */
    int nx, ny, nframes;
    const LzrCandidate *cands;
    const int *order;        /* candidate indices grouped by frame */
    const int *group_start;  /* ngroups+1 offsets into order      */
    int ngroups;
    int next_group;          /* shared work counter (the reference's frame counter) */
    LzrFit *out;
} Pool;

static void *pool_worker(void *arg) {
    Pool *P = (Pool *)arg;
    for (;;) {
        const int g = __atomic_fetch_add(&P->next_group, 1, __ATOMIC_RELAXED);
        if (g >= P->ngroups) break;
        for (int q = P->group_start[g]; q < P->group_start[g + 1]; ++q) {
            const int ci = P->order[q];
            const LzrCandidate *c = &P->cands[ci];
            if (c->frame < 0 || c->frame >= P->nframes) {
                memset(&P->out[ci], 0, sizeof(LzrFit));
                P->out[ci].status = LZR_REJECT_INVALID;
                continue;
            }
            const long base = (long)c->frame * P->nx * P->ny;
            fit_roi(P->s, P->frames64 ? NULL : P->frames + base, P->frames64 ? P->frames64 + base : NULL, P->nx,
                    P->ny, c, &P->out[ci]);
        }
    }
    return NULL;
}

static int cmp_frame_idx(const void *a, const void *b, void *ctx) {
    const LzrCandidate *c = (const LzrCandidate *)ctx;
    const int ia = *(const int *)a, ib = *(const int *)b;
    if (c[ia].frame != c[ib].frame) return c[ia].frame < c[ib].frame ? -1 : 1;
    return ia < ib ? -1 : (ia > ib);
}

static int fit_batch(const OrcSettings *s, const uint16_t *frames, const double *frames64, int nx, int ny,
                     int nframes, const LzrCandidate *cands, int ncands, LzrFit *out, int nthreads) {
    if (ncands <= 0) return 0;
    int *order = (int *)malloc(sizeof(int) * ncands);
    int *starts = (int *)malloc(sizeof(int) * (ncands + 1));
    for (int i = 0; i < ncands; ++i) order[i] = i;
    qsort_r(order, ncands, sizeof(int), cmp_frame_idx, (void *)cands);
    int ng = 0;
    for (int i = 0; i < ncands; ++i)
        if (i == 0 || cands[order[i]].frame != cands[order[i - 1]].frame) starts[ng++] = i;
    starts[ng] = ncands;
    Pool P = {s, frames, frames64, nx, ny, nframes, cands, order, starts, ng, 0, out};
    if (nthreads <= 0) nthreads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > ng) nthreads = ng;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
    for (int t = 1; t < nthreads; ++t) pthread_create(&th[t], NULL, pool_worker, &P);
    pool_worker(&P);
    for (int t = 1; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(th);
    free(order);
    free(starts);
    int acc = 0;
    for (int i = 0; i < ncands; ++i) acc += out[i].status == LZR_ACCEPTED;
    return acc;
}

int orc_fit_batch(const OrcSettings *s, const uint16_t *frames, int nx, int ny, int nframes,
                  const LzrCandidate *cands, int ncands, LzrFit *out, int nthreads) {
    return fit_batch(s, frames, NULL, nx, ny, nframes, cands, ncands, out, nthreads);
}

int orc_fit_batch_f64(const OrcSettings *s, const double *frames, int nx, int ny, int nframes,
                      const LzrCandidate *cands, int ncands, LzrFit *out, int nthreads) {
    return fit_batch(s, NULL, frames, nx, ny, nframes, cands, ncands, out, nthreads);
}
