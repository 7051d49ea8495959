/*
 * xdn_oracle.c -- FP64 host restatement of cross-denoiser's per-frame filtering path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the oracle results are checked against: only tests/,
 * the smoke check of __graft_entry__ and the cpu_baseline leg of bench.py load it. The product
 * library (libxdn.so) never links, calls or falls back to it.
 *
 * Every value path is computed in double precision from the FP32 inputs, following the kernel
 * bodies of the reference (file:line cited per function, paths relative to the reference root).
 * Decision paths whose result is a mask or an index (reprojection validity, temporal screening)
 * are computed in exact FP32 with a pinned operation order, because those must be reproduced bit
 * for bit by the device kernels:
 *   - posPrev = fmaf(mvec, (float)dim, (float)pos)          (reproject.cu.cc:97,103)
 *   - dot(a-b,a-b) = fmaf(dz,dz, fmaf(dx,dx, dy*dy))          (cuda_utils.cuh:260-273)
 *   - the normal test divides in double (the 1e-2 literal is a double, reproject.cu.cc:46)
 *   - float->int conversion truncates toward zero and saturates, NaN -> 0 (cuda_utils.cuh:255-258)
 * Build with -ffp-contract=off and without -ffast-math (see oracle/Makefile).
 *
 * Layouts: every image is one frame, NHWC with the batch dimension removed, row-major
 * [H][W][C] float32. Outputs are double unless stated.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>
#include <omp.h>

#define P6 6 /* COEFFI_DIM_PILOT, regression.cu.cc:30 */

/* reflect-101 boundary used by the regression and the spatial filter
 * (regression.cu.cc:128-129, 413-414). Valid for -n < x < 2n-1. */
static inline int mirror_idx(int x, int n) { return (x >= n) ? 2 * n - 2 - x : abs(x); }

/* iDivUp with C integer division (truncation toward zero), cuda_utils.cuh:275-278.
 * For negative a this is NOT a ceiling: iDivUp(-3, 4) == 1. BlockRecon inherits that. */
static inline int idivup(int a, int b) { return ((a % b) != 0) ? (a / b + 1) : (a / b); }

/* float -> int as the device converts it: truncate toward zero, saturate, NaN -> 0. */
static inline int f2i_trunc(float v)
{
    if (v != v) return 0;
    if (v >= 2147483648.0f) return INT_MAX;
    if (v <= -2147483648.0f) return INT_MIN;
    return (int)v;
}

static inline double sq(double a) { return a * a; }

/* ------------------------------------------------------------------------------------------- */
/* a2: input variance, EstInputVarianceKernel regression.cu.cc:244-276 (winSize 3, :490-491).  */
/* var(p) = || mean_{q in 3x3(p) clamped, q != p} y_q - y_p ||^2                                */
/* ------------------------------------------------------------------------------------------- */
void orc_variance(const float *y, double *var, int H, int W)
{
#pragma omp parallel for schedule(static)
    for (int cy = 0; cy < H; ++cy)
        for (int cx = 0; cx < W; ++cx)
        {
            int sx = cx - 1 < 0 ? 0 : cx - 1, ex = cx + 1 > W - 1 ? W - 1 : cx + 1;
            int sy = cy - 1 < 0 ? 0 : cy - 1, ey = cy + 1 > H - 1 ? H - 1 : cy + 1;
            double acc[3] = {0, 0, 0};
            int n = 0;
            for (int iy = sy; iy <= ey; ++iy)
                for (int ix = sx; ix <= ex; ++ix)
                {
                    if (iy == cy && ix == cx) continue;
                    const float *q = y + 3 * ((size_t)iy * W + ix);
                    acc[0] += q[0]; acc[1] += q[1]; acc[2] += q[2];
                    ++n;
                }
            const float *c = y + 3 * ((size_t)cy * W + cx);
            double v = 0;
            for (int k = 0; k < 3; ++k) v += sq(acc[k] / n - c[k]);
            var[(size_t)cy * W + cx] = v;
        }
}

/* regressionWgt, regression.cu.cc:278-287 (+1e-7 added at the call sites :356, :432). */
static inline double reg_weight(const double *cw, const double *iw, double cvar, double ivar)
{
    double d = sq(iw[0] - cw[0]) + sq(iw[1] - cw[1]) + sq(iw[2] - cw[2]);
    return exp(-d / ((cvar + ivar) + 1e-4)) + 1e-7;
}

/* ------------------------------------------------------------------------------------------- */
/* a3: per-center weighted ridge regression, RegressionKernel regression.cu.cc:367-475,         */
/* cholesky_color :206-218 (SQRT clamp :197-204), forward_backward_solver :220-242.             */
/* beta layout: [H/s][W/s][6][3] (coefficient-major, channel-minor).                            */
/* target = the other half (XtB), feat/var = the own half (weights and delta[1]).              */
/* ------------------------------------------------------------------------------------------- */
static void solve_center(const double A[3][P6][P6], const double XtB[3][P6], double beta[P6][3])
{
    for (int ch = 0; ch < 3; ++ch)
    {
        double M[P6][P6], L[P6][P6], Wv[P6], b[P6];
        for (int i = 0; i < P6; ++i)
            for (int j = 0; j < P6; ++j) M[i][j] = A[ch][i][j];
        for (int i = 0; i < P6; ++i) M[i][i] += 1e-3; /* ridge, :458-461 */
        for (int i = 0; i < P6; ++i)
            for (int j = 0; j <= i; ++j)
            {
                double s = 0;
                for (int k = 0; k < j; ++k) s += L[i][k] * L[j][k];
                if (i == j)
                {
                    double v = M[i][i] - s;
                    L[i][i] = sqrt(v > 1e-5 ? v : 1e-5);
                }
                else
                    L[i][j] = (1.0 / L[j][j]) * (M[j][i] - s);
            }
        Wv[0] = XtB[ch][0] / L[0][0];
        for (int i = 1; i < P6; ++i)
        {
            double s = 0;
            for (int k = 0; k < i; ++k) s += L[i][k] * Wv[k];
            Wv[i] = (XtB[ch][i] - s) / L[i][i];
        }
        b[P6 - 1] = Wv[P6 - 1] / L[P6 - 1][P6 - 1];
        for (int i = P6 - 2; i >= 0; --i)
        {
            double s = 0;
            for (int k = P6 - 1; k > i; --k) s += L[k][i] * b[k];
            b[i] = (Wv[i] - s) / L[i][i];
        }
        for (int i = 0; i < P6; ++i) beta[i][ch] = b[i];
    }
}

static inline void load3(const float *img, size_t idx, double *o)
{
    o[0] = img[3 * idx]; o[1] = img[3 * idx + 1]; o[2] = img[3 * idx + 2];
}

void orc_regression(const float *target, const float *feat, const double *var, const float *alb,
                    const float *nrm, double *beta, int H, int W, int win, int s)
{
    const int half = win / 2, Hs = H / s, Ws = W / s;
#pragma omp parallel for schedule(dynamic, 1)
    for (int by = 0; by < Hs; ++by)
        for (int bx = 0; bx < Ws; ++bx)
        {
            const int cx = bx * s, cy = by * s;
            const size_t c = (size_t)cy * W + cx;
            double cw[3], ca[3], cn[3];
            load3(feat, c, cw); load3(alb, c, ca); load3(nrm, c, cn);
            const double cvar = var[c], cstd = sqrt(cvar);
            double A[3][P6][P6], XtB[3][P6];
            memset(A, 0, sizeof(A));
            memset(XtB, 0, sizeof(XtB));
            for (int iy = cy - half; iy <= cy + half; ++iy)
                for (int ix = cx - half; ix <= cx + half; ++ix)
                {
                    const size_t i = (size_t)mirror_idx(iy, H) * W + mirror_idx(ix, W);
                    double iw[3], ia[3], in[3], it[3];
                    load3(feat, i, iw); load3(alb, i, ia); load3(nrm, i, in); load3(target, i, it);
                    const double ivar = var[i], istd = sqrt(ivar);
                    const double w = reg_weight(cw, iw, cvar, ivar);
                    const double den = (cstd + istd) + 1e-2;
                    for (int ch = 0; ch < 3; ++ch)
                    {
                        const double d[P6] = {1.0, (iw[ch] - cw[ch]) / den, ia[ch] - ca[ch],
                                              in[0] - cn[0], in[1] - cn[1], in[2] - cn[2]};
                        for (int r = 0; r < P6; ++r)
                        {
                            for (int q = r; q < P6; ++q) A[ch][r][q] += w * d[r] * d[q];
                            XtB[ch][r] += w * d[r] * it[ch];
                        }
                    }
                }
            for (int ch = 0; ch < 3; ++ch)
                for (int r = 0; r < P6; ++r)
                    for (int q = 0; q < r; ++q) A[ch][r][q] = A[ch][q][r];
            double b[P6][3];
            solve_center((const double(*)[P6][P6])A, (const double(*)[P6])XtB, b);
            memcpy(beta + ((size_t)by * Ws + bx) * P6 * 3, b, sizeof(b));
        }
}

/* ------------------------------------------------------------------------------------------- */
/* a4: block reconstruction, BlockRecon regression.cu.cc:289-365.                               */
/* Centers within +-half of the pixel, clamped to the center grid with iDivUp (:312-315);       */
/* delta uses the reversed sign (pixel - center, :342-348); out = max(0, sum/max(sumW,1e-7)).  */
/* ------------------------------------------------------------------------------------------- */
void orc_recon(const double *beta, const float *feat, const double *var, const float *alb,
               const float *nrm, double *out, int H, int W, int win, int s, int rs)
{
    const int half = win / 2, Ws = W / s, Hr = H / rs, Wr = W / rs;
#pragma omp parallel for schedule(static)
    for (int ry = 0; ry < Hr; ++ry)
        for (int rx = 0; rx < Wr; ++rx)
        {
            const int cx = rx * rs, cy = ry * rs;
            const size_t c = (size_t)cy * W + cx;
            int sx = idivup(cx - half, s); if (sx < 0) sx = 0;
            int ex = (cx + half) / s; if (ex > (W - 1) / s) ex = (W - 1) / s;
            int sy = idivup(cy - half, s); if (sy < 0) sy = 0;
            int ey = (cy + half) / s; if (ey > (H - 1) / s) ey = (H - 1) / s;
            double cw[3], ca[3], cn[3];
            load3(feat, c, cw); load3(alb, c, ca); load3(nrm, c, cn);
            const double cvar = var[c], cstd = sqrt(cvar);
            double acc[3] = {0, 0, 0}, sumw = 0;
            for (int iy = sy; iy <= ey; ++iy)
                for (int ix = sx; ix <= ex; ++ix)
                {
                    const double *b = beta + ((size_t)iy * Ws + ix) * P6 * 3;
                    const size_t i = (size_t)(iy * s) * W + (size_t)(ix * s);
                    double iw[3], ia[3], in[3];
                    load3(feat, i, iw); load3(alb, i, ia); load3(nrm, i, in);
                    const double ivar = var[i], istd = sqrt(ivar);
                    const double den = (cstd + istd) + 1e-2;
                    const double w = reg_weight(cw, iw, cvar, ivar);
                    for (int ch = 0; ch < 3; ++ch)
                    {
                        const double d[P6] = {1.0, (cw[ch] - iw[ch]) / den, ca[ch] - ia[ch],
                                              cn[0] - in[0], cn[1] - in[1], cn[2] - in[2]};
                        double pred = 0;
                        for (int k = 0; k < P6; ++k) pred += b[k * 3 + ch] * d[k];
                        acc[ch] += w * pred;
                    }
                    sumw += w;
                }
            const double den = sumw > 1e-7 ? sumw : 1e-7;
            const size_t o = (size_t)ry * Wr + rx;
            for (int ch = 0; ch < 3; ++ch)
            {
                const double v = acc[ch] / den;
                out[3 * o + ch] = v > 0 ? v : 0;
            }
        }
}

/* a5: the cross_regression op, CrossRegressionFunc regression.cu.cc:477-500.
 * outa from beta_A (target imgb, features imga, varA); outb symmetric. beta_* are caller
 * scratch of (H/s)*(W/s)*18 doubles each. */
void orc_cross_regression(const float *ya, const float *yb, const float *alb, const float *nrm,
                          double *outa, double *outb, double *vara, double *varb,
                          double *beta_a, double *beta_b, int H, int W, int win, int s, int rs)
{
    orc_variance(ya, vara, H, W);
    orc_variance(yb, varb, H, W);
    orc_regression(yb, ya, vara, alb, nrm, beta_a, H, W, win, s);
    orc_regression(ya, yb, varb, alb, nrm, beta_b, H, W, win, s);
    orc_recon(beta_a, ya, vara, alb, nrm, outa, H, W, win, s, rs);
    orc_recon(beta_b, yb, varb, alb, nrm, outb, H, W, win, s, rs);
}

/* ------------------------------------------------------------------------------------------- */
/* Backward of cross_regression with respect to its four inputs (recon_scale 1).               */
/* The reference registers no gradient for this op; this is the adjoint of the forward above,   */
/* differentiated in FP64 (scatter form, serial). Conventions:                                  */
/*   - beta = M^-1 b with M = sum w d d^T + 1e-3 I; the SQRT clamp of the Cholesky never binds   */
/*     (every pivot of M is >= 1e-3 > 1e-5), so the solve is differentiated as a linear solve:  */
/*     u = M^-1 g_beta, g_b = u, g_M = -u beta^T.                                               */
/*   - max(0, .) of the reconstruction passes the gradient only where the value is > 0; the     */
/*     max(sumW, 1e-7) guard passes it only where sumW > 1e-7.                                  */
/*   - d sqrt(var)/d var is taken as 0 where var == 0.                                         */
/* ------------------------------------------------------------------------------------------- */
typedef struct
{
    const float *f, *a, *n; /* own half (features), albedo, normal */
    const double *v;        /* variance of the own half */
    double *gf, *ga, *gn, *gv;
} CrgHalf;

/* Adjoint of (w, delta) for the pair (P, Q), delta = x_P - x_Q per channel:
 * w = exp(-|f_P - f_Q|^2 / (v_P + v_Q + 1e-4)) + 1e-7, delta_1 = (f_P - f_Q)/(sqrt v_P + sqrt v_Q + 1e-2),
 * delta_2 = a_P - a_Q, delta_3..5 = n_P - n_Q. */
static void crg_pair(const CrgHalf *h, size_t P, size_t Q, double gw, const double gd[3][P6])
{
    double fP[3], fQ[3];
    load3(h->f, P, fP); load3(h->f, Q, fQ);
    const double vP = h->v[P], vQ = h->v[Q];
    double D = 0;
    for (int k = 0; k < 3; ++k) D += sq(fP[k] - fQ[k]);
    const double S = vP + vQ + 1e-4, E = exp(-D / S);
    const double gD = -gw * E / S, gS = gw * E * D / (S * S);
    const double sP = sqrt(vP), sQ = sqrt(vQ), den = sP + sQ + 1e-2;
    double gden = 0;
    for (int k = 0; k < 3; ++k)
    {
        const double df = fP[k] - fQ[k];
        const double gfk = gD * 2.0 * df + gd[k][1] / den;
        h->gf[3 * P + k] += gfk;
        h->gf[3 * Q + k] -= gfk;
        gden -= gd[k][1] * (df / den) / den;
        h->ga[3 * P + k] += gd[k][2];
        h->ga[3 * Q + k] -= gd[k][2];
        for (int j = 0; j < 3; ++j)
        {
            h->gn[3 * P + j] += gd[k][3 + j];
            h->gn[3 * Q + j] -= gd[k][3 + j];
        }
    }
    h->gv[P] += gS + (sP > 0 ? gden * 0.5 / sP : 0.0);
    h->gv[Q] += gS + (sQ > 0 ? gden * 0.5 / sQ : 0.0);
}

/* One half: features f (with its variance v), target t, upstream gradient gout of the
 * reconstruction and gvar_ext of the variance output. Accumulates into gf (features), gt (target),
 * ga, gn. */
static void crg_half(const float *f, const float *t, const float *alb, const float *nrm, const float *gout,
                     const float *gvar_ext, double *gf, double *gt, double *ga, double *gn, int H, int W,
                     int win, int s)
{
    const int half = win / 2, Hs = H / s, Ws = W / s;
    const size_t N = (size_t)H * W;
    double *v = (double *)malloc(N * sizeof(double));
    double *beta = (double *)malloc((size_t)Hs * Ws * P6 * 3 * sizeof(double));
    double *gbeta = (double *)calloc((size_t)Hs * Ws * P6 * 3, sizeof(double));
    double *gv = (double *)calloc(N, sizeof(double));
    orc_variance(f, v, H, W);
    orc_regression(t, f, v, alb, nrm, beta, H, W, win, s);
    CrgHalf h = {f, alb, nrm, v, gf, ga, gn, gv};

    /* reconstruction (orc_recon), adjoint per pixel */
    for (int py = 0; py < H; ++py)
        for (int px = 0; px < W; ++px)
        {
            const size_t p = (size_t)py * W + px;
            int sx = idivup(px - half, s); if (sx < 0) sx = 0;
            int ex = (px + half) / s; if (ex > (W - 1) / s) ex = (W - 1) / s;
            int sy = idivup(py - half, s); if (sy < 0) sy = 0;
            int ey = (py + half) / s; if (ey > (H - 1) / s) ey = (H - 1) / s;
            double cw[3], ca[3], cn[3];
            load3(f, p, cw); load3(alb, p, ca); load3(nrm, p, cn);
            const double cstd = sqrt(v[p]);
            double acc[3] = {0, 0, 0}, sumw = 0;
            for (int pass = 0; pass < 2; ++pass)
            {
                double gacc[3] = {0, 0, 0}, gsumw = 0;
                if (pass == 1)
                {
                    const double dd = sumw > 1e-7 ? sumw : 1e-7;
                    for (int ch = 0; ch < 3; ++ch)
                        gacc[ch] = acc[ch] / dd > 0 ? gout[3 * p + ch] / dd : 0.0;
                    if (sumw > 1e-7)
                        for (int ch = 0; ch < 3; ++ch) gsumw -= gacc[ch] * acc[ch] / dd;
                }
                for (int iy = sy; iy <= ey; ++iy)
                    for (int ix = sx; ix <= ex; ++ix)
                    {
                        const double *b = beta + ((size_t)iy * Ws + ix) * P6 * 3;
                        const size_t c = (size_t)(iy * s) * W + (size_t)(ix * s);
                        double iw[3], ia[3], in[3];
                        load3(f, c, iw); load3(alb, c, ia); load3(nrm, c, in);
                        const double den = (cstd + sqrt(v[c])) + 1e-2;
                        const double w = reg_weight(cw, iw, v[p], v[c]);
                        double pred[3], d[3][P6];
                        for (int ch = 0; ch < 3; ++ch)
                        {
                            d[ch][0] = 1.0; d[ch][1] = (cw[ch] - iw[ch]) / den; d[ch][2] = ca[ch] - ia[ch];
                            d[ch][3] = cn[0] - in[0]; d[ch][4] = cn[1] - in[1]; d[ch][5] = cn[2] - in[2];
                            pred[ch] = 0;
                            for (int k = 0; k < P6; ++k) pred[ch] += b[k * 3 + ch] * d[ch][k];
                        }
                        if (pass == 0)
                        {
                            for (int ch = 0; ch < 3; ++ch) acc[ch] += w * pred[ch];
                            sumw += w;
                            continue;
                        }
                        double gw = gsumw, gd[3][P6];
                        double *gb = gbeta + ((size_t)iy * Ws + ix) * P6 * 3;
                        for (int ch = 0; ch < 3; ++ch)
                        {
                            gw += gacc[ch] * pred[ch];
                            for (int k = 0; k < P6; ++k)
                            {
                                gb[k * 3 + ch] += w * gacc[ch] * d[ch][k];
                                gd[ch][k] = w * gacc[ch] * b[k * 3 + ch];
                            }
                        }
                        crg_pair(&h, p, c, gw, (const double(*)[P6])gd);
                    }
            }
        }

    /* regression (orc_regression), adjoint per center */
    for (int by = 0; by < Hs; ++by)
        for (int bx = 0; bx < Ws; ++bx)
        {
            const int cx = bx * s, cy = by * s;
            const size_t c = (size_t)cy * W + cx;
            const double *b = beta + ((size_t)by * Ws + bx) * P6 * 3;
            const double *gb = gbeta + ((size_t)by * Ws + bx) * P6 * 3;
            double cw[3], ca[3], cn[3];
            load3(f, c, cw); load3(alb, c, ca); load3(nrm, c, cn);
            const double cstd = sqrt(v[c]);
            double A[3][P6][P6], G[3][P6], u[P6][3];
            memset(A, 0, sizeof(A));
            for (int pass = 0; pass < 2; ++pass)
            {
                if (pass == 1)
                {
                    for (int ch = 0; ch < 3; ++ch)
                        for (int r = 0; r < P6; ++r) G[ch][r] = gb[r * 3 + ch];
                    solve_center((const double(*)[P6][P6])A, (const double(*)[P6])G, u);
                }
                for (int iy = cy - half; iy <= cy + half; ++iy)
                    for (int ix = cx - half; ix <= cx + half; ++ix)
                    {
                        const size_t i = (size_t)mirror_idx(iy, H) * W + mirror_idx(ix, W);
                        double iw[3], ia[3], in[3], it[3];
                        load3(f, i, iw); load3(alb, i, ia); load3(nrm, i, in); load3(t, i, it);
                        const double w = reg_weight(cw, iw, v[c], v[i]);
                        const double den = (cstd + sqrt(v[i])) + 1e-2;
                        double d[3][P6];
                        for (int ch = 0; ch < 3; ++ch)
                        {
                            d[ch][0] = 1.0; d[ch][1] = (iw[ch] - cw[ch]) / den; d[ch][2] = ia[ch] - ca[ch];
                            d[ch][3] = in[0] - cn[0]; d[ch][4] = in[1] - cn[1]; d[ch][5] = in[2] - cn[2];
                        }
                        if (pass == 0)
                        {
                            for (int ch = 0; ch < 3; ++ch)
                                for (int r = 0; r < P6; ++r)
                                    for (int q = 0; q < P6; ++q) A[ch][r][q] += w * d[ch][r] * d[ch][q];
                            continue;
                        }
                        double gw = 0, gd[3][P6];
                        for (int ch = 0; ch < 3; ++ch)
                        {
                            double r = it[ch], sdot = 0;
                            for (int k = 0; k < P6; ++k)
                            {
                                r -= d[ch][k] * b[k * 3 + ch];
                                sdot += u[k][ch] * d[ch][k];
                            }
                            gw += sdot * r;
                            for (int k = 0; k < P6; ++k) gd[ch][k] = w * (u[k][ch] * r - b[k * 3 + ch] * sdot);
                            gt[3 * i + ch] += w * sdot;
                        }
                        crg_pair(&h, i, c, gw, (const double(*)[P6])gd);
                    }
            }
        }

    /* variance (orc_variance), adjoint */
    if (gvar_ext)
        for (size_t p = 0; p < N; ++p) gv[p] += gvar_ext[p];
    for (int cy = 0; cy < H; ++cy)
        for (int cx = 0; cx < W; ++cx)
        {
            int sx = cx - 1 < 0 ? 0 : cx - 1, ex = cx + 1 > W - 1 ? W - 1 : cx + 1;
            int sy = cy - 1 < 0 ? 0 : cy - 1, ey = cy + 1 > H - 1 ? H - 1 : cy + 1;
            const int n = (ex - sx + 1) * (ey - sy + 1) - 1;
            const size_t p = (size_t)cy * W + cx;
            double m[3] = {0, 0, 0};
            for (int iy = sy; iy <= ey; ++iy)
                for (int ix = sx; ix <= ex; ++ix)
                    if (iy != cy || ix != cx)
                        for (int k = 0; k < 3; ++k) m[k] += f[3 * ((size_t)iy * W + ix) + k];
            for (int k = 0; k < 3; ++k)
            {
                const double e = gv[p] * 2.0 * (m[k] / n - f[3 * p + k]);
                gf[3 * p + k] -= e;
                for (int iy = sy; iy <= ey; ++iy)
                    for (int ix = sx; ix <= ex; ++ix)
                        if (iy != cy || ix != cx) gf[3 * ((size_t)iy * W + ix) + k] += e / n;
            }
        }
    free(v); free(beta); free(gbeta); free(gv);
}

/* Gradients of L(outa, outb, vara, varb) with respect to (ya, yb, albedo, normal), given the
 * upstream gradients of the four outputs of orc_cross_regression (g_vara/g_varb may be NULL).
 * The four outputs are overwritten. */
void orc_cross_regression_grad(const float *ya, const float *yb, const float *alb, const float *nrm,
                               const float *g_outa, const float *g_outb, const float *g_vara,
                               const float *g_varb, double *g_ya, double *g_yb, double *g_alb, double *g_nrm,
                               int H, int W, int win, int s)
{
    const size_t n3 = (size_t)H * W * 3;
    memset(g_ya, 0, n3 * sizeof(double)); memset(g_yb, 0, n3 * sizeof(double));
    memset(g_alb, 0, n3 * sizeof(double)); memset(g_nrm, 0, n3 * sizeof(double));
    /* half A: features ya, target yb; half B: features yb, target ya */
    crg_half(ya, yb, alb, nrm, g_outa, g_vara, g_ya, g_yb, g_alb, g_nrm, H, W, win, s);
    crg_half(yb, ya, alb, nrm, g_outb, g_varb, g_yb, g_ya, g_alb, g_nrm, H, W, win, s);
}

/* ------------------------------------------------------------------------------------------- */
/* a6: bandwidth-driven bilateral filter, SpatialFilterKernel regression.cu.cc:84-195.          */
/* w^X = exp(-b0|dX|^2 - b2|dalb|^2 - b3|dn|^2 - b4*((dx/h)^2+(dy/h)^2)), X in {A (b0), B (b1)}*/
/* 11x11 reflect-101; outputs the unnormalized sum and the weight sum.                          */
/* When gband != NULL the band gradient moments G_k = sum w*(-dist_k)*(c,1) are also formed    */
/* (:157-177) and contracted with the upstream gradients as SpatialFilterGradKernel does        */
/* (:33-81). band layout [b_colA, b_colB, b_alb, b_nrm, b_pix].                                 */
/* ------------------------------------------------------------------------------------------- */
static void sf_pixel(const float *ia, const float *ib, const float *alb, const float *nrm,
                     const float *band, int H, int W, int win, int cx, int cy, double accA[3],
                     double accB[3], double *wsA, double *wsB, double GA[4][4], double GB[4][4])
{
    const int half = win / 2;
    const size_t c = (size_t)cy * W + cx;
    double cA[3], cB[3], ca[3], cn[3], b[5];
    load3(ia, c, cA); load3(ib, c, cB); load3(alb, c, ca); load3(nrm, c, cn);
    for (int k = 0; k < 5; ++k) b[k] = band[5 * c + k];
    accA[0] = accA[1] = accA[2] = accB[0] = accB[1] = accB[2] = 0;
    *wsA = *wsB = 0;
    if (GA) memset(GA, 0, 16 * sizeof(double)), memset(GB, 0, 16 * sizeof(double));
    for (int iy = cy - half; iy <= cy + half; ++iy)
        for (int ix = cx - half; ix <= cx + half; ++ix)
        {
            const size_t i = (size_t)mirror_idx(iy, H) * W + mirror_idx(ix, W);
            double iA[3], iB[3], iaa[3], in[3];
            load3(ia, i, iA); load3(ib, i, iB); load3(alb, i, iaa); load3(nrm, i, in);
            const double dy = (double)(cy - iy) / half, dx = (double)(cx - ix) / half;
            const double pd = dx * dx + dy * dy;
            const double dA = sq(cA[0] - iA[0]) + sq(cA[1] - iA[1]) + sq(cA[2] - iA[2]);
            const double dB = sq(cB[0] - iB[0]) + sq(cB[1] - iB[1]) + sq(cB[2] - iB[2]);
            const double dal = sq(ca[0] - iaa[0]) + sq(ca[1] - iaa[1]) + sq(ca[2] - iaa[2]);
            const double dn = sq(cn[0] - in[0]) + sq(cn[1] - in[1]) + sq(cn[2] - in[2]);
            const double wA = exp(-b[0] * dA - b[2] * dal - b[3] * dn - b[4] * pd);
            const double wB = exp(-b[1] * dB - b[2] * dal - b[3] * dn - b[4] * pd);
            for (int k = 0; k < 3; ++k) accA[k] += wA * iA[k], accB[k] += wB * iB[k];
            *wsA += wA;
            *wsB += wB;
            if (GA)
            {
                const double dwA[4] = {-wA * dA, -wA * dal, -wA * dn, -wA * pd};
                const double dwB[4] = {-wB * dB, -wB * dal, -wB * dn, -wB * pd};
                for (int k = 0; k < 4; ++k)
                {
                    for (int ch = 0; ch < 3; ++ch) GA[k][ch] += dwA[k] * iA[ch], GB[k][ch] += dwB[k] * iB[ch];
                    GA[k][3] += dwA[k];
                    GB[k][3] += dwB[k];
                }
            }
        }
}

void orc_spatial_filter(const float *ia, const float *ib, const float *alb, const float *nrm,
                        const float *band, double *accA, double *accB, double *wsA, double *wsB,
                        int H, int W, int win)
{
#pragma omp parallel for schedule(static)
    for (int cy = 0; cy < H; ++cy)
        for (int cx = 0; cx < W; ++cx)
        {
            const size_t c = (size_t)cy * W + cx;
            sf_pixel(ia, ib, alb, nrm, band, H, W, win, cx, cy, accA + 3 * c, accB + 3 * c, wsA + c,
                     wsB + c, NULL, NULL);
        }
}

/* a7: band gradient (SpatialFilterGradKernel regression.cu.cc:33-81 applied to the moments the
 * forward stores at :157-177; here recomputed, not stored). */
void orc_spatial_filter_grad(const float *gaccA, const float *gaccB, const float *gwA, const float *gwB,
                             const float *ia, const float *ib, const float *alb, const float *nrm,
                             const float *band, double *gband, int H, int W, int win)
{
#pragma omp parallel for schedule(static)
    for (int cy = 0; cy < H; ++cy)
        for (int cx = 0; cx < W; ++cx)
        {
            const size_t c = (size_t)cy * W + cx;
            double aA[3], aB[3], sA, sB, GA[4][4], GB[4][4];
            sf_pixel(ia, ib, alb, nrm, band, H, W, win, cx, cy, aA, aB, &sA, &sB, GA, GB);
            double oA[4], oB[4];
            for (int k = 0; k < 4; ++k)
            {
                oA[k] = gaccA[3 * c] * GA[k][0] + gaccA[3 * c + 1] * GA[k][1] + gaccA[3 * c + 2] * GA[k][2] + gwA[c] * GA[k][3];
                oB[k] = gaccB[3 * c] * GB[k][0] + gaccB[3 * c + 1] * GB[k][1] + gaccB[3 * c + 2] * GB[k][2] + gwB[c] * GB[k][3];
            }
            gband[5 * c + 0] = oA[0];
            gband[5 * c + 1] = oB[0];
            gband[5 * c + 2] = oA[1] + oB[1];
            gband[5 * c + 3] = oA[2] + oB[2];
            gband[5 * c + 4] = oA[3] + oB[3];
        }
}

/* Gradient of the spatial filter with respect to its image inputs (imga, imgb, albedo, normal).
 * The reference returns None for these (scripts/ops.py:23-32); this is the analytic adjoint of
 * regression.cu.cc:124-179, written in scatter form (serial: every tap updates two pixels). */
void orc_spatial_filter_grad_inputs(const float *gaccA, const float *gaccB, const float *gwA,
                                    const float *gwB, const float *ia, const float *ib,
                                    const float *alb, const float *nrm, const float *band,
                                    double *gia, double *gib, double *galb, double *gnrm, int H,
                                    int W, int win)
{
    const int half = win / 2;
    const size_t N = (size_t)H * W;
    memset(gia, 0, 3 * N * sizeof(double));
    memset(gib, 0, 3 * N * sizeof(double));
    memset(galb, 0, 3 * N * sizeof(double));
    memset(gnrm, 0, 3 * N * sizeof(double));
    for (int cy = 0; cy < H; ++cy)
        for (int cx = 0; cx < W; ++cx)
        {
            const size_t c = (size_t)cy * W + cx;
            double cA[3], cB[3], ca[3], cn[3], b[5], gA[3], gB[3];
            load3(ia, c, cA); load3(ib, c, cB); load3(alb, c, ca); load3(nrm, c, cn);
            for (int k = 0; k < 5; ++k) b[k] = band[5 * c + k];
            for (int k = 0; k < 3; ++k) gA[k] = gaccA[3 * c + k], gB[k] = gaccB[3 * c + k];
            const double hA = gwA[c], hB = gwB[c];
            for (int iy = cy - half; iy <= cy + half; ++iy)
                for (int ix = cx - half; ix <= cx + half; ++ix)
                {
                    const size_t i = (size_t)mirror_idx(iy, H) * W + mirror_idx(ix, W);
                    double iA[3], iB[3], iaa[3], in[3];
                    load3(ia, i, iA); load3(ib, i, iB); load3(alb, i, iaa); load3(nrm, i, in);
                    const double dy = (double)(cy - iy) / half, dx = (double)(cx - ix) / half;
                    const double pd = dx * dx + dy * dy;
                    const double dA = sq(cA[0] - iA[0]) + sq(cA[1] - iA[1]) + sq(cA[2] - iA[2]);
                    const double dB = sq(cB[0] - iB[0]) + sq(cB[1] - iB[1]) + sq(cB[2] - iB[2]);
                    const double dal = sq(ca[0] - iaa[0]) + sq(ca[1] - iaa[1]) + sq(ca[2] - iaa[2]);
                    const double dn = sq(cn[0] - in[0]) + sq(cn[1] - in[1]) + sq(cn[2] - in[2]);
                    const double wA = exp(-b[0] * dA - b[2] * dal - b[3] * dn - b[4] * pd);
                    const double wB = exp(-b[1] * dB - b[2] * dal - b[3] * dn - b[4] * pd);
                    /* dL/dw */
                    const double uA = wA * (gA[0] * iA[0] + gA[1] * iA[1] + gA[2] * iA[2] + hA);
                    const double uB = wB * (gB[0] * iB[0] + gB[1] * iB[1] + gB[2] * iB[2] + hB);
                    for (int k = 0; k < 3; ++k)
                    {
                        /* direct term: d acc / d c_i = w */
                        gia[3 * i + k] += wA * gA[k];
                        gib[3 * i + k] += wB * gB[k];
                        /* through the weights: dw/dc_c = -2 b w (c_c - c_i), dw/dc_i = -dw/dc_c */
                        const double tA = -2.0 * b[0] * uA * (cA[k] - iA[k]);
                        const double tB = -2.0 * b[1] * uB * (cB[k] - iB[k]);
                        gia[3 * c + k] += tA; gia[3 * i + k] -= tA;
                        gib[3 * c + k] += tB; gib[3 * i + k] -= tB;
                        const double ta = -2.0 * b[2] * (uA + uB) * (ca[k] - iaa[k]);
                        const double tn = -2.0 * b[3] * (uA + uB) * (cn[k] - in[k]);
                        galb[3 * c + k] += ta; galb[3 * i + k] -= ta;
                        gnrm[3 * c + k] += tn; gnrm[3 * i + k] -= tn;
                    }
                }
        }
}

/* ------------------------------------------------------------------------------------------- */
/* a8: reprojection, Reproject reproject.cu.cc:53-238 with isReprojValid :32-50.                */
/* Exact-FP32 decisions; output values accumulated in double.                                   */
/* inputs/outputs: n_list buffers with dims[l] channels each (per frame).                       */
/* The 3x3 fallback reads prev_opacity of the same frame (the reference omits the batch offset */
/* at :191, which only differs for batch > 1).                                                  */
/* ------------------------------------------------------------------------------------------- */
static inline float dot3_pinned(float dx, float dy, float dz) { return fmaf(dz, dz, fmaf(dx, dx, dy * dy)); }

static inline int reproj_valid(int W, int H, int px, int py, float Z, float Zp, float fwZ,
                               const float n[3], const float np[3], float fwN)
{
    if (px < 1 || py < 1 || px > W - 1 || py > H - 1) return 0;
    if (fabsf(Zp - Z) / (fwZ + 1e-2f) > 10.f) return 0;
    const float d = sqrtf(dot3_pinned(n[0] - np[0], n[1] - np[1], n[2] - np[2]));
    if ((double)d / ((double)fwN + 1e-2) > 4.0) return 0;
    return 1;
}

void orc_reproject(const float *mvec, int n_list, const float *const *inputs, const int *dims,
                   const float *linZ, const float *prevLinZ, const float *nrm, const float *prevNrm,
                   const float *pnFw, const float *opac, const float *prevOpac, uint8_t *success,
                   double *const *outputs, int H, int W)
{
#pragma omp parallel for schedule(static)
    for (int cy = 0; cy < H; ++cy)
        for (int cx = 0; cx < W; ++cx)
        {
            const size_t c = (size_t)cy * W + cx;
            for (int l = 0; l < n_list; ++l)
                for (int k = 0; k < dims[l]; ++k) outputs[l][c * dims[l] + k] = 0.0;
            if (opac[c] == 0.f)
            {
                success[c] = 0;
                continue;
            }
            const float mx = mvec[3 * c], my = mvec[3 * c + 1];
            const float fwN = pnFw[3 * c + 1];
            const float ppx = fmaf(mx, (float)W, (float)cx); /* floor(posH) + motion*dim */
            const float ppy = fmaf(my, (float)H, (float)cy);
            const int ipx = f2i_trunc(ppx + 0.5f), ipy = f2i_trunc(ppy + 0.5f);
            const float Z = linZ[3 * c], fwZ = linZ[3 * c + 1];
            const float n[3] = {nrm[3 * c], nrm[3 * c + 1], nrm[3 * c + 2]};
            const int bx = f2i_trunc(ppx), by = f2i_trunc(ppy);
            const int ox[4] = {0, 1, 0, 1}, oy[4] = {0, 0, 1, 1};
            int v[4], valid = 0;
            for (int t = 0; t < 4; ++t)
            {
                const int lx = bx + ox[t], ly = by + oy[t];
                v[t] = 0;
                if (lx >= W || lx < 0 || ly >= H || ly < 0) continue;
                const size_t li = (size_t)ly * W + lx;
                if (prevOpac[li] == 0.f) continue;
                const float np[3] = {prevNrm[3 * li], prevNrm[3 * li + 1], prevNrm[3 * li + 2]};
                v[t] = reproj_valid(W, H, ipx, ipy, Z, prevLinZ[3 * li], fwZ, n, np, fwN);
                valid = valid || v[t];
            }
            if (valid)
            {
                const float x = ppx - floorf(ppx), y = ppy - floorf(ppy);
                const float w[4] = {(1 - x) * (1 - y), x * (1 - y), (1 - x) * y, x * y};
                float sumw = 0.f;
                for (int t = 0; t < 4; ++t)
                {
                    if (!v[t]) continue;
                    const size_t li = (size_t)(by + oy[t]) * W + (bx + ox[t]);
                    for (int l = 0; l < n_list; ++l)
                        for (int k = 0; k < dims[l]; ++k)
                            outputs[l][c * dims[l] + k] += (double)w[t] * inputs[l][li * dims[l] + k];
                    sumw += w[t];
                }
                valid = (sumw >= 0.01f);
                for (int l = 0; l < n_list; ++l)
                    for (int k = 0; k < dims[l]; ++k)
                    {
                        double *o = &outputs[l][c * dims[l] + k];
                        *o = valid ? *o / (double)sumw : 0.0;
                    }
            }
            if (!valid)
            {
                int nvalid = 0;
                for (int yy = -1; yy <= 1; ++yy)
                    for (int xx = -1; xx <= 1; ++xx)
                    {
                        const int qx = ipx + xx, qy = ipy + yy;
                        if (qx >= W || qx < 0 || qy >= H || qy < 0) continue;
                        const size_t qi = (size_t)qy * W + qx;
                        if (prevOpac[qi] == 0.f) continue;
                        const float np[3] = {prevNrm[3 * qi], prevNrm[3 * qi + 1], prevNrm[3 * qi + 2]};
                        if (!reproj_valid(W, H, ipx, ipy, Z, prevLinZ[3 * qi], fwZ, n, np, fwN)) continue;
                        for (int l = 0; l < n_list; ++l)
                            for (int k = 0; k < dims[l]; ++k)
                                outputs[l][c * dims[l] + k] += inputs[l][qi * dims[l] + k];
                        ++nvalid;
                    }
                if (nvalid > 0)
                {
                    valid = 1;
                    for (int l = 0; l < n_list; ++l)
                        for (int k = 0; k < dims[l]; ++k) outputs[l][c * dims[l] + k] /= (double)nvalid;
                }
                else
                    for (int l = 0; l < n_list; ++l)
                        for (int k = 0; k < dims[l]; ++k) outputs[l][c * dims[l] + k] = 0.0;
            }
            success[c] = (uint8_t)(valid ? 1 : 0);
        }
}

/* ------------------------------------------------------------------------------------------- */
/* a9: 3x3 erosion of the success mask, scripts/main.py:199-202,270-276                         */
/* (tf.nn.erosion2d with a ones filter, SAME padding, +1): min over in-bounds neighbours.       */
/* ------------------------------------------------------------------------------------------- */
void orc_erosion(const uint8_t *succ, float *out, int H, int W)
{
#pragma omp parallel for schedule(static)
    for (int cy = 0; cy < H; ++cy)
        for (int cx = 0; cx < W; ++cx)
        {
            float m = 1.f;
            for (int iy = cy - 1; iy <= cy + 1; ++iy)
                for (int ix = cx - 1; ix <= cx + 1; ++ix)
                {
                    if (ix < 0 || iy < 0 || ix >= W || iy >= H) continue;
                    const float v = succ[(size_t)iy * W + ix] ? 1.f : 0.f;
                    if (v < m) m = v;
                }
            out[(size_t)cy * W + cx] = m;
        }
}

/* ------------------------------------------------------------------------------------------- */
/* a10: temporal screening, TemporalScreening temporal.cu.cc:57-198 (window walk fo2 :32-55).   */
/* Exact FP32 (outputs are selections of the inputs). FLT_MAX is 1e7 there (:29).               */
/* ------------------------------------------------------------------------------------------- */
#define SCR_BIG 10000000.f
void orc_screening(const float *c1, const float *c2, const float *pv, const float *pv1,
                   const float *pv2, const float *rsucc, float *o, float *o1, float *o2,
                   float *osucc, int H, int W, int win_curr, int win_prev)
{
    const int hc = win_curr / 2, hp = win_prev / 2;
#pragma omp parallel for schedule(static)
    for (int cy = 0; cy < H; ++cy)
        for (int cx = 0; cx < W; ++cx)
        {
            const size_t c = (size_t)cy * W + cx;
            float cur[3], cur1[3], cur2[3];
            for (int k = 0; k < 3; ++k)
            {
                cur1[k] = c1[3 * c + k];
                cur2[k] = c2[3 * c + k];
                cur[k] = (cur1[k] + cur2[k]) * 0.5f;
            }
            if (rsucc[c] + 1e-4f < 1.f)
            {
                for (int k = 0; k < 3; ++k)
                    o[3 * c + k] = cur[k], o1[3 * c + k] = cur1[k], o2[3 * c + k] = cur2[k];
                osucc[c] = 0.f;
                continue;
            }
            float mnC[3][3], mxC[3][3], mnP[3][3], mxP[3][3]; /* [set avg,1,2][ch] */
            for (int s = 0; s < 3; ++s)
                for (int k = 0; k < 3; ++k) mnC[s][k] = mnP[s][k] = SCR_BIG, mxC[s][k] = mxP[s][k] = -SCR_BIG;
            for (int iy = cy - hc; iy <= cy + hc; ++iy)
                for (int ix = cx - hc; ix <= cx + hc; ++ix)
                {
                    if (ix < 0 || ix >= W || iy < 0 || iy >= H) continue;
                    const size_t i = (size_t)iy * W + ix;
                    for (int k = 0; k < 3; ++k)
                    {
                        const float a = c1[3 * i + k], b = c2[3 * i + k], m = (a + b) * 0.5f;
                        mnC[0][k] = fminf(mnC[0][k], m); mxC[0][k] = fmaxf(mxC[0][k], m);
                        mnC[1][k] = fminf(mnC[1][k], a); mxC[1][k] = fmaxf(mxC[1][k], a);
                        mnC[2][k] = fminf(mnC[2][k], b); mxC[2][k] = fmaxf(mxC[2][k], b);
                    }
                }
            for (int iy = cy - hp; iy <= cy + hp; ++iy)
                for (int ix = cx - hp; ix <= cx + hp; ++ix)
                {
                    if (ix < 0 || ix >= W || iy < 0 || iy >= H) continue;
                    const size_t i = (size_t)iy * W + ix;
                    if (rsucc[i] < 1.f) continue;
                    for (int k = 0; k < 3; ++k)
                    {
                        mnP[0][k] = fminf(mnP[0][k], pv[3 * i + k]); mxP[0][k] = fmaxf(mxP[0][k], pv[3 * i + k]);
                        mnP[1][k] = fminf(mnP[1][k], pv1[3 * i + k]); mxP[1][k] = fmaxf(mxP[1][k], pv1[3 * i + k]);
                        mnP[2][k] = fminf(mnP[2][k], pv2[3 * i + k]); mxP[2][k] = fmaxf(mxP[2][k], pv2[3 * i + k]);
                    }
                }
            int keep[3];
            for (int s = 0; s < 3; ++s)
            {
                keep[s] = 1;
                for (int k = 0; k < 3; ++k)
                    if (!(mnP[s][k] < mxC[s][k]) || !(mnC[s][k] < mxP[s][k])) keep[s] = 0;
            }
            for (int k = 0; k < 3; ++k)
            {
                o[3 * c + k] = keep[0] ? pv[3 * c + k] : cur[k];
                o1[3 * c + k] = keep[1] ? pv1[3 * c + k] : cur1[k];
                o2[3 * c + k] = keep[2] ? pv2[3 * c + k] : cur2[k];
            }
            osucc[c] = keep[0] ? 1.f : 0.f;
        }
}

/* Host threads the OpenMP loops above use (for timing the oracle as a CPU baseline). */
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_get_threads(void) { return omp_get_max_threads(); }
