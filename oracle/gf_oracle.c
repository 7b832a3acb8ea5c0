/*
 * gf_oracle.c -- plain-C restatement of the reference GeoClaw Fortran step.
 *
 * TEST INFRASTRUCTURE ONLY (see gf_oracle.h).  Every routine follows the
 * reference Fortran statement by statement, in the same floating-point
 * operation order, so that a build with -ffp-contract=off reproduces the
 * reference bit for bit.  Fortran intrinsics are restated with the semantics
 * LLVM Flang gives them (checked against the reference build):
 *   MAX(a,b)  -> a > b ? a : b        MIN(a,b) -> a < b ? a : b
 *   DMAX1/DMIN1 with three arguments fold left to right
 *   SIGN(a,b) -> copysign(|a|, b)     x**2 -> x*x      x**(7/3) -> pow()
 *   (gfo_pow: the libm's pow, or with gfo_set_pow_mode(1) correctly rounded)
 * Unary minus in Fortran applies to a whole product (-a*b = -(a*b)); IEEE
 * negation commutes with rounding, so C's (-a)*b is bitwise identical.
 */
#include "gf_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MEQN 3
#define MWAVES 3

static inline double FMAX(double a, double b) { return a > b ? a : b; }
static inline double FMIN(double a, double b) { return a < b ? a : b; }

/* The two powers of the step, x**(7.d0/3.d0) (src2) and x**(2.0d0/3.0d0)
 * (hydrograph inflow depth).  By default the host libm's pow, as the
 * reference build calls it.  With gfo_set_pow_mode(1): the correctly rounded
 * value of x^y (the exact real power of the double y), computed as
 * x^(k/3) * x^d with y = k/3 + d exactly (d = 2^-51/3 for 7/3, -2^-53/3 for
 * 2/3): cbrt(x) refined by one Newton step in double-double from exact FMA
 * products (J.-M. Muller et al., Handbook of Floating-Point Arithmetic, 2nd
 * ed., 2018, secs. 4.4, 5.3), x^d = 1 + d ln x (|d ln x| < 1e-13), one final
 * rounding -- the reference's algorithm evaluated without the libm's rare
 * misroundings, which is how the device library evaluates these powers. */
static int g_pow_cr = 0;

void gfo_set_pow_mode(int correctly_rounded) { g_pow_cr = correctly_rounded != 0; }

static double cbrt_dd(double x, double *lo)
{
    const double c0 = cbrt(x), a = c0 * c0, a_lo = fma(c0, c0, -a);
    const double b = a * c0, b_lo = fma(a, c0, -b);
    const double r = (b - x) + (b_lo + a_lo * c0);
    *lo = -r / (3.0 * a);
    return c0;
}

double gfo_pow(double x, double y)
{
    if (!g_pow_cr || !(x > 0x1p-300 && x < 0x1p300) || (y != 7.0 / 3.0 && y != 2.0 / 3.0))
        return pow(x, y);
    double c_lo;
    const double c = cbrt_dd(x, &c_lo);
    if (y == 7.0 / 3.0) {
        const double p = x * x, e = fma(x, x, -p), hi = p * c;
        const double lo = fma(p, c, -hi) + (p * c_lo + e * c);
        const double t = log(x) / 6755399441055744.0;
        return hi + (lo + hi * t);
    }
    const double hi = c * c, lo = fma(c, c, -hi) + 2.0 * (c * c_lo);
    const double t = -log(x) / 27021597764222976.0;
    return hi + (lo + hi * t);
}

/* 1-D slice helpers: Fortran index i in [1-mbc, n+mbc] -> 0-based k = i+mbc-1 */
#define K(i) ((i) + mbc - 1)
#define Q1(m, i) q1d[K(i) * MEQN + (m) - 1]
#define AX(a, i) (a)[K(i) * maux]
#define FW(m, mw, i) fwave[(K(i) * MWAVES + (mw) - 1) * MEQN + (m) - 1]
#define SS(mw, i) s[K(i) * MWAVES + (mw) - 1]
#define V3(a, m, i) (a)[K(i) * MEQN + (m) - 1]

/* ------------------------------------------------------------------------- */
/* riemanntype: geoclaw_riemann_utils.f:566-677                              */
/* ------------------------------------------------------------------------- */
void gfo_riemanntype(double hL, double hR, double uL, double uR, int maxiter,
                     double drytol, double g, double out[3], int rare[2])
{
    double hm, s1m, s2m, u1m, u2m, um, delu;
    double h_max, h_min, h0, F_max, F_min, dfdh, F0, slope, gL, gR;
    int rare1, rare2;

    h_min = FMIN(hR, hL);
    h_max = FMAX(hR, hL);
    delu = uR - uL;

    if (h_min <= drytol) {
        hm = 0.0;
        um = 0.0;
        s1m = uR + uL - 2.0 * sqrt(g * hR) + 2.0 * sqrt(g * hL);
        s2m = uR + uL - 2.0 * sqrt(g * hR) + 2.0 * sqrt(g * hL);
        if (hL <= 0.0) {
            rare2 = 1;
            rare1 = 0;
        } else {
            rare1 = 1;
            rare2 = 0;
        }
    } else {
        F_min = delu + 2.0 * (sqrt(g * h_min) - sqrt(g * h_max));
        F_max = delu + (h_max - h_min) * (sqrt(.5 * g * (h_max + h_min) / (h_max * h_min)));

        if (F_min > 0.0) { /* 2-rarefactions */
            double t = FMAX(0.0, -delu + 2.0 * (sqrt(g * hL) + sqrt(g * hR)));
            hm = (1.0 / (16.0 * g)) * (t * t);
            um = copysign(1.0, hm) * (uL + 2.0 * (sqrt(g * hL) - sqrt(g * hm)));
            s1m = uL + 2.0 * sqrt(g * hL) - 3.0 * sqrt(g * hm);
            s2m = uR - 2.0 * sqrt(g * hR) + 3.0 * sqrt(g * hm);
            rare1 = 1;
            rare2 = 1;
        } else if (F_max <= 0.0) { /* 2 shocks */
            h0 = h_max;
            for (int iter = 1; iter <= maxiter; iter++) {
                gL = sqrt(.5 * g * (1 / h0 + 1 / hL));
                gR = sqrt(.5 * g * (1 / h0 + 1 / hR));
                F0 = delu + (h0 - hL) * gL + (h0 - hR) * gR;
                dfdh = gL - g * (h0 - hL) / (4.0 * (h0 * h0) * gL) + gR -
                       g * (h0 - hR) / (4.0 * (h0 * h0) * gR);
                slope = 2.0 * sqrt(h0) * dfdh;
                double t = sqrt(h0) - F0 / slope;
                h0 = t * t;
            }
            hm = h0;
            u1m = uL - (hm - hL) * sqrt((.5 * g) * (1 / hm + 1 / hL));
            u2m = uR + (hm - hR) * sqrt((.5 * g) * (1 / hm + 1 / hR));
            um = .5 * (u1m + u2m);
            s1m = u1m - sqrt(g * hm);
            s2m = u2m + sqrt(g * hm);
            rare1 = 0;
            rare2 = 0;
        } else { /* one shock one rarefaction */
            h0 = h_min;
            for (int iter = 1; iter <= maxiter; iter++) {
                F0 = delu + 2.0 * (sqrt(g * h0) - sqrt(g * h_max)) +
                     (h0 - h_min) * sqrt(.5 * g * (1 / h0 + 1 / h_min));
                slope = (F_max - F0) / (h_max - h_min);
                h0 = h0 - F0 / slope;
            }
            hm = h0;
            if (hL > hR) {
                um = uL + 2.0 * sqrt(g * hL) - 2.0 * sqrt(g * hm);
                s1m = uL + 2.0 * sqrt(g * hL) - 3.0 * sqrt(g * hm);
                s2m = uL + 2.0 * sqrt(g * hL) - sqrt(g * hm);
                rare1 = 1;
                rare2 = 0;
            } else {
                s2m = uR - 2.0 * sqrt(g * hR) + 3.0 * sqrt(g * hm);
                s1m = uR - 2.0 * sqrt(g * hR) + sqrt(g * hm);
                um = uR - 2.0 * sqrt(g * hR) + 2.0 * sqrt(g * hm);
                rare2 = 1;
                rare1 = 0;
            }
        }
    }
    (void)um;
    out[0] = hm;
    out[1] = s1m;
    out[2] = s2m;
    rare[0] = rare1;
    rare[1] = rare2;
}

/* ------------------------------------------------------------------------- */
/* riemann_aug_JCP: geoclaw_riemann_utils.f:2-311 (rarecorrectortest is      */
/* .false. in the reference, so the middle wave is always the steady-state   */
/* "jump" wave with eigenvector (0,0,1)).                                    */
/* ------------------------------------------------------------------------- */
void gfo_riemann_aug_jcp(int maxiter, const double in[18], double drytol, double g,
                         double sw[3], double fw[9])
{
    double hL = in[0], hR = in[1], huL = in[2], huR = in[3];
    double bL = in[6], bR = in[7], uL = in[8], uR = in[9];
    double vL = in[10], vR = in[11], phiL = in[12], phiR = in[13];
    double pL = in[14], pR = in[15], sE1 = in[16], sE2 = in[17];
    double A[3][3], r[3][3], lambda[3], del[3], beta[3];
    double delh, delhu, delphi, delb, delP, delnorm;
    double criticaltol, criticaltol_2, hustar_interface;
    double s1s2bar, s1s2tilde, hbar, hLstar, hRstar, huRstar, huLstar, uRstar, uLstar, hstarHLL;
    double deldelh, deldelphi, s1m, s2m, hm;
    double det1, det2, det3, determinant;
    int rare[2], sonic;
    double rt[3];

    (void)in[4];
    (void)in[5];
    delh = hR - hL;
    delhu = huR - huL;
    delphi = phiR - phiL;
    delb = bR - bL;
    delP = pR - pL;
    delnorm = delh * delh + delphi * delphi;

    gfo_riemanntype(hL, hR, uL, uR, 1, drytol, g, rt, rare);
    hm = rt[0];
    s1m = rt[1];
    s2m = rt[2];
    (void)hm;

    lambda[0] = FMIN(sE1, s2m);
    lambda[2] = FMAX(sE2, s1m);
    sE1 = lambda[0];
    sE2 = lambda[2];
    lambda[1] = 0.0;

    hstarHLL = FMAX((huL - huR + sE2 * hR - sE1 * hL) / (sE2 - sE1), 0.0);

    /* r(m,mw) stored as r[m][mw] */
    for (int mw = 0; mw < 3; mw++) {
        r[0][mw] = 1.0;
        r[1][mw] = lambda[mw];
        r[2][mw] = lambda[mw] * lambda[mw];
    }
    /* .not.rarecorrector */
    lambda[1] = 0.5 * (lambda[0] + lambda[2]);
    r[0][1] = 0.0;
    r[1][1] = 0.0;
    r[2][1] = 1.0;

    criticaltol = FMAX(drytol * g, 1e-6);
    criticaltol_2 = sqrt(criticaltol);
    deldelh = -delb;
    deldelphi = -0.5 * (hR + hL) * (g * delb);

    hLstar = hL;
    hRstar = hR;
    uLstar = uL;
    uRstar = uR;
    huLstar = uLstar * hLstar;
    huRstar = uRstar * hRstar;

    for (int iter = 1; iter <= maxiter; iter++) {
        double t;
        hbar = FMAX(0.5 * (hLstar + hRstar), 0.0);
        t = uLstar + uRstar;
        s1s2bar = 0.25 * (t * t) - g * hbar;
        s1s2tilde = FMAX(0.0, uLstar * uRstar) - g * hbar;

        sonic = 0;
        if (fabs(s1s2bar) <= criticaltol) {
            sonic = 1;
        } else if (s1s2bar * s1s2tilde <= criticaltol * criticaltol) {
            sonic = 1;
        } else if (s1s2bar * sE1 * sE2 <= criticaltol * criticaltol) {
            sonic = 1;
        } else if (FMIN(fabs(sE1), fabs(sE2)) < criticaltol_2) {
            sonic = 1;
        } else if (sE1 < criticaltol_2 && s1m > -criticaltol_2) {
            sonic = 1;
        } else if (sE2 > -criticaltol_2 && s2m < criticaltol_2) {
            sonic = 1;
        } else if ((uL + sqrt(g * hL)) * (uR + sqrt(g * hR)) < 0.0) {
            sonic = 1;
        } else if ((uL - sqrt(g * hL)) * (uR - sqrt(g * hR)) < 0.0) {
            sonic = 1;
        }

        if (sonic)
            deldelh = -delb;
        else
            deldelh = delb * g * hbar / s1s2bar;

        if (sE1 < -criticaltol && sE2 > criticaltol) {
            deldelh = FMIN(deldelh, hstarHLL * (sE2 - sE1) / sE2);
            deldelh = FMAX(deldelh, hstarHLL * (sE2 - sE1) / sE1);
        } else if (sE1 >= criticaltol) {
            deldelh = FMIN(deldelh, hstarHLL * (sE2 - sE1) / sE1);
            deldelh = FMAX(deldelh, -hL);
        } else if (sE2 <= -criticaltol) {
            deldelh = FMIN(deldelh, hR);
            deldelh = FMAX(deldelh, hstarHLL * (sE2 - sE1) / sE2);
        }

        if (sonic)
            deldelphi = -g * hbar * delb;
        else
            deldelphi = -delb * g * hbar * s1s2tilde / s1s2bar;

        deldelphi = FMIN(deldelphi, g * FMAX(-hLstar * delb, -hRstar * delb));
        deldelphi = FMAX(deldelphi, g * FMIN(-hLstar * delb, -hRstar * delb));

        del[0] = delh - deldelh;
        del[1] = delhu;
        del[2] = delphi - deldelphi;

        det1 = r[0][0] * (r[1][1] * r[2][2] - r[1][2] * r[2][1]);
        det2 = r[0][1] * (r[1][0] * r[2][2] - r[1][2] * r[2][0]);
        det3 = r[0][2] * (r[1][0] * r[2][1] - r[1][1] * r[2][0]);
        determinant = det1 - det2 + det3;

        for (int k = 0; k < 3; k++) {
            for (int mw = 0; mw < 3; mw++) {
                A[0][mw] = r[0][mw];
                A[1][mw] = r[1][mw];
                A[2][mw] = r[2][mw];
            }
            A[0][k] = del[0];
            A[1][k] = del[1];
            A[2][k] = del[2];
            det1 = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]);
            det2 = A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]);
            det3 = A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
            beta[k] = (det1 - det2 + det3) / determinant;
        }

        /* exit if things aren't changing */
        if (fabs(del[0] * del[0] + del[2] * del[2] - delnorm) < 1e-6)
            break;
        delnorm = del[0] * del[0] + del[2] * del[2];
        hLstar = hL;
        hRstar = hR;
        uLstar = uL;
        uRstar = uR;
        huLstar = uLstar * hLstar;
        huRstar = uRstar * hRstar;
        for (int mw = 0; mw < 3; mw++) {
            if (lambda[mw] < 0.0) {
                hLstar = hLstar + beta[mw] * r[0][mw];
                huLstar = huLstar + beta[mw] * r[1][mw];
            }
        }
        for (int mw = 2; mw >= 0; mw--) {
            if (lambda[mw] > 0.0) {
                hRstar = hRstar - beta[mw] * r[0][mw];
                huRstar = huRstar - beta[mw] * r[1][mw];
            }
        }
        if (hLstar > drytol) {
            uLstar = huLstar / hLstar;
        } else {
            hLstar = FMAX(hLstar, 0.0);
            uLstar = 0.0;
        }
        if (hRstar > drytol) {
            uRstar = huRstar / hRstar;
        } else {
            hRstar = FMAX(hRstar, 0.0);
            uRstar = 0.0;
        }
    }
    (void)delP;
    (void)huLstar;
    (void)huRstar;

    /* fw(m,mw) -> fw[mw*3 + m] */
    for (int mw = 0; mw < 3; mw++) {
        sw[mw] = lambda[mw];
        fw[mw * 3 + 0] = beta[mw] * r[1][mw];
        fw[mw * 3 + 1] = beta[mw] * r[2][mw];
        fw[mw * 3 + 2] = beta[mw] * r[1][mw];
    }
    fw[0 * 3 + 2] = fw[0 * 3 + 2] * vL;
    fw[2 * 3 + 2] = fw[2 * 3 + 2] * vR;
    fw[1 * 3 + 2] = 0.0;

    hustar_interface = huL + fw[0 * 3 + 0];
    if (hustar_interface <= 0.0)
        fw[0 * 3 + 2] = fw[0 * 3 + 2] + (hR * uR * vR - hL * uL * vL - fw[0 * 3 + 2] - fw[2 * 3 + 2]);
    else
        fw[2 * 3 + 2] = fw[2 * 3 + 2] + (hR * uR * vR - hL * uL * vL - fw[0 * 3 + 2] - fw[2 * 3 + 2]);
}

/* ------------------------------------------------------------------------- */
/* rpn2: rpn2_geoclaw.f:2-386                                                */
/* ------------------------------------------------------------------------- */
void gfo_rpn2(const gfo_params *p, int ixy, int mx, int mbc, int maux,
              double *q1d, const double *aux2,
              double *fwave, double *s, double *amdq, double *apdq)
{
    const double g = p->grav, drytol = p->dry_tolerance;
    const int mu = (ixy == 1) ? 2 : 3;
    const int nv = (ixy == 1) ? 3 : 2;

    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        double hR, hL, huR, huL, uR, uL, hvR, hvL, vR, vL, phiR, phiL;
        double bR, bL, sL, sR, sRoe1, sRoe2, sE1, sE2, uhat, chat;
        double hstartest, wall[3], sw[3], fw[9], rt[3], in[18];
        int rare[2];

        for (int mw = 1; mw <= MWAVES; mw++) {
            SS(mw, i) = 0.0;
            FW(1, mw, i) = 0.0;
            FW(2, mw, i) = 0.0;
            FW(3, mw, i) = 0.0;
        }

        /* zero (small) negative values (ql and qr alias q1d) */
        if (Q1(1, i - 1) < 0.0) {
            Q1(1, i - 1) = 0.0;
            Q1(2, i - 1) = 0.0;
            Q1(3, i - 1) = 0.0;
        }
        if (Q1(1, i) < 0.0) {
            Q1(1, i) = 0.0;
            Q1(2, i) = 0.0;
            Q1(3, i) = 0.0;
        }

        if (Q1(1, i - 1) <= drytol && Q1(1, i) <= drytol)
            continue;

        hL = Q1(1, i - 1);
        hR = Q1(1, i);
        huL = Q1(mu, i - 1);
        huR = Q1(mu, i);
        bL = AX(aux2, i - 1);
        bR = AX(aux2, i);
        hvL = Q1(nv, i - 1);
        hvR = Q1(nv, i);

        if (hR > drytol) {
            uR = huR / hR;
            vR = hvR / hR;
            phiR = 0.5 * g * (hR * hR) + (huR * huR) / hR;
        } else {
            hR = 0.0;
            huR = 0.0;
            hvR = 0.0;
            uR = 0.0;
            vR = 0.0;
            phiR = 0.0;
        }
        if (hL > drytol) {
            uL = huL / hL;
            vL = hvL / hL;
            phiL = 0.5 * g * (hL * hL) + (huL * huL) / hL;
        } else {
            hL = 0.0;
            huL = 0.0;
            hvL = 0.0;
            uL = 0.0;
            vL = 0.0;
            phiL = 0.0;
        }

        wall[0] = 1.0;
        wall[1] = 1.0;
        wall[2] = 1.0;
        if (hR <= drytol) {
            gfo_riemanntype(hL, hL, uL, -uL, 1, drytol, g, rt, rare);
            hstartest = FMAX(hL, rt[0]);
            if (hstartest + bL < bR) {
                wall[1] = 0.0;
                wall[2] = 0.0;
                hR = hL;
                huR = -huL;
                bR = bL;
                phiR = phiL;
                uR = -uL;
                vR = vL;
            } else if (hL + bL < bR) {
                bR = hL + bL;
            }
        } else if (hL <= drytol) {
            gfo_riemanntype(hR, hR, -uR, uR, 1, drytol, g, rt, rare);
            hstartest = FMAX(hR, rt[0]);
            if (hstartest + bR < bL) {
                wall[0] = 0.0;
                wall[1] = 0.0;
                hL = hR;
                huL = -huR;
                bL = bR;
                phiL = phiR;
                uL = -uR;
                vL = vR;
            } else if (hR + bR < bL) {
                bL = hR + bR;
            }
        }

        sL = uL - sqrt(g * hL);
        sR = uR + sqrt(g * hR);
        uhat = (sqrt(g * hL) * uL + sqrt(g * hR) * uR) / (sqrt(g * hR) + sqrt(g * hL));
        chat = sqrt(g * 0.5 * (hR + hL));
        sRoe1 = uhat - chat;
        sRoe2 = uhat + chat;
        sE1 = FMIN(sL, sRoe1);
        sE2 = FMAX(sR, sRoe2);

        in[0] = hL; in[1] = hR; in[2] = huL; in[3] = huR; in[4] = hvL; in[5] = hvR;
        in[6] = bL; in[7] = bR; in[8] = uL; in[9] = uR; in[10] = vL; in[11] = vR;
        in[12] = phiL; in[13] = phiR; in[14] = 0.0; in[15] = 0.0; in[16] = sE1; in[17] = sE2;
        gfo_riemann_aug_jcp(1, in, drytol, g, sw, fw);

        for (int mw = 0; mw < 3; mw++) {
            sw[mw] = sw[mw] * wall[mw];
            fw[mw * 3 + 0] = fw[mw * 3 + 0] * wall[mw];
            fw[mw * 3 + 1] = fw[mw * 3 + 1] * wall[mw];
            fw[mw * 3 + 2] = fw[mw * 3 + 2] * wall[mw];
        }
        for (int mw = 1; mw <= MWAVES; mw++) {
            SS(mw, i) = sw[mw - 1];
            FW(1, mw, i) = fw[(mw - 1) * 3 + 0];
            FW(mu, mw, i) = fw[(mw - 1) * 3 + 1];
            FW(nv, mw, i) = fw[(mw - 1) * 3 + 2];
        }
    }

    /* fluctuations: amdq(1:3,:) = 0, apdq(1:3,:) = 0 over the whole slice */
    for (int i = 1 - mbc; i <= mx + mbc; i++)
        for (int m = 1; m <= MEQN; m++) {
            V3(amdq, m, i) = 0.0;
            V3(apdq, m, i) = 0.0;
        }
    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        for (int mw = 1; mw <= MWAVES; mw++) {
            if (SS(mw, i) < 0.0) {
                for (int m = 1; m <= 3; m++)
                    V3(amdq, m, i) = V3(amdq, m, i) + FW(m, mw, i);
            } else if (SS(mw, i) > 0.0) {
                for (int m = 1; m <= 3; m++)
                    V3(apdq, m, i) = V3(apdq, m, i) + FW(m, mw, i);
            } else {
                for (int m = 1; m <= 3; m++) {
                    V3(amdq, m, i) = V3(amdq, m, i) + 0.5 * FW(m, mw, i);
                    V3(apdq, m, i) = V3(apdq, m, i) + 0.5 * FW(m, mw, i);
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* rpt2: rpt2_geoclaw.f:2-190 (Cartesian coordinates)                        */
/* ------------------------------------------------------------------------- */
void gfo_rpt2(const gfo_params *p, int ixy, int imp, int mx, int mbc, int maux,
              const double *q1d, const double *aux1, const double *aux2,
              const double *aux3, const double *asdq,
              double *bmasdq, double *bpasdq)
{
    const double g = p->grav, tol = p->dry_tolerance;
    const int mu = (ixy == 1) ? 2 : 3;
    const int mv = (ixy == 1) ? 3 : 2;

    for (int i = 1 - mbc; i <= mx + mbc; i++)
        for (int m = 1; m <= MEQN; m++) {
            V3(bmasdq, m, i) = 0.0;
            V3(bpasdq, m, i) = 0.0;
        }

    for (int i = 2 - mbc; i <= mx + mbc; i++) {
        double h, u, v, eta, topo1, topo3, s[3], beta[3], r[3][3];
        double delf1, delf2, delf3, dxdcm, dxdcp;
        const int ic = (imp == 1) ? i - 1 : i;

        h = Q1(1, ic);
        if (h <= tol)
            continue;
        u = Q1(mu, ic) / h;
        v = Q1(mv, ic) / h;
        eta = h + AX(aux2, ic);
        topo1 = AX(aux1, ic);
        topo3 = AX(aux3, ic);

        if (eta < FMIN(topo1, topo3))
            continue;

        dxdcp = 1.0;
        dxdcm = 1.0;

        s[0] = v - sqrt(g * h);
        s[1] = v;
        s[2] = v + sqrt(g * h);

        delf1 = V3(asdq, 1, i);
        delf2 = V3(asdq, mu, i);
        delf3 = V3(asdq, mv, i);

        beta[0] = (s[2] * delf1 - delf3) / (s[2] - s[0]);
        beta[1] = -u * delf1 + delf2;
        beta[2] = (delf3 - s[0] * delf1) / (s[2] - s[0]);

        r[0][0] = 1.0; r[1][0] = u;   r[2][0] = s[0];
        r[0][1] = 0.0; r[1][1] = 1.0; r[2][1] = 0.0;
        r[0][2] = 1.0; r[1][2] = u;   r[2][2] = s[2];

        for (int mw = 0; mw < 3; mw++) {
            if ((s[mw] < 0.0) && (eta >= topo1)) {
                V3(bmasdq, 1, i) = V3(bmasdq, 1, i) + dxdcm * s[mw] * beta[mw] * r[0][mw];
                V3(bmasdq, mu, i) = V3(bmasdq, mu, i) + dxdcm * s[mw] * beta[mw] * r[1][mw];
                V3(bmasdq, mv, i) = V3(bmasdq, mv, i) + dxdcm * s[mw] * beta[mw] * r[2][mw];
            } else if ((s[mw] > 0.0) && (eta >= topo3)) {
                V3(bpasdq, 1, i) = V3(bpasdq, 1, i) + dxdcp * s[mw] * beta[mw] * r[0][mw];
                V3(bpasdq, mu, i) = V3(bpasdq, mu, i) + dxdcp * s[mw] * beta[mw] * r[1][mw];
                V3(bpasdq, mv, i) = V3(bpasdq, mv, i) + dxdcp * s[mw] * beta[mw] * r[2][mw];
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* philim.f:65-126 and geoclaw_limiter_fort.f:4-61                           */
/* ------------------------------------------------------------------------- */
double gfo_philim(double a, double b, int meth)
{
    double r, c;
    if (a == 0.0)
        return 0.0;
    r = b / a;
    switch (meth) {
    case 1:
        return FMAX(0.0, FMIN(1.0, r));
    case 2:
        return FMAX(FMAX(0.0, FMIN(1.0, 2.0 * r)), FMIN(2.0, r));
    case 3:
        return (r + fabs(r)) / (1.0 + fabs(r));
    case 4:
        c = (1.0 + r) / 2.0;
        return FMAX(0.0, FMIN(FMIN(c, 2.0), 2.0 * r));
    case 5:
        return r;
    default:
        /* a computed GO TO whose index is out of range falls through to the
         * next statement, which is the minmod branch */
        return FMAX(0.0, FMIN(1.0, r));
    }
}

void gfo_limiter(int mx, int mbc, double *wave, const double *s, const int mthlim[3])
{
    double *fwave = wave;
    for (int mw = 1; mw <= MWAVES; mw++) {
        double dotr, dotl, wnorm2, wlimitr;
        if (mthlim[mw - 1] == 0)
            continue;
        dotr = 0.0;
        for (int i = 0; i <= mx + 1; i++) {
            wnorm2 = 0.0;
            dotl = dotr;
            dotr = 0.0;
            for (int m = 1; m <= MEQN; m++) {
                wnorm2 = wnorm2 + FW(m, mw, i) * FW(m, mw, i);
                dotr = dotr + FW(m, mw, i) * FW(m, mw, i + 1);
            }
            if (i == 0)
                continue;
            if (wnorm2 == 0.0)
                continue;
            if (SS(mw, i) > 0.0)
                wlimitr = gfo_philim(wnorm2, dotl, mthlim[mw - 1]);
            else
                wlimitr = gfo_philim(wnorm2, dotr, mthlim[mw - 1]);
            for (int m = 1; m <= MEQN; m++)
                FW(m, mw, i) = wlimitr * FW(m, mw, i);
        }
    }
}

/* ------------------------------------------------------------------------- */
/* fc2d_geoclaw_flux2: geoclaw_flux2fw_fort.f:4-269                          */
/* ------------------------------------------------------------------------- */
void gfo_flux2(const gfo_params *p, int ixy, int mx, int mbc, int maux,
               double *q1d, double *dtdx1d,
               const double *aux1, const double *aux2, const double *aux3,
               double *faddm, double *faddp, double *gaddm, double *gaddp,
               double *cfl1d_out)
{
    const int n = mx + 2 * mbc;
    double *fwave = (double *)malloc(sizeof(double) * (size_t)n * MEQN * MWAVES);
    double *s = (double *)malloc(sizeof(double) * (size_t)n * MWAVES);
    double *amdq = (double *)malloc(sizeof(double) * (size_t)n * MEQN);
    double *apdq = (double *)malloc(sizeof(double) * (size_t)n * MEQN);
    double *cqxx = (double *)calloc((size_t)n * MEQN, sizeof(double));
    double *bmasdq = (double *)malloc(sizeof(double) * (size_t)n * MEQN);
    double *bpasdq = (double *)malloc(sizeof(double) * (size_t)n * MEQN);
    int limit = 0;
    double cfl1d;
#define DT(i) dtdx1d[K(i)]
#define G2(a, m, i, side) (a)[((side) - 1) * n * MEQN + K(i) * MEQN + (m) - 1]

    for (int mw = 0; mw < MWAVES; mw++)
        if (p->mthlim[mw] > 0)
            limit = 1;

    for (int jside = 1; jside <= 2; jside++)
        for (int i = 1 - mbc; i <= mx + mbc; i++)
            for (int m = 1; m <= MEQN; m++) {
                V3(faddm, m, i) = 0.0;
                V3(faddp, m, i) = 0.0;
                G2(gaddm, m, i, jside) = 0.0;
                G2(gaddp, m, i, jside) = 0.0;
            }

    gfo_rpn2(p, ixy, mx, mbc, maux, q1d, aux2, fwave, s, amdq, apdq);

    for (int i = 2 - mbc; i <= mx + mbc - 1; i++)
        for (int m = 1; m <= MEQN; m++) {
            V3(faddp, m, i) = V3(faddp, m, i) - V3(apdq, m, i);
            V3(faddm, m, i) = V3(faddm, m, i) + V3(amdq, m, i);
        }

    cfl1d = 0.0;
    for (int mw = 1; mw <= MWAVES; mw++)
        for (int i = 2 - mbc; i <= mx + mbc; i++) {
            cfl1d = FMAX(cfl1d, DT(i) * SS(mw, i));
            cfl1d = FMAX(cfl1d, -DT(i - 1) * SS(mw, i));
        }

    if (p->order != 1) {
        if (limit)
            gfo_limiter(mx, mbc, fwave, s, p->mthlim);
        for (int i = 2 - mbc; i <= mx + mbc; i++) {
            DT(i - 1) = 0.5 * (DT(i - 1) + DT(i));
            for (int m = 1; m <= MEQN; m++) {
                V3(cqxx, m, i) = 0.0;
                for (int mw = 1; mw <= MWAVES; mw++)
                    V3(cqxx, m, i) = V3(cqxx, m, i) + copysign(1.0, SS(mw, i)) *
                                     (1.0 - fabs(SS(mw, i)) * DT(i - 1)) * FW(m, mw, i);
                V3(faddm, m, i) = V3(faddm, m, i) + 0.5 * V3(cqxx, m, i);
                V3(faddp, m, i) = V3(faddp, m, i) + 0.5 * V3(cqxx, m, i);
            }
        }
    }

    if (p->transverse != 0) {
        if (p->order > 1 && p->transverse == 2) {
            for (int i = 2 - mbc; i <= mx + mbc - 1; i++)
                for (int m = 1; m <= MEQN; m++) {
                    V3(amdq, m, i) = V3(amdq, m, i) + V3(cqxx, m, i);
                    V3(apdq, m, i) = V3(apdq, m, i) - V3(cqxx, m, i);
                }
        }
        gfo_rpt2(p, ixy, 1, mx, mbc, maux, q1d, aux1, aux2, aux3, amdq, bmasdq, bpasdq);
        for (int i = 2 - mbc; i <= mx + mbc; i++)
            for (int m = 1; m <= MEQN; m++) {
                double gupdate = 0.5 * DT(i - 1) * V3(bmasdq, m, i);
                G2(gaddm, m, i - 1, 1) = G2(gaddm, m, i - 1, 1) - gupdate;
                G2(gaddp, m, i - 1, 1) = G2(gaddp, m, i - 1, 1) - gupdate;
                gupdate = 0.5 * DT(i - 1) * V3(bpasdq, m, i);
                G2(gaddm, m, i - 1, 2) = G2(gaddm, m, i - 1, 2) - gupdate;
                G2(gaddp, m, i - 1, 2) = G2(gaddp, m, i - 1, 2) - gupdate;
            }
        gfo_rpt2(p, ixy, 2, mx, mbc, maux, q1d, aux1, aux2, aux3, apdq, bmasdq, bpasdq);
        for (int i = 2 - mbc; i <= mx + mbc; i++)
            for (int m = 1; m <= MEQN; m++) {
                double gupdate = 0.5 * DT(i - 1) * V3(bmasdq, m, i);
                G2(gaddm, m, i, 1) = G2(gaddm, m, i, 1) - gupdate;
                G2(gaddp, m, i, 1) = G2(gaddp, m, i, 1) - gupdate;
                gupdate = 0.5 * DT(i - 1) * V3(bpasdq, m, i);
                G2(gaddm, m, i, 2) = G2(gaddm, m, i, 2) - gupdate;
                G2(gaddp, m, i, 2) = G2(gaddp, m, i, 2) - gupdate;
            }
    }
#undef DT
#undef G2
    *cfl1d_out = cfl1d;
    free(fwave);
    free(s);
    free(amdq);
    free(apdq);
    free(cqxx);
    free(bmasdq);
    free(bpasdq);
}

/* ------------------------------------------------------------------------- */
/* fc2d_geoclaw_step2 (geoclaw_step2_fort.f90:1-294) + the update of         */
/* fc2d_geoclaw_step2_wrap (geoclaw_step2_wrap_fort.f:94-115), mcapa = 0.     */
/* ------------------------------------------------------------------------- */
#define P2(a, nc, m, i, j) (a)[((size_t)((j) + mbc - 1) * (mx + 2 * mbc) + (size_t)((i) + mbc - 1)) * (nc) + (m) - 1]

double gfo_step2_wrap(const gfo_params *p, int mx, int my, int mbc, int maux,
                      double *qold, double *aux, double dx, double dy, double dt)
{
    const int maxm = mx > my ? mx : my;
    const int n = maxm + 2 * mbc;
    const size_t ncell = (size_t)(mx + 2 * mbc) * (my + 2 * mbc);
    double *fm = (double *)calloc(ncell * MEQN, sizeof(double));
    double *fp = (double *)calloc(ncell * MEQN, sizeof(double));
    double *gm = (double *)calloc(ncell * MEQN, sizeof(double));
    double *gp = (double *)calloc(ncell * MEQN, sizeof(double));
    double *faddm = (double *)malloc(sizeof(double) * n * MEQN);
    double *faddp = (double *)malloc(sizeof(double) * n * MEQN);
    double *gaddm = (double *)malloc(sizeof(double) * n * MEQN * 2);
    double *gaddp = (double *)malloc(sizeof(double) * n * MEQN * 2);
    double *q1d = (double *)malloc(sizeof(double) * n * MEQN);
    double *aux1 = (double *)malloc(sizeof(double) * n * maux);
    double *aux2 = (double *)malloc(sizeof(double) * n * maux);
    double *aux3 = (double *)malloc(sizeof(double) * n * maux);
    double *dtdx1d = (double *)malloc(sizeof(double) * n);
    double cflgrid, cfl1d, dtdx, dtdy;

    cflgrid = 0.0;
    dtdx = dt / dx;
    dtdy = dt / dy;

    /* x-sweeps */
    for (int j = 2 - mbc; j <= my + mbc - 1; j++) {
        for (int i = 1 - mbc; i <= mx + mbc; i++)
            for (int m = 1; m <= MEQN; m++)
                V3(q1d, m, i) = P2(qold, MEQN, m, i, j);
        for (int k = 0; k < n; k++)
            dtdx1d[k] = dtdx;
        for (int i = 1 - mbc; i <= mx + mbc; i++)
            for (int ma = 0; ma < maux; ma++) {
                aux1[K(i) * maux + ma] = P2(aux, maux, ma + 1, i, j - 1);
                aux2[K(i) * maux + ma] = P2(aux, maux, ma + 1, i, j);
                aux3[K(i) * maux + ma] = P2(aux, maux, ma + 1, i, j + 1);
            }
        gfo_flux2(p, 1, mx, mbc, maux, q1d, dtdx1d, aux1, aux2, aux3,
                  faddm, faddp, gaddm, gaddp, &cfl1d);
        cflgrid = FMAX(cflgrid, cfl1d);
        const int nx = mx + 2 * mbc;
        for (int i = 2 - mbc; i <= mx + mbc; i++)
            for (int m = 1; m <= MEQN; m++) {
                P2(fm, MEQN, m, i, j) = P2(fm, MEQN, m, i, j) + V3(faddm, m, i);
                P2(fp, MEQN, m, i, j) = P2(fp, MEQN, m, i, j) + V3(faddp, m, i);
                P2(gm, MEQN, m, i, j) = P2(gm, MEQN, m, i, j) + gaddm[K(i) * MEQN + m - 1];
                P2(gp, MEQN, m, i, j) = P2(gp, MEQN, m, i, j) + gaddp[K(i) * MEQN + m - 1];
                P2(gm, MEQN, m, i, j + 1) = P2(gm, MEQN, m, i, j + 1) + gaddm[nx * MEQN + K(i) * MEQN + m - 1];
                P2(gp, MEQN, m, i, j + 1) = P2(gp, MEQN, m, i, j + 1) + gaddp[nx * MEQN + K(i) * MEQN + m - 1];
            }
    }

    /* y-sweeps */
    for (int i = 2 - mbc; i <= mx + mbc - 1; i++) {
        for (int j = 1 - mbc; j <= my + mbc; j++)
            for (int m = 1; m <= MEQN; m++)
                V3(q1d, m, j) = P2(qold, MEQN, m, i, j);
        for (int k = 0; k < n; k++)
            dtdx1d[k] = dtdy;
        for (int j = 1 - mbc; j <= my + mbc; j++)
            for (int ma = 0; ma < maux; ma++) {
                aux1[K(j) * maux + ma] = P2(aux, maux, ma + 1, i - 1, j);
                aux2[K(j) * maux + ma] = P2(aux, maux, ma + 1, i, j);
                aux3[K(j) * maux + ma] = P2(aux, maux, ma + 1, i + 1, j);
            }
        gfo_flux2(p, 2, my, mbc, maux, q1d, dtdx1d, aux1, aux2, aux3,
                  faddm, faddp, gaddm, gaddp, &cfl1d);
        cflgrid = FMAX(cflgrid, cfl1d);
        const int ny = my + 2 * mbc;
        for (int j = 2 - mbc; j <= my + mbc; j++)
            for (int m = 1; m <= MEQN; m++) {
                P2(gm, MEQN, m, i, j) = P2(gm, MEQN, m, i, j) + V3(faddm, m, j);
                P2(gp, MEQN, m, i, j) = P2(gp, MEQN, m, i, j) + V3(faddp, m, j);
                P2(fm, MEQN, m, i, j) = P2(fm, MEQN, m, i, j) + gaddm[K(j) * MEQN + m - 1];
                P2(fp, MEQN, m, i, j) = P2(fp, MEQN, m, i, j) + gaddp[K(j) * MEQN + m - 1];
                P2(fm, MEQN, m, i + 1, j) = P2(fm, MEQN, m, i + 1, j) + gaddm[ny * MEQN + K(j) * MEQN + m - 1];
                P2(fp, MEQN, m, i + 1, j) = P2(fp, MEQN, m, i + 1, j) + gaddp[ny * MEQN + K(j) * MEQN + m - 1];
            }
    }

    /* update q (step2_wrap), ghost cells included */
    for (int m = 1; m <= MEQN; m++)
        for (int i = 1 - mbc; i <= mx + mbc - 1; i++)
            for (int j = 1 - mbc; j <= my + mbc - 1; j++)
                P2(qold, MEQN, m, i, j) = P2(qold, MEQN, m, i, j)
                    - dtdx * (P2(fm, MEQN, m, i + 1, j) - P2(fp, MEQN, m, i, j))
                    - dtdy * (P2(gm, MEQN, m, i, j + 1) - P2(gp, MEQN, m, i, j));

    free(fm); free(fp); free(gm); free(gp);
    free(faddm); free(faddp); free(gaddm); free(gaddp);
    free(q1d); free(aux1); free(aux2); free(aux3); free(dtdx1d);
    return cflgrid;
}

/* ------------------------------------------------------------------------- */
void gfo_b4step2(const gfo_params *p, int mx, int my, int mbc, double *q)
{
    for (int j = 1 - mbc; j <= my + mbc; j++)
        for (int i = 1 - mbc; i <= mx + mbc; i++)
            if (P2(q, MEQN, 1, i, j) < p->dry_tolerance) {
                P2(q, MEQN, 1, i, j) = FMAX(P2(q, MEQN, 1, i, j), 0.0);
                P2(q, MEQN, 2, i, j) = 0.0;
                P2(q, MEQN, 3, i, j) = 0.0;
            }
}

void gfo_src2(const gfo_params *p, int mx, int my, int mbc, int maux,
              double *q, const double *aux, double dt)
{
    const double g = p->grav;
    /* manning_break(num_manning) = rinfinity (geoclaw_module.f90:110),
     * rinfinity = 10.e32, a default-real literal (amr_module.f90:171) */
    const double manning_break = (double)10.e32f;
    (void)aux;
    (void)maux;
    if (!p->friction_forcing)
        return;
    for (int j = 1; j <= my; j++)
        for (int i = 1; i <= mx; i++) {
            if (P2(q, MEQN, 1, i, j) < 1.0e-30) {
                P2(q, MEQN, 2, i, j) = 0.0;
                P2(q, MEQN, 3, i, j) = 0.0;
            } else if (P2(q, MEQN, 1, i, j) <= p->friction_depth) {
                double coeff = 0.0, gamma, dgamma, h = P2(q, MEQN, 1, i, j);
                double hu = P2(q, MEQN, 2, i, j), hv = P2(q, MEQN, 3, i, j);
                if (P2(aux, maux, 1, i, j) < manning_break)
                    coeff = p->manning_n;
                gamma = sqrt(hu * hu + hv * hv) * g * (coeff * coeff) / gfo_pow(h, 7.0 / 3.0);
                dgamma = 1.0 + dt * gamma;
                P2(q, MEQN, 2, i, j) = hu / dgamma;
                P2(q, MEQN, 3, i, j) = P2(q, MEQN, 3, i, j) / dgamma;
            }
        }
}

/* ------------------------------------------------------------------------- */
/* fc2d_geoclaw_bc2 (geoflood_bc2_fort.f90): extrapolation and solid walls.  */
/* ------------------------------------------------------------------------- */
void gfo_bc2(int mx, int my, int mbc, int maux, double *q, double *aux, const int mthbc[4])
{
    /* left */
    if (mthbc[0] == 1) {
        for (int j = 1 - mbc; j <= my + mbc; j++)
            for (int ibc = 1; ibc <= mbc; ibc++) {
                P2(aux, maux, 1, 1 - ibc, j) = P2(aux, maux, 1, 1, j);
                for (int m = 1; m <= MEQN; m++)
                    P2(q, MEQN, m, 1 - ibc, j) = P2(q, MEQN, m, 1, j);
            }
    } else if (mthbc[0] == 3) {
        for (int j = 1 - mbc; j <= my + mbc; j++)
            for (int ibc = 1; ibc <= mbc; ibc++) {
                P2(aux, maux, 1, 1 - ibc, j) = P2(aux, maux, 1, ibc, j);
                for (int m = 1; m <= MEQN; m++)
                    P2(q, MEQN, m, 1 - ibc, j) = P2(q, MEQN, m, ibc, j);
            }
        for (int j = 1 - mbc; j <= my + mbc; j++)
            for (int ibc = 1; ibc <= mbc; ibc++)
                P2(q, MEQN, 2, 1 - ibc, j) = -P2(q, MEQN, 2, ibc, j);
    }
    /* right */
    if (mthbc[1] == 1) {
        for (int j = 1 - mbc; j <= my + mbc; j++)
            for (int ibc = 1; ibc <= mbc; ibc++) {
                P2(aux, maux, 1, mx + ibc, j) = P2(aux, maux, 1, mx, j);
                for (int m = 1; m <= MEQN; m++)
                    P2(q, MEQN, m, mx + ibc, j) = P2(q, MEQN, m, mx, j);
            }
    } else if (mthbc[1] == 3) {
        for (int j = 1 - mbc; j <= my + mbc; j++)
            for (int ibc = 1; ibc <= mbc; ibc++) {
                P2(aux, maux, 1, mx + ibc, j) = P2(aux, maux, 1, mx + 1 - ibc, j);
                for (int m = 1; m <= MEQN; m++)
                    P2(q, MEQN, m, mx + ibc, j) = P2(q, MEQN, m, mx + 1 - ibc, j);
            }
        for (int j = 1 - mbc; j <= my + mbc; j++)
            for (int ibc = 1; ibc <= mbc; ibc++)
                P2(q, MEQN, 2, mx + ibc, j) = -P2(q, MEQN, 2, mx + 1 - ibc, j);
    }
    /* bottom */
    if (mthbc[2] == 1) {
        for (int jbc = 1; jbc <= mbc; jbc++)
            for (int i = 1 - mbc; i <= mx + mbc; i++) {
                P2(aux, maux, 1, i, 1 - jbc) = P2(aux, maux, 1, i, 1);
                for (int m = 1; m <= MEQN; m++)
                    P2(q, MEQN, m, i, 1 - jbc) = P2(q, MEQN, m, i, 1);
            }
    } else if (mthbc[2] == 3) {
        for (int jbc = 1; jbc <= mbc; jbc++)
            for (int i = 1 - mbc; i <= mx + mbc; i++) {
                P2(aux, maux, 1, i, 1 - jbc) = P2(aux, maux, 1, i, jbc);
                for (int m = 1; m <= MEQN; m++)
                    P2(q, MEQN, m, i, 1 - jbc) = P2(q, MEQN, m, i, jbc);
            }
        for (int jbc = 1; jbc <= mbc; jbc++)
            for (int i = 1 - mbc; i <= mx + mbc; i++)
                P2(q, MEQN, 3, i, 1 - jbc) = -P2(q, MEQN, 3, i, jbc);
    }
    /* top */
    if (mthbc[3] == 1) {
        for (int jbc = 1; jbc <= mbc; jbc++)
            for (int i = 1 - mbc; i <= mx + mbc; i++) {
                P2(aux, maux, 1, i, my + jbc) = P2(aux, maux, 1, i, my);
                for (int m = 1; m <= MEQN; m++)
                    P2(q, MEQN, m, i, my + jbc) = P2(q, MEQN, m, i, my);
            }
    } else if (mthbc[3] == 3) {
        for (int jbc = 1; jbc <= mbc; jbc++)
            for (int i = 1 - mbc; i <= mx + mbc; i++) {
                P2(aux, maux, 1, i, my + jbc) = P2(aux, maux, 1, i, my + 1 - jbc);
                for (int m = 1; m <= MEQN; m++)
                    P2(q, MEQN, m, i, my + jbc) = P2(q, MEQN, m, i, my + 1 - jbc);
            }
        for (int jbc = 1; jbc <= mbc; jbc++)
            for (int i = 1 - mbc; i <= mx + mbc; i++)
                P2(q, MEQN, 3, i, my + jbc) = -P2(q, MEQN, 3, i, my + 1 - jbc);
    }
}

/* ------------------------------------------------------------------------- */
/* Hydrograph inflow                                                          */
/* ------------------------------------------------------------------------- */
double gfo_hydro_interpolate(const gfo_hydrograph *h, double t)
{
    /* hydrograh_module.f90:186-196: clamp outside the table, else the first
     * interval [time(i), time(i+1)] holding t, linear in t. */
    const int n = h->nrows;
    if (t < h->time[0])
        return h->hu[0];
    if (t > h->time[n - 1])
        return h->hu[n - 1];
    for (int i = 0; i < n - 1; i++)
        if (t >= h->time[i] && t <= h->time[i + 1])
            return h->hu[i] + (h->hu[i + 1] - h->hu[i]) / (h->time[i + 1] - h->time[i]) * (t - h->time[i]);
    return 0.0;
}

/* tol = 1.0e-6 is a default-real literal in the reference: its double value. */
#define HYDRO_TOL ((double)1.0e-6f)
#define HYDRO_MAXITER 100

void gfo_riemann_invariants(double q0[2], const double q1[2], double grav)
{
    /* hydrograh_module.f90:213-242 (discharge branch), Fr = 0.5 */
    const double u1 = q1[1] / q1[0];
    const double Fr = 0.5;
    q0[0] = gfo_pow(q0[1] / sqrt(grav) * Fr, 2.0 / 3.0);
    for (int it = 0; it < HYDRO_MAXITER; it++) {
        const double fxn = q0[1] / q0[0] - 2.0 * sqrt(grav * q0[0]) - u1 + 2.0 * sqrt(grav * q1[0]);
        if (fabs(fxn) < HYDRO_TOL)
            return;
        const double dfxn = -(q0[1] / (q0[0] * q0[0])) - sqrt(grav / q0[0]);
        q0[0] = q0[0] - fxn / dfxn;
    }
}

void gfo_two_shock(double q0[2], const double q1[2], double grav)
{
    /* hydrograh_module.f90:274-303 (discharge branch), Fr = 1 */
    const double u1 = q1[1] / q1[0];
    const double Fr = 1.0;
    q0[0] = gfo_pow(q0[1] / sqrt(grav) * Fr, 2.0 / 3.0);
    for (int it = 0; it < HYDRO_MAXITER; it++) {
        const double fxn = q0[1] / q0[0] - u1 -
                           (q0[0] - q1[0]) * sqrt((grav / 2.0) * (1.0 / q0[0] + 1.0 / q1[0]));
        if (fabs(fxn) < HYDRO_TOL)
            return;
        const double num = sqrt(grav * (1.0 / q0[0] + 1.0 / q1[0]));
        const double deno = 2.0 * sqrt(2.0) * (q0[0] * q0[0]) * num;
        const double dfxn = -(q0[1] / (q0[0] * q0[0])) - num / sqrt(2.0) + grav * (q0[0] - q1[0]) / deno;
        q0[0] = q0[0] - fxn / dfxn;
    }
}

/* One side of bc2 with mthbc = 0 (geoflood_bc2_fort.f90:52-102 left,
 * :141-183 right, :229-272 bottom, :318-361 top).  For the boundary cell of
 * each row (column) the interpolated discharge sets the ghost state inside
 * the channel strip: dry boundary cell -> critical depth with the discharge,
 * wet with non-zero discharge -> depth from the Riemann invariant (two-shock
 * when that depth exceeds the cell's), zero discharge -> wall.  Outside the
 * strip the side is a wall.  Wall branches also mirror aux; the inflow
 * branches leave aux alone.  The inflow state puts the discharge into q(2)
 * on every side, as the reference does. */
static void bc2_inflow_side(const gfo_params *p, int side, int mx, int my, int mbc, int maux,
                            double *q, double *aux, double xlower, double ylower,
                            double dx, double dy, double t, const gfo_hydrograph *hg)
{
    const double g = p->grav;
    const int xside = side < 2;
    const int n = xside ? my : mx;
    const int nrm = xside ? 2 : 3;
    for (int r = 1 - mbc; r <= n + mbc; r++) {
        /* boundary cell (bi, bj) of this row */
        const int bi = xside ? (side == 0 ? 1 : mx) : r;
        const int bj = xside ? r : (side == 2 ? 1 : my);
        const double c = xside ? ylower + ((double)r - 0.5) * dy : xlower + ((double)r - 0.5) * dx;
        const double c0 = xside ? hg->y0 : hg->x0;
        const double q1[2] = {P2(q, MEQN, 1, bi, bj), P2(q, MEQN, 2, bi, bj)};
        double q0[2];
        q0[1] = gfo_hydro_interpolate(hg, t);
        const int in_strip = fabs(c - c0) <= hg->b / 2.0;
        for (int k = 1; k <= mbc; k++) {
            int gi, gj, mi_, mj_;
            if (side == 0) { gi = 1 - k; gj = r; mi_ = k; mj_ = r; }
            else if (side == 1) { gi = mx + k; gj = r; mi_ = mx + 1 - k; mj_ = r; }
            else if (side == 2) { gi = r; gj = 1 - k; mi_ = r; mj_ = k; }
            else { gi = r; gj = my + k; mi_ = r; mj_ = my + 1 - k; }
            int mirror = !in_strip;
            if (in_strip) {
                if (q1[0] < p->dry_tolerance) {
                    P2(q, MEQN, 1, gi, gj) = FMAX(gfo_pow(q0[1] / sqrt(g), 2.0 / 3.0), 0.001);
                    P2(q, MEQN, 2, gi, gj) = q0[1];
                    P2(q, MEQN, 3, gi, gj) = 0.0;
                } else if (q0[1] != 0.0) {
                    gfo_riemann_invariants(q0, q1, g);
                    if (q0[0] > q1[0])
                        gfo_two_shock(q0, q1, g);
                    P2(q, MEQN, 1, gi, gj) = q0[0];
                    P2(q, MEQN, 2, gi, gj) = q0[1];
                    P2(q, MEQN, 3, gi, gj) = 0.0;
                } else {
                    mirror = 1;
                }
            }
            if (mirror) {
                P2(aux, maux, 1, gi, gj) = P2(aux, maux, 1, mi_, mj_);
                for (int m = 1; m <= MEQN; m++)
                    P2(q, MEQN, m, gi, gj) = P2(q, MEQN, m, mi_, mj_);
                P2(q, MEQN, nrm, gi, gj) = -P2(q, MEQN, nrm, mi_, mj_);
            }
        }
    }
}

int gfo_bc2_ex(const gfo_params *p, int mx, int my, int mbc, int maux, double *q,
               double *aux, const int mthbc[4], double xlower, double ylower,
               double dx, double dy, double t, const gfo_hydrograph *h)
{
    for (int side = 0; side < 4; side++) {
        if (mthbc[side] == 0) {
            if (!h || !h->side[side] || h->nrows < 1)
                return -1;
            bc2_inflow_side(p, side, mx, my, mbc, maux, q, aux, xlower, ylower, dx, dy, t, h);
        } else {
            int one[4] = {-1, -1, -1, -1};
            one[side] = mthbc[side];
            gfo_bc2(mx, my, mbc, maux, q, aux, one);
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Refinement tagging                                                         */
/* ------------------------------------------------------------------------- */
int gfo_flag2refine(const gfo_params *p, const gfo_refine *r, int mx, int my, int mbc, int maux,
                    const double *q, const double *aux, double x1, double y1, double x2,
                    double y2, double t, int level, int maxlevel, int init_flag,
                    int is_coarsening, int i, int j)
{
    (void)my;
    /* no coarsening while the initial grids are built (:27-31) */
    if (init_flag != 0 && is_coarsening)
        return 0;
    /* allowflag (geolib_source/allowflag.f) is always true. */
    /* forced refinement in a region (:42-50) */
    for (int m = 0; m < r->mregions; m++) {
        const double *g = r->region + 8 * m;
        if (level < (int)g[6] && t >= g[4] && t <= g[5])
            if (x2 > g[0] && x1 < g[2] && y2 > g[1] && y1 < g[3])
                return 1;
    }
    const int max_num_speeds = r->num_speeds < maxlevel ? r->num_speeds : maxlevel;
    const double hu = P2(q, MEQN, 2, i, j), hv = P2(q, MEQN, 3, i, j);
    const double momentum = sqrt(hu * hu + hv * hv);
    const double depth = P2(q, MEQN, 1, i, j);
    const double eta = depth + P2(aux, maux, 1, i, j);
    /* th_factor is an INTEGER in the reference: 1.0 -> 1, 0.5 -> 0 (:59-61) */
    const int th_factor = is_coarsening ? 0 : 1;
    for (int m = 0; m < r->num_flowgrades; m++) {
        int needed = 0;
        switch (r->iflowgradevariable[m]) {
        case 1: needed = depth > r->flowgradevalue[m]; break;
        case 2: needed = momentum > r->flowgradevalue[m]; break;
        case 3:
            if (depth > p->dry_tolerance)
                needed = fabs(eta) > r->flowgradevalue[m];
            break;
        default: break;
        }
        if (needed && level < r->iflowgrademinlevel[m])
            return 1;
    }
    double speed = 0.0;
    if (depth > p->dry_tolerance) {
        speed = momentum / depth;
        for (int n = 1; n <= max_num_speeds; n++)
            if (speed > th_factor * r->speed_tolerance[n - 1] && level <= n)
                return 1;
    }
    if (r->num_flowgrades == 0 && depth > p->dry_tolerance) {
        if (fabs(eta - r->sea_level) > th_factor * r->wave_tolerance) {
            if (level < r->max_level_deep)
                return 1;
            if (fabs(P2(aux, maux, 1, i, j)) < r->deep_depth)
                return 1;
        }
        if (fabs(P2(aux, maux, 1, i, j)) > r->deep_depth && speed < r->speed_tolerance[0])
            return 0;
        if (speed / depth > r->max_velocity_depth_product)
            return 1;
    }
    return 0;
}

int gfo_tag4refinement(const gfo_params *p, const gfo_refine *r, int mx, int my, int mbc,
                       int maux, double xlower, double ylower, double dx, double dy, double t,
                       const double *q, const double *aux, int level, int maxlevel,
                       int init_flag)
{
    for (int j = 1; j <= my; j++) {
        const double y1 = ylower + (j - 1) * dy, y2 = ylower + j * dy;
        for (int i = 1; i <= mx; i++) {
            const double x1 = xlower + (i - 1) * dx, x2 = xlower + i * dx;
            if (gfo_flag2refine(p, r, mx, my, mbc, maux, q, aux, x1, y1, x2, y2, t, level, maxlevel,
                                init_flag, 0, i, j) > 0)
                return 1;
        }
    }
    return 0;
}

int gfo_tag4coarsening(const gfo_params *p, const gfo_refine *r, int mx, int my, int mbc,
                       int maux, const double xlower[4], const double ylower[4], double dx,
                       double dy, double t, const double *const q[4],
                       const double *const aux[4], int level, int maxlevel, int init_flag)
{
    for (int k = 0; k < 4; k++)
        for (int j = 1; j <= my; j++) {
            const double y1 = ylower[k] + (j - 1) * dy, y2 = ylower[k] + j * dy;
            for (int i = 1; i <= mx; i++) {
                const double x1 = xlower[k] + (i - 1) * dx, x2 = xlower[k] + i * dx;
                if (gfo_flag2refine(p, r, mx, my, mbc, maux, q[k], aux[k], x1, y1, x2, y2, t, level,
                                    maxlevel, init_flag, 1, i, j) > 0)
                    return 0;
            }
        }
    return 1;
}

/* ------------------------------------------------------------------------- */
double gfo_patch_update(const gfo_params *p, int mx, int my, int mbc, int maux,
                        double *q, double *aux, double dx, double dy,
                        double t, double dt)
{
    double maxcfl;
    (void)t;
    gfo_b4step2(p, mx, my, mbc, q);
    maxcfl = gfo_step2_wrap(p, mx, my, mbc, maux, q, aux, dx, dy, dt);
    if (maxcfl > 1000.0) /* MAXCFL_CAP, fc2d_cpucuda.h:43 */
        maxcfl = 0.0;
    if (p->src_term > 0)
        gfo_src2(p, mx, my, mbc, maux, q, aux, dt);
    return maxcfl;
}

static double builtin_update(const void *ctx, int mx, int my, int mbc, int maux,
                             double *q, double *aux, double dx, double dy,
                             double t, double dt)
{
    return gfo_patch_update((const gfo_params *)ctx, mx, my, mbc, maux, q, aux, dx, dy, t, dt);
}

/* ------------------------------------------------------------------------- */
/* Uniform brick driver                                                      */
/* ------------------------------------------------------------------------- */
int gfo_brick_ghost_fill_ex(const gfo_params *p, int mi, int mj, int mx, int my, int mbc,
                            int maux, double **q, double **aux, const int mthbc[4],
                            double xlo, double ylo, double dx, double dy, double t,
                            const gfo_hydrograph *h)
{
    int rc = 0;
    /* same-level face copies (q only; aux ghosts are set at creation) */
    for (int pj = 0; pj < mj; pj++)
        for (int pi = 0; pi < mi; pi++) {
            double *d = q[pj * mi + pi];
            if (pi > 0) {
                const double *s = q[pj * mi + pi - 1];
                for (int j = 1; j <= my; j++)
                    for (int i = 1 - mbc; i <= 0; i++)
                        for (int m = 1; m <= MEQN; m++)
                            P2(d, MEQN, m, i, j) = P2(s, MEQN, m, mx + i, j);
            }
            if (pi < mi - 1) {
                const double *s = q[pj * mi + pi + 1];
                for (int j = 1; j <= my; j++)
                    for (int i = mx + 1; i <= mx + mbc; i++)
                        for (int m = 1; m <= MEQN; m++)
                            P2(d, MEQN, m, i, j) = P2(s, MEQN, m, i - mx, j);
            }
            if (pj > 0) {
                const double *s = q[(pj - 1) * mi + pi];
                for (int j = 1 - mbc; j <= 0; j++)
                    for (int i = 1; i <= mx; i++)
                        for (int m = 1; m <= MEQN; m++)
                            P2(d, MEQN, m, i, j) = P2(s, MEQN, m, i, my + j);
            }
            if (pj < mj - 1) {
                const double *s = q[(pj + 1) * mi + pi];
                for (int j = my + 1; j <= my + mbc; j++)
                    for (int i = 1; i <= mx; i++)
                        for (int m = 1; m <= MEQN; m++)
                            P2(d, MEQN, m, i, j) = P2(s, MEQN, m, i, j - my);
            }
        }
    /* same-level corner copies */
    for (int pj = 0; pj < mj; pj++)
        for (int pi = 0; pi < mi; pi++) {
            double *d = q[pj * mi + pi];
            for (int cy = 0; cy < 2; cy++)
                for (int cx = 0; cx < 2; cx++) {
                    int ni = pi + (cx ? 1 : -1), nj = pj + (cy ? 1 : -1);
                    if (ni < 0 || ni >= mi || nj < 0 || nj >= mj)
                        continue;
                    const double *s = q[nj * mi + ni];
                    for (int jb = 1; jb <= mbc; jb++)
                        for (int ib = 1; ib <= mbc; ib++) {
                            int i = cx ? mx + ib : 1 - ib, j = cy ? my + jb : 1 - jb;
                            int si = cx ? i - mx : i + mx, sj = cy ? j - my : j + my;
                            for (int m = 1; m <= MEQN; m++)
                                P2(d, MEQN, m, i, j) = P2(s, MEQN, m, si, sj);
                        }
                }
        }
    /* physical boundaries */
    for (int pj = 0; pj < mj; pj++)
        for (int pi = 0; pi < mi; pi++) {
            int bc[4];
            bc[0] = (pi == 0) ? mthbc[0] : -1;
            bc[1] = (pi == mi - 1) ? mthbc[1] : -1;
            bc[2] = (pj == 0) ? mthbc[2] : -1;
            bc[3] = (pj == mj - 1) ? mthbc[3] : -1;
            if (gfo_bc2_ex(p, mx, my, mbc, maux, q[pj * mi + pi], aux[pj * mi + pi], bc,
                           xlo + pi * mx * dx, ylo + pj * my * dy, dx, dy, t, h) != 0)
                rc = -1;
        }
    return rc;
}

void gfo_brick_ghost_fill(int mi, int mj, int mx, int my, int mbc, int maux,
                          double **q, double **aux, const int mthbc[4])
{
    gfo_params p;
    memset(&p, 0, sizeof(p));
    gfo_brick_ghost_fill_ex(&p, mi, mj, mx, my, mbc, maux, q, aux, mthbc, 0.0, 0.0, 1.0, 1.0, 0.0, NULL);
}

static double brick_advance(const gfo_params *p, int mi, int mj, int mx, int my, int mbc,
                            int maux, double **q, double **aux, const int mthbc[4],
                            double xlo, double ylo, double dx, double dy,
                            const gfo_hydrograph *h, double t, double dt, gfo_update_fn fn,
                            const void *ctx, int nthreads)
{
    const int np = mi * mj;
    double maxcfl = 0.0;
    double *cfl = (double *)malloc(sizeof(double) * np);
    if (!fn) {
        fn = builtin_update;
        ctx = p;
    }
    gfo_brick_ghost_fill_ex(p, mi, mj, mx, my, mbc, maux, q, aux, mthbc, xlo, ylo, dx, dy, t, h);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
    for (int k = 0; k < np; k++)
        cfl[k] = fn(ctx, mx, my, mbc, maux, q[k], aux[k], dx, dy, t, dt);
    for (int k = 0; k < np; k++)
        maxcfl = FMAX(maxcfl, cfl[k]);
    free(cfl);
    return maxcfl;
}

double gfo_brick_advance(const gfo_params *p, int mi, int mj, int mx, int my,
                         int mbc, int maux, double **q, double **aux,
                         const int mthbc[4], double dx, double dy, double t,
                         double dt, gfo_update_fn fn, const void *ctx,
                         int nthreads)
{
    return brick_advance(p, mi, mj, mx, my, mbc, maux, q, aux, mthbc, 0.0, 0.0, dx, dy, NULL, t, dt,
                         fn, ctx, nthreads);
}

int gfo_brick_run_ex(const gfo_params *p, int mi, int mj, int mx, int my, int mbc,
                     int maux, double **q, double **aux, const int mthbc[4],
                     double xlo, double ylo, double dx, double dy,
                     const gfo_hydrograph *h, double t0, double initial_dt,
                     double desired_cfl, double max_cfl, int nsteps,
                     gfo_update_fn fn, const void *ctx, int nthreads,
                     double *dt_out, double *cfl_out, double *mass_out)
{
    return gfo_brick_run_dtmax(p, mi, mj, mx, my, mbc, maux, q, aux, mthbc, xlo, ylo, dx, dy, h, t0,
                               initial_dt, desired_cfl, max_cfl, 0.0, nsteps, fn, ctx, nthreads, dt_out,
                               cfl_out, mass_out);
}

int gfo_brick_run_dtmax(const gfo_params *p, int mi, int mj, int mx, int my, int mbc,
                        int maux, double **q, double **aux, const int mthbc[4],
                        double xlo, double ylo, double dx, double dy,
                        const gfo_hydrograph *h, double t0, double initial_dt,
                        double desired_cfl, double max_cfl, double dt_max, int nsteps,
                        gfo_update_fn fn, const void *ctx, int nthreads,
                        double *dt_out, double *cfl_out, double *mass_out)
{
    const int np = mi * mj;
    const size_t nq = (size_t)MEQN * (mx + 2 * mbc) * (my + 2 * mbc);
    double *save = (double *)malloc(sizeof(double) * nq * np);
    double t = t0, dt = initial_dt;
    int n = 0, retakes = 0;

    while (n < nsteps) {
        double maxcfl;
        for (int k = 0; k < np; k++)
            memcpy(save + nq * k, q[k], sizeof(double) * nq);
        maxcfl = brick_advance(p, mi, mj, mx, my, mbc, maux, q, aux, mthbc, xlo, ylo, dx, dy, h,
                               t, dt, fn, ctx, nthreads);
        if (maxcfl > max_cfl) {
            for (int k = 0; k < np; k++)
                memcpy(q[k], save + nq * k, sizeof(double) * nq);
            dt = dt * desired_cfl / maxcfl;
            retakes++;
            continue;
        }
        if (dt_out)
            dt_out[n] = dt;
        if (cfl_out)
            cfl_out[n] = maxcfl;
        if (mass_out) {
            double mass = 0.0;
            for (int k = 0; k < np; k++) {
                double sum = 0.0;
                for (int j = 1; j <= my; j++)
                    for (int i = 1; i <= mx; i++)
                        sum += P2(q[k], MEQN, 1, i, j);
                mass += sum * dx * dy;
            }
            mass_out[n] = mass;
        }
        t = t + dt;
        if (dt_max > 0.0) {
            /* the library's bound (no counterpart in the reference): a zero
             * CFL number takes dt_max, any other is capped at it */
            const double d = maxcfl > 0.0 ? dt * desired_cfl / maxcfl : dt;
            dt = (maxcfl == 0.0 || d > dt_max) ? dt_max : d;
        } else {
            dt = dt * desired_cfl / maxcfl;
        }
        n++;
    }
    free(save);
    return retakes;
}

int gfo_brick_run(const gfo_params *p, int mi, int mj, int mx, int my, int mbc,
                  int maux, double **q, double **aux, const int mthbc[4],
                  double dx, double dy, double t0, double initial_dt,
                  double desired_cfl, double max_cfl, int nsteps,
                  gfo_update_fn fn, const void *ctx, int nthreads,
                  double *dt_out, double *cfl_out, double *mass_out)
{
    return gfo_brick_run_ex(p, mi, mj, mx, my, mbc, maux, q, aux, mthbc, 0.0, 0.0, dx, dy, NULL, t0,
                            initial_dt, desired_cfl, max_cfl, nsteps, fn, ctx, nthreads, dt_out,
                            cfl_out, mass_out);
}

/* ======================================================================
 * ASCII frame output and gauges
 * ====================================================================== */

/* Fortran Ew.d as Flang writes it (probed against the reference writer
 * built by oracle/Makefile): the value rounded to d significant digits
 * (nearest, ties to even, as C's %e), printed as 0.<digits> with the
 * exponent one higher than C's; "E+dd" for |exponent| <= 99, "+ddd"
 * without the E above that; zero keeps its sign. */
void gfo_format_E(char *out, double v, int w, int d)
{
    char body[64];
    int n;
    if (isnan(v)) {
        n = snprintf(body, sizeof body, "NaN");
    } else if (isinf(v)) {
        n = snprintf(body, sizeof body, "%sInf", v < 0 ? "-" : "");
    } else {
        char digits[40];
        int e10 = 0;
        const int neg = signbit(v) ? 1 : 0;
        if (v == 0.0) {
            memset(digits, '0', (size_t)d);
            e10 = 0;
        } else {
            char tmp[64];
            snprintf(tmp, sizeof tmp, "%.*e", d - 1, fabs(v));  /* D.DDDDe+XX */
            digits[0] = tmp[0];
            memcpy(digits + 1, tmp + 2, (size_t)(d - 1));
            e10 = atoi(strchr(tmp, 'e') + 1) + 1;
        }
        digits[d] = 0;
        const int ae = e10 < 0 ? -e10 : e10;
        if (ae <= 99)
            n = snprintf(body, sizeof body, "%s0.%sE%c%02d", neg ? "-" : "", digits, e10 < 0 ? '-' : '+', ae);
        else
            n = snprintf(body, sizeof body, "%s0.%s%c%03d", neg ? "-" : "", digits, e10 < 0 ? '-' : '+', ae);
    }
    if (n > w) {
        memset(out, '*', (size_t)w);
        return;
    }
    memset(out, ' ', (size_t)(w - n));
    memcpy(out + (w - n), body, (size_t)n);
}

/* Fortran I5: right-justified, "*****" when it does not fit. */
static int put_i5(char *out, int v)
{
    char tmp[32];
    const int n = snprintf(tmp, sizeof tmp, "%d", v);
    if (n > 5) {
        memcpy(out, "*****", 5);
    } else {
        memset(out, ' ', (size_t)(5 - n));
        memcpy(out + 5 - n, tmp, (size_t)n);
    }
    return 5;
}

typedef struct { char *p; int len, cap, bad; } obuf;

static void ob_put(obuf *b, const char *s, int n)
{
    if (b->bad || b->len + n > b->cap) { b->bad = 1; return; }
    memcpy(b->p + b->len, s, (size_t)n);
    b->len += n;
}

static void ob_i5_line(obuf *b, int v, const char *label)
{
    char t[8];
    put_i5(t, v);
    ob_put(b, t, 5);
    ob_put(b, "                 ", 17);
    ob_put(b, label, (int)strlen(label));
    ob_put(b, "\n", 1);
}

static void ob_E(obuf *b, double v, int w, int d)
{
    char t[64];
    gfo_format_E(t, v, w, d);
    ob_put(b, t, w);
}

/* fc2d_geoclaw_fort_write_header (fc2d_geoclaw_output_ascii_fort.f:1-36):
 * format 1000 = e30.20 'time', then i5 meqn+1, ngrids, maux, 2, mbc. */
int gfo_frame_header(char *out, int cap, double time, int meqn, int maux, int ngrids, int mbc)
{
    obuf b = {out, 0, cap, 0};
    ob_E(&b, time, 30, 20);
    ob_put(&b, "    time\n", 9);
    ob_i5_line(&b, meqn + 1, "meqn");
    ob_i5_line(&b, ngrids, "ngrids");
    ob_i5_line(&b, maux, "num_aux");
    ob_i5_line(&b, 2, "num_dim");
    ob_i5_line(&b, mbc, "num_ghost");
    return b.bad ? -1 : b.len;
}

/* fc2d_geoclaw_fort_write_file (fc2d_geoclaw_output_ascii_fort.f:38-117):
 * format 1001 (six I5 lines), format 1002 (four e24.16 lines and an empty
 * record), then per row j = 1..my and cell i = 1..mx: q with |q| < 1e-99
 * set to 0 in place, eta = q(1) + aux(1) (0 when |eta| < 1e-99), written as
 * 5E26.16 (four values here), and after each row a list-directed ' ' (a
 * record of two blanks). */
int gfo_frame_patch(char *out, int cap, int mx, int my, int meqn, int maux, int mbc,
                    double xlower, double ylower, double dx, double dy, double *q,
                    const double *aux, int patch_num, int level, int blockno, int mpirank)
{
    obuf b = {out, 0, cap, 0};
    const int nx = mx + 2 * mbc;
    ob_i5_line(&b, patch_num, "grid_number");
    ob_i5_line(&b, level, "AMR_level");
    ob_i5_line(&b, blockno, "block_number");
    ob_i5_line(&b, mpirank, "mpi_rank");
    ob_i5_line(&b, mx, "mx");
    ob_i5_line(&b, my, "my");
    ob_E(&b, xlower, 24, 16); ob_put(&b, "    xlow\n", 9);
    ob_E(&b, ylower, 24, 16); ob_put(&b, "    ylow\n", 9);
    ob_E(&b, dx, 24, 16); ob_put(&b, "    dx\n", 7);
    ob_E(&b, dy, 24, 16); ob_put(&b, "    dy\n", 7);
    ob_put(&b, "\n", 1);
    for (int j = 1; j <= my; j++) {
        for (int i = 1; i <= mx; i++) {
            double *qc = q + ((size_t)(j - 1 + mbc) * nx + (i - 1 + mbc)) * meqn;
            for (int m = 0; m < meqn; m++)
                if (fabs(qc[m]) < 1e-99) qc[m] = 0.0;
            double eta = qc[0] + aux[((size_t)(j - 1 + mbc) * nx + (i - 1 + mbc)) * maux];
            if (fabs(eta) < 1e-99) eta = 0.0;
            for (int m = 0; m < meqn; m++) ob_E(&b, qc[m], 26, 16);
            ob_E(&b, eta, 26, 16);
            ob_put(&b, "\n", 1);
        }
        ob_put(&b, "  \n", 3);
    }
    return b.bad ? -1 : b.len;
}

/* fc2d_geoclaw_update_gauge (geoclaw_gauges_fort.f90:1-77). */
void gfo_update_gauge(const gfo_params *p, int mx, int my, int mbc, int maux, double xlower,
                      double ylower, double dx, double dy, const double *q, const double *aux,
                      double xc, double yc, double qvar[3], double avar[1])
{
    (void)my;
    const int nx = mx + 2 * mbc;
#define GQ(m, i, j) q[((size_t)((j) - 1 + mbc) * nx + ((i) - 1 + mbc)) * 3 + (m)]
#define GA(i, j) aux[((size_t)((j) - 1 + mbc) * nx + ((i) - 1 + mbc)) * maux]
    const int ii = (int)((xc - xlower) / dx) + 1;
    const int jj = (int)((yc - ylower) / dy) + 1;
    const double xcent = xlower + (ii - .5) * dx;
    const double ycent = ylower + (jj - .5) * dy;
    const double xoff = (xc - xcent) / dx;
    const double yoff = (yc - ycent) / dy;
    const double drytol2 = 0.1 * p->dry_tolerance;
    const double h1 = GQ(0, ii, jj), h2 = GQ(0, ii + 1, jj), h3 = GQ(0, ii, jj + 1), h4 = GQ(0, ii + 1, jj + 1);
    double topo;
    if (h1 < drytol2 || h2 < drytol2 || h3 < drytol2 || h4 < drytol2) {
        for (int m = 0; m < 3; m++) qvar[m] = GQ(m, ii, jj);
        topo = GA(ii, jj);
    } else {
        for (int m = 0; m < 3; m++)
            qvar[m] = (1.0 - xoff) * (1.0 - yoff) * GQ(m, ii, jj) + xoff * (1.0 - yoff) * GQ(m, ii + 1, jj) +
                      (1.0 - xoff) * yoff * GQ(m, ii, jj + 1) + xoff * yoff * GQ(m, ii + 1, jj + 1);
        topo = (1.0 - xoff) * (1.0 - yoff) * GA(ii, jj) + xoff * (1.0 - yoff) * GA(ii + 1, jj) +
               (1.0 - xoff) * yoff * GA(ii, jj + 1) + xoff * yoff * GA(ii + 1, jj + 1);
    }
    avar[0] = topo;
    for (int m = 0; m < 3; m++)
        if (fabs(qvar[m]) < 1e-99) qvar[m] = 0.0;
#undef GQ
#undef GA
}

/* geoclaw_create_gauge_files_default (fc2d_geoclaw_gauges_default.c:172-200):
 * the two header lines of gaugeNNNNN.txt, num_eqn = 4. */
int gfo_gauge_header(char *out, int cap, int id, double xc, double yc)
{
    const int n = snprintf(out, (size_t)cap,
                           "# gauge_id= %5d location=( %17.10e %17.10e ) num_eqn= %2d\n"
                           "# Columns: level time h    hu    hv    eta\n", id, xc, yc, 4);
    return (n < 0 || n >= cap) ? -1 : n;
}

/* One record of geoclaw_print_gauges_default (fc2d_geoclaw_gauges_default.c:
 * 275-305): eta = h + b, 0 when |eta| < 1e-99. */
int gfo_gauge_line(char *out, int cap, int level, double t, const double qvar[3], const double avar[1])
{
    double eta = qvar[0] + avar[0];
    eta = fabs(eta) < 1e-99 ? 0 : eta;
    const int n = snprintf(out, (size_t)cap, "%5d %15.7e %15.7e %15.7e %15.7e %15.7e\n", level, t, qvar[0],
                           qvar[1], qvar[2], eta);
    return (n < 0 || n >= cap) ? -1 : n;
}

/* ======================================================================
 * Regrid and time interpolation between levels
 * ====================================================================== */

#define AQ(q, m, i, j) (q)[((size_t)((j) - 1 + mbc) * nx + ((i) - 1 + mbc)) * 3 + (m)]
#define AA(a, i, j) (a)[((size_t)((j) - 1 + mbc) * nx + ((i) - 1 + mbc)) * maux]

/* One coarse cell (i1, j1) from the four fine cells (i2[m], j2[m]), the
 * per-cell rule shared by fc2d_geoclaw_average_fort.f:482-603 (average2coarse),
 * :33-320 (average_face) and :323-480 (average_corner), mcapa = 0 (capa =
 * capac = 1): the wet-cell mean of eta, the mean depth capped by it over the
 * coarse bathymetry, momentum rescaled. */
static void average_cell(const gfo_params *p, int nx, int mbc, int maux, double *qcoarse, const double *qfine,
                         const double *aux_coarse, const double *aux_fine, int i1, int j1, const int i2[4],
                         const int j2[4])
{
    const double capac = 1.0;
    double etasum = 0.0, hsum = 0.0, husum = 0.0, hvsum = 0.0;
    int nwet = 0;
    for (int m = 0; m < 4; m++) {
        const double capa = 1.0;
        const double hf = AQ(qfine, 0, i2[m], j2[m]) * capa;
        const double bf = AA(aux_fine, i2[m], j2[m]) * capa;
        double huf = AQ(qfine, 1, i2[m], j2[m]) * capa;
        double hvf = AQ(qfine, 2, i2[m], j2[m]) * capa;
        double etaf;
        if (hf > p->dry_tolerance) {
            etaf = hf + bf;
            nwet++;
        } else {
            etaf = 0.0;
            huf = 0.0;
            hvf = 0.0;
        }
        hsum = hsum + hf;
        husum = husum + huf;
        hvsum = hvsum + hvf;
        etasum = etasum + etaf;
    }
    double hc, huc, hvc;
    if (nwet > 0) {
        const double etaav = etasum / (double)nwet;
        const double hav = hsum / (double)nwet;
        const double e = etaav - AA(aux_coarse, i1, j1) * capac;
        const double hmx = e > 0.0 ? e : 0.0;  /* Fortran MAX(a, b) = a > b ? a : b */
        hc = hav < hmx ? hav : hmx;            /* MIN */
        huc = (hc / hsum) * husum;
        hvc = (hc / hsum) * hvsum;
    } else {
        hc = 0.0;
        huc = 0.0;
        hvc = 0.0;
    }
    AQ(qcoarse, 0, i1, j1) = hc / capac;
    AQ(qcoarse, 1, i1, j1) = huc / capac;
    AQ(qcoarse, 2, i1, j1) = hvc / capac;
}

/* fc2d_geoclaw_average_fort.f:482-603 with mcapa = 0 (capa = capac = 1). */
void gfo_average2coarse(const gfo_params *p, int mx, int my, int mbc, int maux, double *qcoarse,
                        const double *qfine, const double *aux_coarse, const double *aux_fine, int igrid)
{
    const int nx = mx + 2 * mbc;
    const int ig = igrid % 2, jg = (igrid - ig) / 2;
    const int ic_add = ig * mx / 2, jc_add = jg * my / 2;
    for (int j = 1; j <= my / 2; j++)
        for (int i = 1; i <= mx / 2; i++) {
            const int i1 = i + ic_add, j1 = j + jc_add;
            int i2[4], j2[4], m = 0;
            for (int jj = 1; jj <= 2; jj++)
                for (int ii = 1; ii <= 2; ii++) {
                    i2[m] = (i - 1) * 2 + ii;
                    j2[m] = (j - 1) * 2 + jj;
                    m++;
                }
            average_cell(p, nx, mbc, maux, qcoarse, qfine, aux_coarse, aux_fine, i1, j1, i2, j2);
        }
}

/* The index map of a half-size neighbour with the coarse patch's own
 * orientation (same block, or a brick block without rotation), in place of
 * ForestClaw's fclaw2d_clawpatch_transform_face_half / _corner_half, which
 * the reference tree does not hold: the fine patch covers coarse cells
 * ox+1 .. ox+mx/2 by oy+1 .. oy+my/2, so coarse cell (ic, jc) is fine cells
 * (2(ic-ox)-1 + s, 2(jc-oy)-1 + t), s, t in {0, 1}, listed x fastest.
 * side 0..3: the fine patch across face 0..3 of the coarse patch (left,
 * right, bottom, top) in half igrid (0: lower coordinate along the face);
 * side 4..7: the fine patch across corner 0..3 (lower-left, lower-right,
 * upper-left, upper-right). */
static void half_origin(int side, int igrid, int mx, int my, int *ox, int *oy)
{
    switch (side) {
    case 0: *ox = -mx / 2; *oy = igrid * my / 2; break;
    case 1: *ox = mx;      *oy = igrid * my / 2; break;
    case 2: *ox = igrid * mx / 2; *oy = -my / 2; break;
    case 3: *ox = igrid * mx / 2; *oy = my;      break;
    case 4: *ox = -mx / 2; *oy = -my / 2; break;
    case 5: *ox = mx;      *oy = -my / 2; break;
    case 6: *ox = -mx / 2; *oy = my;      break;
    default: *ox = mx;     *oy = my;      break;
    }
}

static void transform_half(int ox, int oy, int ic, int jc, int i2[4], int j2[4])
{
    const int i0 = 2 * (ic - ox) - 1, j0 = 2 * (jc - oy) - 1;
    i2[0] = i0;     j2[0] = j0;
    i2[1] = i0 + 1; j2[1] = j0;
    i2[2] = i0;     j2[2] = j0 + 1;
    i2[3] = i0 + 1; j2[3] = j0 + 1;
}

/* fc2d_geoclaw_fort_average_face (fc2d_geoclaw_average_fort.f:33-320) and
 * _average_corner (:323-480), mcapa = 0, for side 0..7 as in half_origin:
 * the coarse ghost cells that the fine neighbour covers. */
int gfo_average_ghost(const gfo_params *p, int mx, int my, int mbc, int maux, double *qcoarse,
                      const double *qfine, const double *aux_coarse, const double *aux_fine, int side, int igrid)
{
    const int nx = mx + 2 * mbc;
    int ox, oy, i2[4], j2[4];
    if (side < 0 || side > 7 || igrid < 0 || igrid > 1) return -1;
    half_origin(side, igrid, mx, my, &ox, &oy);
    if (side >= 4) {
        const int icorner = side - 4;
        for (int ibc = 1; ibc <= mbc; ibc++)
            for (int jbc = 1; jbc <= mbc; jbc++) {
                const int i1 = (icorner & 1) ? mx + ibc : 1 - ibc;
                const int j1 = (icorner & 2) ? my + jbc : 1 - jbc;
                transform_half(ox, oy, i1, j1, i2, j2);
                average_cell(p, nx, mbc, maux, qcoarse, qfine, aux_coarse, aux_fine, i1, j1, i2, j2);
            }
        return 0;
    }
    const int idir = side / 2;
    for (int a = 1; a <= (idir == 0 ? my : mbc); a++)
        for (int b = 1; b <= (idir == 0 ? mbc : mx); b++) {
            /* idir 0: jc = a, ibc = b; idir 1: jbc = a, ic = b */
            const int ic = idir == 0 ? (side == 0 ? 1 - b : mx + b) : b;
            const int jc = idir == 0 ? a : (side == 2 ? 1 - a : my + a);
            transform_half(ox, oy, ic, jc, i2, j2);
            int skip = 0;  /* fclaw2d_clawpatch_is_valid_average: 1 <= i <= mx, 1 <= j <= my */
            for (int m = 0; m < 4; m++)
                if (!(i2[m] >= 1 && i2[m] <= mx && j2[m] >= 1 && j2[m] <= my)) skip = 1;
            if (!skip) average_cell(p, nx, mbc, maux, qcoarse, qfine, aux_coarse, aux_fine, ic, jc, i2, j2);
        }
    return 0;
}

/* ForestClaw fclaw2d_clawpatch_compute_slopes, mth = 5: the smallest of
 * 2|sl|, 2|sr| and |sc| (sc = (sl+sr)/2) with the sign of sc, and 0 when
 * sl and sr differ in sign or one is 0.  Restated from ForestClaw's source
 * (not vendored in the reference); mth <= 4 (philim) is not used here. */
double gfo_compute_slopes(double sl, double sr, int mth)
{
    (void)mth;
    const double sc = (sl + sr) / 2.0;
    double a = 2 * fabs(sl), b = 2 * fabs(sr), c = fabs(sc);
    double v = a < b ? a : b;
    v = v < c ? v : c;
    const double mn = sl < sr ? sl : sr, mxv = sl > sr ? sl : sr;
    if (mn <= 0 && mxv >= 0) return 0.0;
    return copysign(v, sc);
}

/* fc2d_geoclaw_interpolate_fort.f:405-579 (the extremum check there is
 * commented out and not restated). */
void gfo_interpolate2fine(const gfo_params *p, int mx, int my, int mbc, int maux, const double *qcoarse,
                          double *qfine, const double *aux_coarse, const double *aux_fine, int igrid)
{
    const int nx = mx + 2 * mbc;
    const double sea_level = 0.0;
    const int mth = 5;
    const int ig = igrid % 2, jg = (igrid - ig) / 2;
    const int i1 = 1 - ig, i2 = mx / 2 + (1 - ig), ic_add = ig * mx / 2;
    const int j1 = 1 - jg, j2 = my / 2 + (1 - jg), jc_add = jg * my / 2;
    for (int mq = 0; mq < 3; mq++)
        for (int i = i1; i <= i2; i++)
            for (int j = j1; j <= j2; j++) {
                const int ic = i + ic_add, jc = j + jc_add;
                double etab[3][3];
                for (int ii = -1; ii <= 1; ii++)
                    for (int jj = -1; jj <= 1; jj++) {
                        const double h = AQ(qcoarse, 0, ic + ii, jc + jj);
                        const double b = AA(aux_coarse, ic + ii, jc + jj);
                        etab[ii + 1][jj + 1] = h < p->dry_tolerance ? sea_level : h + b;
                    }
                double qc, gradx, grady;
                if (mq == 0) {
                    qc = etab[1][1];
                    gradx = gfo_compute_slopes(qc - etab[0][1], etab[2][1] - qc, mth);
                    grady = gfo_compute_slopes(qc - etab[1][0], etab[1][2] - qc, mth);
                } else {
                    qc = AQ(qcoarse, mq, ic, jc);
                    gradx = gfo_compute_slopes(qc - AQ(qcoarse, mq, ic - 1, jc), AQ(qcoarse, mq, ic + 1, jc) - qc, mth);
                    grady = gfo_compute_slopes(qc - AQ(qcoarse, mq, ic, jc - 1), AQ(qcoarse, mq, ic, jc + 1) - qc, mth);
                }
                for (int ii = 1; ii <= 2; ii++)
                    for (int jj = 1; jj <= 2; jj++) {
                        const double shiftx = (ii - 2 / 2.0 - 0.5) / 2;
                        const double shifty = (jj - 2 / 2.0 - 0.5) / 2;
                        const int iff = (i - 1) * 2 + ii, jf = (j - 1) * 2 + jj;
                        if (mq == 0)
                            AQ(qfine, 0, iff, jf) = qc + shiftx * gradx + shifty * grady - AA(aux_fine, iff, jf);
                        else
                            AQ(qfine, mq, iff, jf) = qc + shiftx * gradx + shifty * grady;
                    }
            }
}

/* The coarse cell (ic, jc)'s limited reconstruction written to the fine
 * cells i2[0] + (0|1), j2[0] + (0|1) (x fastest) as the face and corner
 * interpolations do (fc2d_geoclaw_interpolate_fort.f:117-171, :347-396):
 * eta from the dry-limited 3x3 coarse stencil (sea_level = 0), depth =
 * max(eta - b_fine, 0), momentum components on their own. */
static void interpolate_cell(const gfo_params *p, int nx, int mbc, int maux, const double *qcoarse, double *qfine,
                             const double *aux_coarse, const double *aux_fine, int ic, int jc, int i0, int j0)
{
    const double sea_level = 0.0;
    const int mth = 5;
    for (int mq = 0; mq < 3; mq++) {
        double qc, gradx, grady;
        if (mq == 0) {
            double etab[3][3];
            for (int ii = -1; ii <= 1; ii++)
                for (int jj = -1; jj <= 1; jj++) {
                    const double h = AQ(qcoarse, 0, ic + ii, jc + jj);
                    const double b = AA(aux_coarse, ic + ii, jc + jj);
                    etab[ii + 1][jj + 1] = h < p->dry_tolerance ? sea_level : h + b;
                }
            qc = etab[1][1];
            gradx = gfo_compute_slopes(qc - etab[0][1], etab[2][1] - qc, mth);
            grady = gfo_compute_slopes(qc - etab[1][0], etab[1][2] - qc, mth);
        } else {
            qc = AQ(qcoarse, mq, ic, jc);
            gradx = gfo_compute_slopes(qc - AQ(qcoarse, mq, ic - 1, jc), AQ(qcoarse, mq, ic + 1, jc) - qc, mth);
            grady = gfo_compute_slopes(qc - AQ(qcoarse, mq, ic, jc - 1), AQ(qcoarse, mq, ic, jc + 1) - qc, mth);
        }
        for (int m = 0; m < 4; m++) {
            const int ii = m & 1, jj = m >> 1;
            const double shiftx = (ii - 0.5) / 2.0, shifty = (jj - 0.5) / 2.0;
            const int iff = i0 + ii, jff = j0 + jj;
            if (mq == 0) {
                const double value = qc + gradx * shiftx + grady * shifty - AA(aux_fine, iff, jff);
                AQ(qfine, 0, iff, jff) = value > 0.0 ? value : 0.0;  /* MAX(value, 0.0) */
            } else {
                AQ(qfine, mq, iff, jff) = qc + gradx * shiftx + grady * shifty;
            }
        }
    }
}

/* fc2d_geoclaw_fort_interpolate_face (fc2d_geoclaw_interpolate_fort.f:24-259)
 * and _interpolate_corner (:261-403), side 0..7 as in half_origin: the fine
 * neighbour's ghost cells that the coarse patch's boundary cells cover
 * (faces: every coarse cell along the face whose four fine cells lie in the
 * fine array, 1-mbc .. mx+mbc, which includes the hanging corners; the face
 * loop runs over jc = 1..mx for idir = 0 as the reference does, so mx = my
 * is required).  Coarse ghost cells are read as they are stored. */
int gfo_interpolate_ghost(const gfo_params *p, int mx, int my, int mbc, int maux, const double *qcoarse,
                          double *qfine, const double *aux_coarse, const double *aux_fine, int side, int igrid)
{
    const int nx = mx + 2 * mbc;
    int ox, oy, i2[4], j2[4];
    if (side < 0 || side > 7 || igrid < 0 || igrid > 1 || mx != my) return -1;
    half_origin(side, igrid, mx, my, &ox, &oy);
    if (side >= 4) {
        const int icorner = side - 4;
        const int ic = (icorner & 1) ? mx : 1, jc = (icorner & 2) ? my : 1;
        transform_half(ox, oy, ic, jc, i2, j2);
        interpolate_cell(p, nx, mbc, maux, qcoarse, qfine, aux_coarse, aux_fine, ic, jc, i2[0], j2[0]);
        return 0;
    }
    const int idir = side / 2;
    for (int t = 1; t <= mx; t++) {
        const int ic = idir == 0 ? (side == 0 ? 1 : mx) : t;
        const int jc = idir == 0 ? t : (side == 2 ? 1 : my);
        transform_half(ox, oy, ic, jc, i2, j2);
        int skip = 0;  /* fclaw2d_clawpatch_is_valid_interp: 1-mbc <= i <= mx+mbc, same in j */
        for (int m = 0; m < 4; m++)
            if (!(i2[m] >= 1 - mbc && i2[m] <= mx + mbc && j2[m] >= 1 - mbc && j2[m] <= my + mbc)) skip = 1;
        if (!skip) interpolate_cell(p, nx, mbc, maux, qcoarse, qfine, aux_coarse, aux_fine, ic, jc, i2[0], j2[0]);
    }
    return 0;
}

#undef AQ
#undef AA

/* fc2d_geoclaw_timeinterp_fort.f:1-69 (ng = 0, mint = 2). */
int gfo_timeinterp(int mx, int my, int mbc, int meqn, int psize, const double *qcurr,
                   const double *qlast, double *qinterp, double alpha)
{
    const int nx = mx + 2 * mbc, mint = 2, ng = 0;
    int k = 1;
#define IDX(m, i, j) (((size_t)((j) - 1 + mbc) * nx + ((i) - 1 + mbc)) * meqn + (m))
#define TI(m, i, j) qinterp[IDX(m, i, j)] = qlast[IDX(m, i, j)] + alpha * (qcurr[IDX(m, i, j)] - qlast[IDX(m, i, j)]); k++
    for (int m = 0; m < meqn; m++) {
        for (int j = 1 - ng; j <= my - mint; j++)
            for (int i = 1 - ng; i <= mint; i++) { TI(m, i, j); }
        for (int j = 1 - ng; j <= mint; j++)
            for (int i = mint + 1; i <= mx + ng; i++) { TI(m, i, j); }
        for (int j = mint + 1; j <= my + ng; j++)
            for (int i = mx - mint + 1; i <= mx + ng; i++) { TI(m, i, j); }
        for (int j = my - mint + 1; j <= my + ng; j++)
            for (int i = 1 - ng; i <= mx - mint; i++) { TI(m, i, j); }
    }
#undef TI
#undef IDX
    return (k - 1 != psize) ? 2 : 0;
}
