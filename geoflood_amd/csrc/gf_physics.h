// gf_physics.h -- FP64 device routines of the GeoClaw shallow-water step.
//
// Each routine restates the reference Fortran in the same floating-point
// operation order (the library is built with -ffp-contract=off), in a
// direction-local form: vectors are ordered (h, normal momentum, transverse
// momentum) with respect to the face being solved, which is how rpn2/rpt2
// index them through mu/nv (rpn2_geoclaw.f:101-107, rpt2_geoclaw.f:46-54).
//
// Fortran intrinsics keep LLVM Flang's semantics: MAX(a,b) = a > b ? a : b,
// MIN(a,b) = a < b ? a : b, SIGN(1,x) = copysign(1,x).
#pragma once

#include <hip/hip_runtime.h>

#include "gf_pow.h"

namespace gf {

struct Params {
    double grav;
    double dry_tolerance;
    double manning_n;
    double friction_depth;
    int friction_forcing;
    int src_term;
    int order;       // method(2)
    int transverse;  // method(3)
    int mthlim[3];
    // Formed once on the host from grav and dry_tolerance with the operations
    // the solver would otherwise repeat per face: halfg = 0.5*grav (the .5*g
    // of rpn2 and riemanntype), crit = max(drytol*g, 1d-6), crit2 = sqrt(crit)
    // and critsq = crit*crit (riemann_aug_JCP's criticaltol, criticaltol_2
    // and criticaltol**2, geoclaw_riemann_utils.f:135-136,177), inv16g =
    // 1/(16*grav) (riemanntype's two-rarefaction depth).
    double halfg, crit, crit2, critsq, inv16g;
};

__device__ __forceinline__ double fmaxf64(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double fminf64(double a, double b) { return a < b ? a : b; }

// riemanntype (geoclaw_riemann_utils.f:566-677) with maxiter = 1; returns
// hm and the middle-state speeds s1m, s2m (rare1/rare2 are unused by callers).
// sgL, sgR are sqrt(g*hL), sqrt(g*hR), which the caller has already formed:
// every sqrt(g*h) of hL, hR, h_min or h_max below is one of them, bit for bit.
__device__ __forceinline__ void riemanntype(double hL, double hR, double uL, double uR,
                                            double sgL, double sgR, double drytol, double g, double halfg, double inv16g,
                                            double& hm, double& s1m, double& s2m)
{
    const double h_min = fminf64(hR, hL);
    const double h_max = fmaxf64(hR, hL);
    const double sg_min = hR < hL ? sgR : sgL;  // sqrt(g * h_min)
    const double sg_max = hR > hL ? sgR : sgL;  // sqrt(g * h_max)
    const double delu = uR - uL;
    if (h_min <= drytol) {
        hm = 0.0;
        s1m = uR + uL - 2.0 * sgR + 2.0 * sgL;
        s2m = s1m;
        return;
    }
    const double F_min = delu + 2.0 * (sg_min - sg_max);
    const double F_max = delu + (h_max - h_min) * (sqrt(halfg * (h_max + h_min) / (h_max * h_min)));
    // The three branches leave hm (and, for two shocks, u1m, u2m); sqrt(g*hm)
    // and the speeds are formed once after them, so that a wave whose lanes
    // take different branches pays for the shared part only once.
    double u1m = 0.0, u2m = 0.0;
    int kind;  // 0: two rarefactions, 1: two shocks, 2: one shock, one rarefaction
    if (F_min > 0.0) {
        const double t = fmaxf64(0.0, -delu + 2.0 * (sgL + sgR));
        hm = inv16g * (t * t);
        kind = 0;
    } else if (F_max <= 0.0) {  // one Newton step on sqrt(h)
        double h0 = h_max;
        const double gL = sqrt(halfg * (1 / h0 + 1 / hL));
        const double gR = sqrt(halfg * (1 / h0 + 1 / hR));
        const double F0 = delu + (h0 - hL) * gL + (h0 - hR) * gR;
        const double dfdh = gL - g * (h0 - hL) / (4.0 * (h0 * h0) * gL) + gR -
                            g * (h0 - hR) / (4.0 * (h0 * h0) * gR);
        const double slope = 2.0 * sqrt(h0) * dfdh;
        const double t = sqrt(h0) - F0 / slope;
        h0 = t * t;
        hm = h0;
        u1m = uL - (hm - hL) * sqrt(halfg * (1 / hm + 1 / hL));
        u2m = uR + (hm - hR) * sqrt(halfg * (1 / hm + 1 / hR));
        kind = 1;
    } else {
        // F0 = delu + 2(sqrt(g h0) - sqrt(g h_max)) + (h0 - h_min) sqrt(g/2 (1/h0 + 1/h_min))
        // at h0 = h_min: the last product is (+0) times a finite positive
        // whenever g/h_min cannot overflow, so F0 is F_min + 0.0 exactly
        double h0 = h_min;
        const double F0 = (h_min >= 1e-150 && g <= 1e150)
                              ? F_min + 0.0
                              : delu + 2.0 * (sg_min - sg_max) + (h0 - h_min) * sqrt(halfg * (1 / h0 + 1 / h_min));
        const double slope = (F_max - F0) / (h_max - h_min);
        h0 = h0 - F0 / slope;
        hm = h0;
        kind = 2;
    }
    const double sghm = sqrt(g * hm);
    // two rarefactions:      s1m = (uL + 2 sqrt(g hL)) - 3 sghm, s2m = (uR - 2 sqrt(g hR)) + 3 sghm
    // shock-rarefaction:     hL > hR: a = uL + 2 sqrt(g hL), s1m = a - 3 sghm, s2m = a - sghm
    //                        else:    a = uR - 2 sqrt(g hR), s2m = a + 3 sghm, s1m = a + sghm
    // two shocks:            s1m = u1m - sghm, s2m = u2m + sghm
    const double A = uL + 2.0 * sgL, B = uR - 2.0 * sgR, s3 = 3.0 * sghm;
    if (kind == 1) {
        s1m = u1m - sghm;
        s2m = u2m + sghm;
    } else if (kind == 0 || hL > hR) {
        s1m = A - s3;
        s2m = kind == 0 ? B + s3 : A - sghm;
    } else {
        s1m = B + sghm;
        s2m = B + s3;
    }
}

// hm of riemanntype(h, h, u, -u), the wall test of rpn2 (rpn2_geoclaw.f:
// 185-242), whose speeds the caller does not use.  With hL = hR = h every
// (h_max - h_min) factor is an exact +0: F_min = F_max = delu + 0.0, a
// positive F takes the two-rarefaction branch and any other the two-shock
// branch with gL = gR and dfdh = gL + gL.  Outside the range where those
// zeros are exact (g/h overflowing) the general routine is used.
__device__ __forceinline__ double wall_hstar(double h, double u, double sg, double drytol, double g, double halfg, double inv16g)
{
    if (!(h >= 1e-150 && g <= 1e150) || h <= drytol) {
        double hm, s1m, s2m;
        riemanntype(h, h, u, -u, sg, sg, drytol, g, halfg, inv16g, hm, s1m, s2m);
        return hm;
    }
    const double delu = -u - u;
    const double F = delu + 0.0;
    if (F > 0.0) {
        const double t = fmaxf64(0.0, -delu + 2.0 * (sg + sg));
        return inv16g * (t * t);
    }
    const double gL = sqrt(halfg * (1 / h + 1 / h));
    const double F0 = (delu + 0.0) + 0.0;
    const double dfdh = gL + gL;
    const double slope = 2.0 * sqrt(h) * dfdh;
    const double t = sqrt(h) - F0 / slope;
    return t * t;
}

// One face of rpn2 (rpn2_geoclaw.f:84-330) followed by riemann_aug_JCP
// (geoclaw_riemann_utils.f:2-311, maxiter = 1, no rarefaction corrector).
// Inputs are the left/right states in direction-local order and the left/right
// bathymetry.  Outputs: speeds s[3] and f-waves w[mw][k] with k in
// direction-local order.  Returns false when both sides are dry (the
// reference leaves s and the f-waves at zero).
// With RC, rc holds the right-hand cell's huR/hR, hvR/hR and sqrt(g*hR) as the
// caller formed them for a wet cell (the same operations on the same values).
struct RightCell {
    double u, v, sg;
};

template <bool RC = false>
__device__ __forceinline__ bool rpn2_face(double hL, double huL, double hvL, double bL,
                                          double hR, double huR, double hvR, double bR,
                                          const Params& P,
                                          double s[3], double w[3][3], const RightCell& rc = RightCell{})
{
    const double g = P.grav, drytol = P.dry_tolerance, halfg = P.halfg;
    if (hL <= drytol && hR <= drytol) {
#pragma unroll
        for (int mw = 0; mw < 3; mw++) {
            s[mw] = 0.0;
            w[mw][0] = 0.0; w[mw][1] = 0.0; w[mw][2] = 0.0;
        }
        return false;
    }
    double uR, vR, phiR, uL, vL, phiL;
    const bool r_wet = hR > drytol;  // then hR keeps its value below
    if (hR > drytol) {
        uR = RC ? rc.u : huR / hR;
        vR = RC ? rc.v : hvR / hR;
        phiR = halfg * (hR * hR) + (huR * huR) / hR;
    } else {
        hR = 0.0; huR = 0.0; hvR = 0.0; uR = 0.0; vR = 0.0; phiR = 0.0;
    }
    if (hL > drytol) {
        uL = huL / hL;
        vL = hvL / hL;
        phiL = halfg * (hL * hL) + (huL * huL) / hL;
    } else {
        hL = 0.0; huL = 0.0; hvL = 0.0; uL = 0.0; vL = 0.0; phiL = 0.0;
    }

    double wall0 = 1.0, wall1 = 1.0, wall2 = 1.0;
    // the wall test of whichever side is dry: riemanntype(hL, hL, uL, -uL) or
    // riemanntype(hR, hR, -uR, uR), one call for both kinds of lane
    double hstar = 0.0;
    if (hR <= drytol || hL <= drytol) {
        const double hw = hR <= drytol ? hL : hR;
        const double uw = hR <= drytol ? uL : -uR;
        hstar = wall_hstar(hw, uw, sqrt(g * hw), drytol, g, halfg, P.inv16g);
    }
    if (hR <= drytol) {
        const double hstartest = fmaxf64(hL, hstar);
        if (hstartest + bL < bR) {
            wall1 = 0.0; wall2 = 0.0;
            hR = hL; huR = -huL; bR = bL; phiR = phiL; uR = -uL; vR = vL;
        } else if (hL + bL < bR) {
            bR = hL + bL;
        }
    } else if (hL <= drytol) {
        const double hstartest = fmaxf64(hR, hstar);
        if (hstartest + bR < bL) {
            wall0 = 0.0; wall1 = 0.0;
            hL = hR; huL = -huR; bL = bR; phiL = phiR; uL = -uR; vL = vR;
        } else if (hR + bR < bL) {
            bL = hR + bR;
        }
    }

    const double sghL = sqrt(g * hL);
    const double sghR = (RC && r_wet) ? rc.sg : sqrt(g * hR);
    const double sL = uL - sghL;
    const double sR = uR + sghR;
    const double uhat = (sghL * uL + sghR * uR) / (sghR + sghL);
    const double chat = sqrt(halfg * (hR + hL));
    const double sRoe1 = uhat - chat;
    const double sRoe2 = uhat + chat;
    double sE1 = fminf64(sL, sRoe1);
    double sE2 = fmaxf64(sR, sRoe2);

    // ---- riemann_aug_JCP ----
    const double delh = hR - hL;
    const double delhu = huR - huL;
    const double delphi = phiR - phiL;
    const double delb = bR - bL;

    double hm, s1m, s2m;
    riemanntype(hL, hR, uL, uR, sghL, sghR, drytol, g, halfg, P.inv16g, hm, s1m, s2m);

    const double lam0 = fminf64(sE1, s2m);
    const double lam2 = fmaxf64(sE2, s1m);
    sE1 = lam0;
    sE2 = lam2;
    const double hstarHLL = fmaxf64((huL - huR + sE2 * hR - sE1 * hL) / (sE2 - sE1), 0.0);
    const double lam1 = 0.5 * (lam0 + lam2);

    // r(:,1) = (1, lam0, lam0^2), r(:,2) = (0, 0, 1), r(:,3) = (1, lam2, lam2^2)
    const double r10 = lam0, r20 = lam0 * lam0;
    const double r12 = lam2, r22 = lam2 * lam2;

    const double criticaltol = P.crit;
    const double criticaltol_2 = P.crit2;

    const double hbar = fmaxf64(0.5 * (hL + hR), 0.0);
    const double tu = uL + uR;
    const double s1s2bar = 0.25 * (tu * tu) - g * hbar;
    const double s1s2tilde = fmaxf64(0.0, uL * uR) - g * hbar;

    bool sonic;
    if (fabs(s1s2bar) <= criticaltol) sonic = true;
    else if (s1s2bar * s1s2tilde <= P.critsq) sonic = true;
    else if (s1s2bar * sE1 * sE2 <= P.critsq) sonic = true;
    else if (fminf64(fabs(sE1), fabs(sE2)) < criticaltol_2) sonic = true;
    else if (sE1 < criticaltol_2 && s1m > -criticaltol_2) sonic = true;
    else if (sE2 > -criticaltol_2 && s2m < criticaltol_2) sonic = true;
    else if ((uL + sghL) * (uR + sghR) < 0.0) sonic = true;
    else if ((uL - sghL) * (uR - sghR) < 0.0) sonic = true;
    else sonic = false;

    double deldelh = sonic ? -delb : delb * g * hbar / s1s2bar;
    // the bounds' two quotients are formed once for every lane (the three
    // cases use them in different combinations)
    const double q2 = hstarHLL * (sE2 - sE1) / sE2;
    const double q1 = hstarHLL * (sE2 - sE1) / sE1;
    if (sE1 < -criticaltol && sE2 > criticaltol) {
        deldelh = fminf64(deldelh, q2);
        deldelh = fmaxf64(deldelh, q1);
    } else if (sE1 >= criticaltol) {
        deldelh = fminf64(deldelh, q1);
        deldelh = fmaxf64(deldelh, -hL);
    } else if (sE2 <= -criticaltol) {
        deldelh = fminf64(deldelh, hR);
        deldelh = fmaxf64(deldelh, q2);
    }
    double deldelphi = sonic ? -g * hbar * delb : -delb * g * hbar * s1s2tilde / s1s2bar;
    deldelphi = fminf64(deldelphi, g * fmaxf64(-hL * delb, -hR * delb));
    deldelphi = fmaxf64(deldelphi, g * fminf64(-hL * delb, -hR * delb));

    const double d0 = delh - deldelh;
    const double d1 = delhu;
    const double d2 = delphi - deldelphi;

    // Cramer's rule exactly as the reference writes it (geoclaw_riemann_utils.f:230-249)
    // with r = [[1,0,1],[r10,0,r12],[r20,1,r22]].
    double det1 = 1.0 * (0.0 * r22 - r12 * 1.0);
    double det2 = 0.0 * (r10 * r22 - r12 * r20);
    double det3 = 1.0 * (r10 * 1.0 - 0.0 * r20);
    const double determinant = det1 - det2 + det3;
    double beta0, beta1, beta2;
    det1 = d0 * (0.0 * r22 - r12 * 1.0);
    det2 = 0.0 * (d1 * r22 - r12 * d2);
    det3 = 1.0 * (d1 * 1.0 - 0.0 * d2);
    beta0 = (det1 - det2 + det3) / determinant;
    det1 = 1.0 * (d1 * r22 - r12 * d2);
    det2 = d0 * (r10 * r22 - r12 * r20);
    det3 = 1.0 * (r10 * d2 - d1 * r20);
    beta1 = (det1 - det2 + det3) / determinant;
    det1 = 1.0 * (0.0 * d2 - d1 * 1.0);
    det2 = 0.0 * (r10 * d2 - d1 * r20);
    det3 = d0 * (r10 * 1.0 - 0.0 * r20);
    beta2 = (det1 - det2 + det3) / determinant;

    // f-waves (geoclaw_riemann_utils.f:289-306)
    double fw00 = beta0 * r10, fw01 = beta0 * r20, fw02 = beta0 * r10;
    double fw10 = beta1 * 0.0, fw11 = beta1 * 1.0, fw12;
    double fw20 = beta2 * r12, fw21 = beta2 * r22, fw22 = beta2 * r12;
    fw02 = fw02 * vL;
    fw22 = fw22 * vR;
    fw12 = 0.0;
    const double hustar_interface = huL + fw00;
    if (hustar_interface <= 0.0)
        fw02 = fw02 + (hR * uR * vR - hL * uL * vL - fw02 - fw22);
    else
        fw22 = fw22 + (hR * uR * vR - hL * uL * vL - fw02 - fw22);

    // eliminate ghost fluxes for the wall problem (rpn2_geoclaw.f:303-309)
    s[0] = lam0 * wall0;
    s[1] = lam1 * wall1;
    s[2] = lam2 * wall2;
    w[0][0] = fw00 * wall0; w[0][1] = fw01 * wall0; w[0][2] = fw02 * wall0;
    w[1][0] = fw10 * wall1; w[1][1] = fw11 * wall1; w[1][2] = fw12 * wall1;
    w[2][0] = fw20 * wall2; w[2][1] = fw21 * wall2; w[2][2] = fw22 * wall2;
    return true;
}

// Fluctuations of one face (rpn2_geoclaw.f:359-376), direction-local order.
__device__ __forceinline__ void fluctuations(const double s[3], const double w[3][3],
                                             double amdq[3], double apdq[3])
{
    amdq[0] = 0.0; amdq[1] = 0.0; amdq[2] = 0.0;
    apdq[0] = 0.0; apdq[1] = 0.0; apdq[2] = 0.0;
#pragma unroll
    for (int mw = 0; mw < 3; mw++) {
        if (s[mw] < 0.0) {
#pragma unroll
            for (int k = 0; k < 3; k++) amdq[k] = amdq[k] + w[mw][k];
        } else if (s[mw] > 0.0) {
#pragma unroll
            for (int k = 0; k < 3; k++) apdq[k] = apdq[k] + w[mw][k];
        } else {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                amdq[k] = amdq[k] + 0.5 * w[mw][k];
                apdq[k] = apdq[k] + 0.5 * w[mw][k];
            }
        }
    }
}

// philim (philim.f:65-126); an out-of-range method falls through to minmod.
__device__ __forceinline__ double philim(double a, double b, int meth)
{
    if (a == 0.0) return 0.0;
    const double r = b / a;
    switch (meth) {
    case 2:
        return fmaxf64(fmaxf64(0.0, fminf64(1.0, 2.0 * r)), fminf64(2.0, r));
    case 3:
        return (r + fabs(r)) / (1.0 + fabs(r));
    case 4: {
        const double c = (1.0 + r) / 2.0;
        return fmaxf64(0.0, fminf64(fminf64(c, 2.0), 2.0 * r));
    }
    case 5:
        return r;
    default:
        return fmaxf64(0.0, fminf64(1.0, r));
    }
}

// Eigen-split of a transverse fluctuation for one cell (rpt2_geoclaw.f:61-187):
// the cell's state (h, un, vt in direction-local order with respect to the
// normal solve), its surface eta = h + b and the bathymetry of its two
// transverse neighbours.  Applies the split to two fluctuations at once (the
// imp = 1 split of amdq of the face on the cell's right/top and the imp = 2
// split of apdq of the face on its left/bottom share the same Jacobian).
struct Rpt2Cell {
    bool active;
    double u, s0, s1, s2, eta, topo1, topo3;
};

// The same with the cell's hun/h, hvt/h and sqrt(g*h) formed by the caller
// (used only when h > tol, where they are what rpt2_setup would form).
__device__ __forceinline__ Rpt2Cell rpt2_setup_c(double h, double u_c, double v_c, double sg_c, double b,
                                                 double topo1, double topo3, double tol)
{
    Rpt2Cell c;
    c.active = false;
    c.u = 0.0; c.s0 = 0.0; c.s1 = 0.0; c.s2 = 0.0; c.eta = 0.0; c.topo1 = topo1; c.topo3 = topo3;
    if (h <= tol) return c;
    const double eta = h + b;
    if (eta < fminf64(topo1, topo3)) return c;
    c.active = true;
    c.u = u_c;
    c.s0 = v_c - sg_c;
    c.s1 = v_c;
    c.s2 = v_c + sg_c;
    c.eta = eta;
    return c;
}

__device__ __forceinline__ Rpt2Cell rpt2_setup(double h, double hun, double hvt, double b,
                                               double topo1, double topo3, double g, double tol)
{
    Rpt2Cell c;
    c.active = false;
    c.u = 0.0; c.s0 = 0.0; c.s1 = 0.0; c.s2 = 0.0; c.eta = 0.0; c.topo1 = topo1; c.topo3 = topo3;
    if (h <= tol) return c;
    const double u = hun / h;
    const double v = hvt / h;
    const double eta = h + b;
    if (eta < fminf64(topo1, topo3)) return c;
    const double sgh = sqrt(g * h);
    c.active = true;
    c.u = u;
    c.s0 = v - sgh;
    c.s1 = v;
    c.s2 = v + sgh;
    c.eta = eta;
    return c;
}

// bmasdq/bpasdq in direction-local order for one asdq.
__device__ __forceinline__ void rpt2_apply(const Rpt2Cell& c, const double asdq[3],
                                           double bm[3], double bp[3])
{
    bm[0] = 0.0; bm[1] = 0.0; bm[2] = 0.0;
    bp[0] = 0.0; bp[1] = 0.0; bp[2] = 0.0;
    if (!c.active) return;
    const double delf1 = asdq[0], delf2 = asdq[1], delf3 = asdq[2];
    const double beta0 = (c.s2 * delf1 - delf3) / (c.s2 - c.s0);
    const double beta1 = -c.u * delf1 + delf2;
    const double beta2 = (delf3 - c.s0 * delf1) / (c.s2 - c.s0);
    const double sv[3] = {c.s0, c.s1, c.s2};
    const double bv[3] = {beta0, beta1, beta2};
    // r(:,1) = (1, u, s0), r(:,2) = (0, 1, 0), r(:,3) = (1, u, s2)
    const double r0[3] = {1.0, 0.0, 1.0};
    const double r1[3] = {c.u, 1.0, c.u};
    const double r2[3] = {c.s0, 0.0, c.s2};
#pragma unroll
    for (int mw = 0; mw < 3; mw++) {
        if ((sv[mw] < 0.0) && (c.eta >= c.topo1)) {
            bm[0] = bm[0] + 1.0 * sv[mw] * bv[mw] * r0[mw];
            bm[1] = bm[1] + 1.0 * sv[mw] * bv[mw] * r1[mw];
            bm[2] = bm[2] + 1.0 * sv[mw] * bv[mw] * r2[mw];
        } else if ((sv[mw] > 0.0) && (c.eta >= c.topo3)) {
            bp[0] = bp[0] + 1.0 * sv[mw] * bv[mw] * r0[mw];
            bp[1] = bp[1] + 1.0 * sv[mw] * bv[mw] * r1[mw];
            bp[2] = bp[2] + 1.0 * sv[mw] * bv[mw] * r2[mw];
        }
    }
}

// Manning friction of fc2d_geoclaw_src2 (geoclaw_src2_fort.f90:41-71) for one
// interior cell; manning_break(1) = rinfinity (default-real 10.e32).
__device__ __forceinline__ void src2_cell(double h, double& hu, double& hv, double b,
                                          double dt, const Params& P)
{
    if (!P.friction_forcing) return;
    if (h < 1.0e-30) {
        hu = 0.0;
        hv = 0.0;
    } else if (h <= P.friction_depth) {
        const double coeff = (b < (double)10.e32f) ? P.manning_n : 0.0;
        const double gamma = sqrt(hu * hu + hv * hv) * P.grav * (coeff * coeff) / pow_7_3(h);
        const double dgamma = 1.0 + dt * gamma;
        hu = hu / dgamma;
        hv = hv / dgamma;
    }
}

}  // namespace gf
