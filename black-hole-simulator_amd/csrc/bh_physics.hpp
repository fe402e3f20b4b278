// Pointwise GRMHD physics shared by the gfx950 kernels and the host-side
// tree code (initial data, regrid).  Compiled for both targets from this one
// source with contraction off, so host and device round identically.
//
// Geometry enters as precomputed metric values (the metric is static); no
// transcendental function is evaluated here except sqrt, which is correctly
// rounded on both targets.  Each routine keeps the reference's operation
// order so that results match it bit for bit; reference quirks are kept and
// marked "quirk" with the reference file:line.
#pragma once

#include <cfloat>

#include "bh_common.hpp"

namespace bh {

struct Prim { double rho, eps, v[3], B[3]; };
struct Cons { double D, S[3], tau, B[3]; };

BH_HD Cons czero() {
    Cons c;
    c.D = 0.0; c.tau = 0.0;
    for (int i = 0; i < 3; i++) { c.S[i] = 0.0; c.B[i] = 0.0; }
    return c;
}

// Tree-metric values (Kerr-Newman-Schild with the run's a, Q) at one point.
// The zero entries of gamma_ij / gamma^ij are exact +0.0 in the reference
// (spacetime.cpp:20-49 zero the arrays, then set the non-zero entries).
struct GPt {
    double sqrtdetg, sqrtdetgam, alpha, beta0;
    double g11, g13, g22, g33;  // gamma_ij = g_{i+1,j+1}
    double gi13, gi33;          // gamma^{02}, gamma^{22}
};

// Cell-metric values: every reference cell carries a default-constructed
// metric (M=1, a=0, Q=0), cell.hpp:26 / metric.hpp:30 (quirk Q2).
// gamma_02 = -H*0*sin^2 = -0.0 and beta^1 = beta^2 = +0.0 exactly.
struct GCell {
    double sqrtdetg, alpha, beta0;
    double g11, g22, g33;        // gamma_ij
    double gi00, gi01, gi11, gi22, gi33;  // g^{mu nu}
};
constexpr double CELL_G13 = -0.0;

// ----------------------------------------------------------------- EOS
// Ideal gas, thermo.cpp:10-35.  rho <= 0 throws std::domain_error in the
// reference; here it is counted in nf and the arithmetic continues.
BH_HD double eos_press(double gm, double rho, double eps, int& nf) {
    nf += (rho <= 0.0);
    return (gm - 1.0) * rho * eps;
}
BH_HD double eos_enth(double gm, double rho, double eps, int& nf) {
    nf += (rho <= 0.0);
    return 1.0 + gm * eps;
}
BH_HD double eos_temp(double rho, double p, int& nf) {
    nf += (rho <= 0.0);
    return p / rho;
}
BH_HD double eos_cs2(double gm, double rho, double eps, int& nf) {
    nf += (rho <= 0.0);
    double p = eos_press(gm, rho, eps, nf);
    double h = eos_enth(gm, rho, eps, nf);
    return (gm * p) / (rho * h);
}

// q / l for a numerator that is often exactly zero (the magnetic terms in
// field-free regions).  When q is +-0 and l finite and non-zero in every
// lane of the wavefront, the quotient is the signed zero IEEE division gives
// (the sign of q flipped for negative l) and the division sequence is
// skipped; otherwise every lane divides.  Same bits as q / l in all cases.
BH_HD double div_z(double q, double l) {
#if defined(__HIP_DEVICE_COMPILE__)
    if (__all(q == 0.0 && l != 0.0 && fabs(l) <= DBL_MAX)) return l > 0.0 ? q : -q;
#endif
    return q / l;
}

// ------------------------------------------------------------- ptoc
// conserved::ptoc, conserved.cpp:17-73 (quirk Q7: b^2 divided by W).
BH_HD Cons ptoc(const GPt& m, double gm, const Prim& pv, int& nf) {
    const double gam[3][3] = {{m.g11, 0.0, m.g13}, {0.0, m.g22, 0.0}, {m.g13, 0.0, m.g33}};
    double sg = m.sqrtdetgam;
    double v2 = 0.0;
    double vlow[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) vlow[i] += gam[i][j] * pv.v[j];
        v2 += vlow[i] * pv.v[i];
    }
    v2 = smin(v2, 1.0 - 1e-10);
    double lor = 1.0 / sqrt(1.0 - v2);
    double p = eos_press(gm, pv.rho, pv.eps, nf);
    double h = eos_enth(gm, pv.rho, pv.eps, nf);
    double Blow[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Blow[i] += gam[i][j] * pv.B[j];
    double Bv = 0.0;
    for (int i = 0; i < 3; i++) Bv += Blow[i] * pv.v[i];
    double Bsq = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Bsq += Blow[i] * gam[i][j] * pv.B[j];
    double b2 = (Bv * Bv + Bsq) / lor;
    double bt = lor * Bv / m.alpha;
    double pstar = p + b2 / 2;
    double w = pv.rho * h + b2;
    Cons cv;
    cv.D = sg * pv.rho * lor;
    cv.tau = sg * (w * lor * lor - pstar - bt * bt) - cv.D;
    for (int i = 0; i < 3; i++) {
        double bi = (Blow[i] + m.alpha * bt * lor * vlow[i]) / lor;
        cv.S[i] = sg * (w * lor * lor * vlow[i] - m.alpha * bt * bi);
        cv.B[i] = sg * pv.B[i];
    }
    return cv;
}

// ------------------------------------------------------------- ctop
// Energy residual and derivative, conserved.cpp:80-123 (quirk Q8).
BH_HD double c2p_fdf(double gm, double xi, double Ssq, double Bsq, double SB, double Dn, double taun, double& df) {
    double Yi = SB * SB;
    double K = Bsq;
    double X = xi * xi;
    double xipK = xi + K;
    double vsq = (Ssq * X + Yi * (2.0 * xi + K)) / (X + xipK * xipK);
    if (vsq >= 1.0) { df = 0.0; return 1e30; }
    double ltz2 = 1.0 / (1.0 - vsq);
    double ltz = sqrt(ltz2);
    double rho = Dn / ltz;
    double G = gm - 1.0;
    double p = (G / gm) * (xi / ltz2 - Dn / ltz);
    double pstar = p + div_z(K, 2 * ltz2) + div_z(Yi, 2 * X);
    if (p < 0.0 || rho < 0.0) { df = 0.0; return 1e32; }
    double f = xi - taun - Dn - pstar;
    double N = Ssq * X + Yi * (2.0 * xi + K);
    double Mq = X * xipK * xipK;
    double dN = 2.0 * Ssq * xi + 2.0 * Yi;
    double dM = 2.0 * xi * xipK * xipK + 2.0 * X * X;
    double dvv = (dN * Mq - dM * N) / (Mq * Mq);
    double dltz2 = ltz2 * ltz2 * dvv;
    double dltz = dltz2 / (2.0 * ltz);
    double dp = (G / gm) * ((1.0 / ltz2) - (xi * dltz2) / (ltz2 * ltz2) + Dn * dltz / ltz2);
    double dpstar = dp - div_z(K * dltz2, 2.0 * ltz2 * ltz2) - div_z(Yi, xi * X);
    df = 1 - dpstar;
    return f;
}
// Residual only (conserved::fres).  The derivative tail of c2p_fdf is dead
// code here and drops out after inlining.
BH_HD double c2p_f(double gm, double xi, double Ssq, double Bsq, double SB, double Dn, double taun) {
    double Yi = SB * SB;
    double K = Bsq;
    double X = xi * xi;
    double xipK = xi + K;
    double vsq = (Ssq * X + Yi * (2.0 * xi + K)) / (X + xipK * xipK);
    if (vsq >= 1.0) return 1e30;
    double ltz2 = 1.0 / (1.0 - vsq);
    double ltz = sqrt(ltz2);
    double rho = Dn / ltz;
    double G = gm - 1.0;
    double p = (G / gm) * (xi / ltz2 - Dn / ltz);
    double pstar = p + div_z(K, 2 * ltz2) + div_z(Yi, 2 * X);
    if (p < 0.0 || rho < 0.0) return 1e32;
    return xi - taun - Dn - pstar;
}

// Reciprocal and square root with relative error below 2^-50 (device: the
// hardware estimate refined by two Newton steps; host: the correctly rounded
// operations).  Only used where an error bound covers them.
BH_HD double approx_rcp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
#else
    return 1.0 / x;
#endif
}
BH_HD double approx_sqrt(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double r = __builtin_amdgcn_rsq(x);
    double g = x * r, h = 0.5 * r;
    double e = fma(-g, h, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    e = fma(-g, h, 0.5);
    return fma(g, e, g);
#else
    return sqrt(x);
#endif
}

// Sign of c2p_f(gm, xi, ...) without its five divisions and square root:
// +1 when c2p_f returns a value > 0 (its 1e30 and 1e32 exits included), -1
// when it returns a value < 0, 0 when the cheap evaluation cannot prove
// which (the caller then evaluates c2p_f).
//
// The numerator and denominator of v^2 are formed with c2p_f's own
// operations (same bits).  From there the quantities carry approximation
// error: v^2 within 2^-49.8 relative, w = 1 - v^2 within 2^-49.7 (1 + kappa),
// kappa = v^2 / w (cancellation in w), and every later term within a few
// 2^-53 more; c2p_f's own roundings are of the same order.  Each decision
// below requires a margin of 2^-44 (1 + kappa) times the magnitudes that
// enter it (32x the first-order bound) plus 2^-1000 (so that a sign cannot
// be lost to underflow); kappa is capped at 2^30 so that first order holds.
BH_HD int c2p_f_sign(double gm, double xi, double Ssq, double Bsq, double SB, double Dn, double taun) {
    if (!(Dn >= 0.0)) return 0;  // rho = Dn / ltz < 0, or NaN: exact path
    double Yi = SB * SB;
    double K = Bsq;
    double X = xi * xi;
    double xipK = xi + K;
    const double num = Ssq * X + Yi * (2.0 * xi + K);
    const double den = X + xipK * xipK;
    if (!(den >= 0x1p-900 && den <= 0x1p+900 && num >= 0.0 && num <= 0x1p+900)) return 0;
    const double vsq = num * approx_rcp(den);
    if (!(vsq < 1.0 - 0x1p-46)) return vsq >= 1.0 + 0x1p-46 ? 1 : 0;  // c2p_f returns 1e30 when vsq >= 1
    const double w = 1.0 - vsq;                                       // 1 / ltz2
    const double kappa = vsq * approx_rcp(w);
    if (!(kappa <= 0x1p+30)) return 0;
    const double tol = 0x1p-44 * (1.0 + kappa);
    const double t1 = xi * w;                // xi / ltz2
    const double t2 = Dn * approx_sqrt(w);   // Dn / ltz
    const double d = t1 - t2;                // sign of p
    const double ep = tol * (fabs(t1) + fabs(t2)) + 0x1p-1000;
    if (d < -ep) return 1;                   // p < 0: c2p_f returns 1e32
    if (!(d > ep)) return 0;
    const double c = (gm - 1.0) / gm;
    const double pr = c * d;
    const double t3 = K * w * 0.5;           // K / (2 ltz2)
    if (!(X >= 0x1p-900 && X <= 0x1p+900)) return 0;
    const double t4 = Yi * approx_rcp(2.0 * X);  // Yi / (2 X)
    const double pstar = pr + t3 + t4;
    const double f = xi - taun - Dn - pstar;
    const double e = tol * (fabs(c) * (fabs(t1) + fabs(t2)) + fabs(t3)) +
                     0x1p-44 * (fabs(t4) + fabs(pr) + fabs(t3) + fabs(pstar) + fabs(f)) + 0x1p-1000;
    if (!(fabs(f) <= 0x1p+1000 && e <= 0x1p+1000)) return 0;
    if (f > e) return 1;
    if (f < -e) return -1;
    return 0;
}
// c2p_f where only its sign is used: the sign as +-1 when c2p_f_sign proves
// it, else c2p_f itself.
BH_HD double c2p_f_signed(double gm, double xi, double Ssq, double Bsq, double SB, double Dn, double taun) {
    const int sg = c2p_f_sign(gm, xi, Ssq, Bsq, SB, Dn, taun);
    return sg ? (double)sg : c2p_f(gm, xi, Ssq, Bsq, SB, Dn, taun);
}

BH_HD void dswap(double& a, double& b) { double t = a; a = b; b = t; }

// The Brent loop below when Ssq = Bsq = SB = 0 (no momentum or field in the
// components the inversion reads, quirk Q6), the usual state of the
// atmosphere.  Then vsq = +0, ltz = 1, rho = D and the residual evaluates
// exactly as f(xi) = ((xi - tau) - D) - c (xi - D), c = (gm - 1)/gm, whose
// rounding error is below 7u (xi + |tau| + D), u = 2^-53 (five roundings,
// each at most u times a term bounded by xi + |tau| + D).  So for xi above
// the last midpoint s_n of the loop the residual is positive once
// (s_n - D)(1 - c) - tau > 7u (s_n + |tau| + D); the condition tested
// below, with 32u, also covers the rounding of its own evaluation.  With
// fa > 0 and xb < xa every midpoint then moves xa (fa * fs < 0 is false),
// xb never moves, and the loop ends on root = xb after the same number of
// halvings, or fails after 100; the midpoints themselves are replayed
// exactly, and so is the test that keeps the step a bisection.  Returns 1
// (root = xb), 0 (the 100 passes end without convergence), or -1 when the
// argument does not apply and the loop must run.
BH_HD int c2p_brent_atmosphere(double gm, double xa, double xb, double D, double tau, double tol, double& root) {
    const double c = (gm - 1.0) / gm;
    // Short form.  For 0 < xb < 2^-16 every pass halves the width w = xa - xb
    // up to a relative 2^-30 (the midpoint's rounding is below 2^-53 (xa + xb)
    // and 2^-53 xb << tol <= w), so from w <= 2^80 tol the loop converges
    // within 82 passes; each midpoint exceeds xb by at least ~tol/2 (the last
    // one halves a width >= tol), so the residual test at xb + 0.49 tol covers
    // all of them; and (3 xa + 4)/4 >= 1 > xb keeps every pass a bisection.
    if (xb > 0.0 && xb < 0x1p-16 && xa - xb <= 0x1p80 * tol) {
        const double s = xb + 0.49 * tol;
        if (!((s - D) * (1.0 - c) - tau > 0x1p-48 * (fabs(s) + fabs(tau) + D))) return -1;
        root = xb;
        return 1;
    }
    double a = xa, s = xa;
    int n = 0;
    bool converged = false;
    for (int i = 0; i < 100; i++) {
        if (fabs(xb - a) < tol) { converged = true; break; }
        if ((3 * a + 4) / 4 < xb) return -1;  // the interpolant could be taken
        s = (a + xb) / 2;
        a = s;
        n = i + 1;
    }
    // the midpoints decrease from xa towards xb, so s is the smallest
    if (n > 0 && !((s - D) * (1.0 - c) - tau > 0x1p-48 * (fabs(s) + fabs(tau) + D))) return -1;
    if (!converged) return 0;
    root = xb;
    return 1;
}

// Brent-Dekker fallback, conserved.cpp:232-285 (quirk Q13: bisection test
// (3a+4)/4; fa and fb are never refreshed inside the loop).
BH_HD bool c2p_brent(double gm, double xa, double xb, double Ssq, double Bsq, double SB, double D, double tau,
                     double& root, double tol) {
    double fa = c2p_f(gm, xa, Ssq, Bsq, SB, D, tau);
    double fb = c2p_f(gm, xb, Ssq, Bsq, SB, D, tau);
    if (fa * fb > 0.0) return false;
    if (fabs(fa) < fabs(fb)) { dswap(xa, xb); dswap(fa, fb); }
    if (Ssq == 0.0 && Bsq == 0.0 && SB == 0.0 && fa > 0.0 && D > 0.0 && xb < xa && isfinite(xa) &&
        isfinite(tau)) {
        const int r = c2p_brent_atmosphere(gm, xa, xb, D, tau, tol, root);
        if (r >= 0) return r == 1;
    }
    double xc = xa, fc = fa;
    bool mflag = true;
    double s = 0.0, d = 0.0;
    for (int i = 0; i < 100; i++) {
        if (fabs(xb - xa) < tol) { root = xb; return true; }
        // The interpolated point is kept only if lo < s < xb.  When lo < xb
        // fails no s can pass (NaN included), the step bisects, and the
        // interpolant, which nothing else reads, is not evaluated.
        const double lo = (3 * xa + 4) / 4;
        bool bisect = true;
        if (lo < xb) {
            if (fa != fc && fb != fc) {
                s = (xa * fb * fc) / ((fa - fb) - (fa - fc)) + (xb * fa * fc) / ((fb - fa) - (fb - fc)) +
                    (xc * fa * fb) / ((fc - fa) - (fc - fb));
            } else {
                s = xb - fb * (xb - xa) / (fb - fa);
            }
            bool c1 = !(lo < s && s < xb);
            bool c2 = mflag && fabs(s - xb) >= fabs(xb - xc) / 2;
            bool c3 = !mflag && fabs(s - xb) >= fabs(xc - d) / 2;
            bisect = c1 || c2 || c3;
        }
        if (bisect) { s = (xa + xb) / 2; mflag = true; }
        else mflag = false;
        // fs only enters through the sign of fa * fs (fa and fb are never
        // refreshed, quirk Q13)
        double fs = c2p_f_signed(gm, s, Ssq, Bsq, SB, D, tau);
        d = xc; xc = xb; fc = fb;
        if (fa * fs < 0.0) xb = s;
        else xa = s;
        if (fabs(fa) < fabs(fb)) { dswap(xa, xb); dswap(fa, fb); }
    }
    return false;
}

// Conserved-to-primitive inversion, conserved.cpp:135-230, in three parts
// so that kernels can run the Newton stage for every cell and the rare,
// long Brent fallback over a compacted list of cells.
// quirk Q6 with the uninitialised read at conserved.cpp:148-149: only the
// j=2 term of gamma^{ij}S_j survives, and for i=0,1 S[2]/B[2] are read before
// assignment; the deterministic build of the reference reads 0 there.
// quirks Q9 (bracket expansion updates f_lo), Q10, Q11, Q12 (500 Newton
// iterations without convergence count as success).
struct C2P {
    double D, tau, Ssq, Bsq, SB, B[3], Bu[3], Su[3];
};
BH_HD void ctop_setup(const GPt& m, const Cons& cv, C2P& s) {
    const double gami2[3] = {m.gi13, 0.0, m.gi33};
    double sg = m.sqrtdetgam;
    s.D = cv.D / sg;
    s.tau = cv.tau / sg;
    double S[3];
    s.Bsq = 0.0; s.Ssq = 0.0; s.SB = 0.0;
    for (int i = 0; i < 3; i++) {
        s.B[i] = div_z(cv.B[i], sg);
        S[i] = cv.S[i] / sg;
        double B2 = (i == 2) ? s.B[2] : 0.0;
        double S2 = (i == 2) ? S[2] : 0.0;
        s.Bu[i] = gami2[i] * B2;
        s.Su[i] = gami2[i] * S2;
        s.Ssq += S[i] * s.Su[i];
        s.Bsq += s.B[i] * s.Bu[i];
        s.SB += S[i] * s.Bu[i];
    }
}
// Residuals at the ends of the Brent bracket [D, xi_h], when the Newton
// stage has already evaluated them (exact = the bracket was not widened).
struct C2PEnds {
    double fD, fh;
    bool exact;
};
// Bracket [xi_l, xi_h] and Newton iteration.  Returns false when the
// reference falls back to Brent (then xi_h is the bracket to use).
BH_HD bool ctop_newton(double gm, const C2P& s, double& xi, double& xi_h, C2PEnds* ends = nullptr) {
    const int maxiter = 500;
    const double tol = 1e-10;
    const double D = s.D, tau = s.tau, Ssq = s.Ssq, Bsq = s.Bsq, SB = s.SB;
    double xi_l = D;
    xi_h = tau + D + Bsq / 2.0 + 1e3;
    double fl = c2p_f(gm, xi_l, Ssq, Bsq, SB, D, tau);
    double fh = c2p_f(gm, xi_h, Ssq, Bsq, SB, D, tau);
    if (ends) { ends->fD = fl; ends->fh = fh; ends->exact = true; }
    while (fl * fh > 0.0 && xi_h < 1e8 * D) {
        xi_h *= 2.0;
        fl = c2p_f_signed(gm, xi_h, Ssq, Bsq, SB, D, tau);  // fl only enters through fl * fh > 0
        if (ends) ends->exact = false;
    }
    xi = (xi_l + xi_h) / 2.0;
    // The iterate is a pure function of the previous one, so once it is NaN,
    // or repeats with period 1 or 2, every remaining pass is determined: the
    // loop stops early with the value the full 500 passes would end on (NaN
    // stays NaN; a 2-cycle ends on the parity of the remaining count).
    double xprev = xi;
    if (xi == xi) {
        for (int it = 0; it < maxiter; it++) {
            double dF;
            double F = c2p_fdf(gm, xi, Ssq, Bsq, SB, D, tau, dF);
            if (fabs(F) < tol) break;
            if (fabs(dF) < 1e-30) return false;
            double xin = xi - F / dF;
            if (xin < xi_l || xin > xi_h) return false;
            if (xin != xin || xin == xi) { xi = xin; break; }
            if (it > 0 && xin == xprev) {
                if (((maxiter - 1 - it) & 1) == 0) xi = xin;  // else the iterate stays at xi
                break;
            }
            xprev = xi;
            xi = xin;
        }
    }
    return true;
}
BH_HD bool ctop_brent(double gm, const C2P& s, double xi_h, double& xi) {
    return c2p_brent(gm, s.D, xi_h, s.Ssq, s.Bsq, s.SB, s.D, s.tau, xi, 1e-10);
}
// ctop_brent where its outcome is known without the loop: the end-point
// test, then the atmosphere form (c2p_brent_atmosphere).  1 = root xi,
// 0 = Brent fails, -1 = the loop must run.
// The residuals at D and xi_h are taken from the Newton stage where it
// computed them with these same arguments.
BH_HD int ctop_brent_short(double gm, const C2P& s, double xi_h, double& xi, const C2PEnds* ends = nullptr) {
    if (!(s.Ssq == 0.0 && s.Bsq == 0.0 && s.SB == 0.0)) return -1;
    double xa = s.D, xb = xi_h;
    const bool known = ends && ends->exact;
    double fa = known ? ends->fD : c2p_f(gm, xa, s.Ssq, s.Bsq, s.SB, s.D, s.tau);
    double fb = known ? ends->fh : c2p_f(gm, xb, s.Ssq, s.Bsq, s.SB, s.D, s.tau);
    if (fa * fb > 0.0) return 0;
    if (fabs(fa) < fabs(fb)) { dswap(xa, xb); dswap(fa, fb); }
    if (fa > 0.0 && s.D > 0.0 && xb < xa && isfinite(xa) && isfinite(s.tau))
        return c2p_brent_atmosphere(gm, xa, xb, s.D, s.tau, 1e-10, xi);
    return -1;
}
// Primitive recovery from the root xi (conserved.cpp:196-229)
BH_HD bool ctop_finish(double gm, const C2P& s, double xi, Prim& pv, int& nf) {
    const double Bsq = s.Bsq, SB = s.SB, Ssq = s.Ssq;
    double xpQ = xi + Bsq;
    double sdb2 = SB * SB;
    double vsq = (Ssq * xpQ * xpQ + sdb2 * (2.0 * xi + Bsq)) / (xpQ * xpQ * xi * xi);
    vsq = smin(vsq, 1.0 - 1e-10);
    double ltz2 = 1.0 / (1.0 - vsq);
    double ltz = sqrt(ltz2);
    pv.rho = s.D / ltz;
    double h = xi / (pv.rho * ltz2);
    pv.eps = (h - 1.0) / (gm);
    double p = eos_press(gm, pv.rho, pv.eps, nf);
    eos_enth(gm, pv.rho, pv.eps, nf);
    eos_temp(pv.rho, p, nf);
    for (int i = 0; i < 3; i++) {
        pv.v[i] = (s.Su[i] + div_z(SB, xi) * s.Bu[i]) / xpQ;
        pv.B[i] = s.B[i];
    }
    if (pv.rho <= 0.0 || pv.eps <= 0.0 || p <= 0.0) return false;
    return true;
}
BH_HD bool ctop(const GPt& m, double gm, const Cons& cv, Prim& pv, int& nf) {
    C2P s;
    ctop_setup(m, cv, s);
    double xi, xi_h;
    if (!ctop_newton(gm, s, xi, xi_h)) {
        if (!ctop_brent(gm, s, xi_h, xi)) return false;
    }
    return ctop_finish(gm, s, xi, pv, nf);
}

// ------------------------------------------------------- reconstruction
// Slope limiters, recon.hpp:16-34, and the two-face scalar reconstruction,
// reconstruction.cpp:28-38.
BH_HD double minmod(double a, double b) {
    if (a * b <= 0.0) return 0.0;
    return (fabs(a) < fabs(b)) ? a : b;
}
BH_HD double vanleer(double a, double b) {
    if (a * b <= 0.0) return 0.0;
    return 2.0 * a * b / (a + b);
}
BH_HD double mclim(double a, double b) {
    if (a * b <= 0.0) return 0.0;
    double c = (a + b) / 2;
    double s = (a > 0.0) ? 1.0 : -1.0;
    return s * minmod(fabs(2 * a), minmod(fabs(2 * b), fabs(c)));
}
enum { LIM_MINMOD = 0, LIM_VANLEER = 1, LIM_MC = 2 };
BH_HD double limit(double dl, double dr, int lim) {
    if (lim == LIM_VANLEER) return vanleer(dl, dr);
    if (lim == LIM_MC) return mclim(dl, dr);
    return minmod(dl, dr);
}
BH_HD void recon_scal(double Wm, double Wc, double Wp, double Wpp, int lim, double& fL, double& fR) {
    double cslop = limit(Wc - Wm, Wp - Wc, lim);
    fL = Wc + cslop / 2.0;
    double pslop = limit(Wp - Wc, Wpp - Wp, lim);
    fR = Wp - pslop / 2.0;
}
// Causality rescale of a reconstructed velocity with the cell metric,
// reconstruction.cpp:71-83 (quirk Q2).
BH_HD void causal_v(double v[3], const double g11, const double g22, const double g33) {
    const double gam[3][3] = {{g11, 0.0, CELL_G13}, {0.0, g22, 0.0}, {CELL_G13, 0.0, g33}};
    double v2 = 0.0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) v2 += gam[a][b] * v[a] * v[b];
    if (v2 >= 1.0) {
        double s = (1 - 1e-10) / sqrt(v2);
        for (int a = 0; a < 3; a++) v[a] *= s;
    }
}

// ------------------------------------------------------------------ HLLD
struct Side {
    double rho, eps, p, h, vn, vt1, vt2, Bn, Bt1, Bt2, ltz2, b2, pt, cf, D, tau, Sn, St1, St2, Bc, E;
};
struct CS { double D, Sn, St1, St2, tau, Bn, Bt1, Bt2; };

// side_ + magcomp + fastspeed, hlld.cpp:13-72
BH_HD Side hlld_side(const Prim& W, const Cons& U, const GPt& m, double gm, double sqrtg, int dim, int& nf) {
    const double gam[3][3] = {{m.g11, 0.0, m.g13}, {0.0, m.g22, 0.0}, {m.g13, 0.0, m.g33}};
    Side s;
    int t1 = (dim + 1) % 3, t2 = (dim + 2) % 3;
    s.rho = W.rho; s.eps = W.eps;
    s.p = eos_press(gm, W.rho, W.eps, nf);
    s.h = eos_enth(gm, W.rho, W.eps, nf);
    double v2 = 0.0;
    double vlow[3] = {0.0, 0.0, 0.0}, Blow[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) {
            vlow[i] += gam[i][j] * W.v[j];
            Blow[i] += gam[i][j] * W.B[j];
        }
        v2 += vlow[i] * W.v[i];
    }
    v2 = smin(v2, 1 - 1e-10);
    s.ltz2 = 1.0 / (1.0 - v2);
    double Bv = 0.0;
    for (int i = 0; i < 3; i++) Bv += Blow[i] * W.v[i];
    double Bsq = 0.0;
    for (int i = 0; i < 3; i++) Bsq += Blow[i] * W.B[i];
    s.b2 = Bv * Bv + Bsq / s.ltz2;
    s.vn = W.v[dim]; s.vt1 = W.v[t1]; s.vt2 = W.v[t2];
    s.Bn = W.B[dim]; s.Bt1 = W.B[t1]; s.Bt2 = W.B[t2];
    s.pt = s.p + s.b2 / 2;
    double cs2 = eos_cs2(gm, W.rho, W.eps, nf);
    double va2 = s.b2 / (s.rho * s.h + s.b2);
    double cf2 = cs2 + va2 * (1 - cs2);
    cf2 = smin(cf2, 1 - 1e-10);
    s.cf = sqrt(cf2);
    s.D = U.D / sqrtg; s.tau = U.tau / sqrtg;
    s.Sn = U.S[dim] / sqrtg; s.St1 = U.S[t1] / sqrtg; s.St2 = U.S[t2] / sqrtg;
    s.Bc = U.B[dim] / sqrtg;
    s.E = s.tau + s.D;
    return s;
}
// physflux with the ADM correction, hlld.cpp:75-120
BH_HD CS hlld_physflux(const Side& s, double alpha, double btn) {
    CS F;
    double vn = s.vn;
    F.D = s.D * vn;
    F.tau = (s.E + s.pt) * vn - s.Bn * (s.Bn * vn + s.Bt1 * s.vt1 + s.Bt2 * s.vt2) / s.ltz2;
    F.Sn = s.Sn * vn - s.Bn * s.Bn / s.ltz2 + s.pt;
    F.St1 = s.St1 * vn - s.Bt1 * s.Bn / s.ltz2;
    F.St2 = s.St2 * vn - s.Bt2 * s.Bn / s.ltz2;
    F.Bn = 0.0;
    F.Bt1 = s.Bt1 * vn - s.Bn * s.vt1;
    F.Bt2 = s.Bt2 * vn - s.Bn * s.vt2;
    F.D = alpha * F.D - btn * s.D;
    F.tau = alpha * F.tau - btn * s.tau;
    F.Sn = alpha * F.Sn - btn * s.Sn;
    F.St1 = alpha * F.St1 - btn * s.St1;
    F.St2 = alpha * F.St2 - btn * s.St2;
    F.Bn = alpha * F.Bn - btn * s.Bn;
    F.Bt1 = alpha * F.Bt1 - btn * s.Bt1;
    F.Bt2 = alpha * F.Bt2 - btn * s.Bt2;
    return F;
}
// HLL fallback (Toro ch. 10.3), hlld.cpp:122-137
BH_HD CS hll_flux(const Side& L, const Side& R, const CS& FL, const CS& FR, double SL, double SR) {
    double dS = SR - SL;
    CS F;
    F.D = (SR * FL.D - SL * FR.D + SL * SR * (R.D - L.D)) / dS;
    F.tau = (SR * FL.tau - SL * FR.tau + SL * SR * (R.tau - L.tau)) / dS;
    F.Sn = (SR * FL.Sn - SL * FR.Sn + SL * SR * (R.Sn - L.Sn)) / dS;
    F.St1 = (SR * FL.St1 - SL * FR.St1 + SL * SR * (R.St1 - L.St1)) / dS;
    F.St2 = (SR * FL.St2 - SL * FR.St2 + SL * SR * (R.St2 - L.St2)) / dS;
    F.Bn = (SR * FL.Bn - SL * FR.Bn + SL * SR * (R.Bn - L.Bn)) / dS;
    F.Bt1 = (SR * FL.Bt1 - SL * FR.Bt1 + SL * SR * (R.Bt1 - L.Bt1)) / dS;
    F.Bt2 = (SR * FL.Bt2 - SL * FR.Bt2 + SL * SR * (R.Bt2 - L.Bt2)) / dS;
    return F;
}
// Star state across a fast wave, hlld.cpp:140-158 (quirk Q16)
BH_HD CS hlld_star(const Side& s, double SK, double SM, double pt_x) {
    CS Ux;
    Ux.D = s.D * (SK - s.vn) / (SK - SM);
    Ux.tau = (s.E * (SK - s.vn) - s.pt * s.vn + pt_x * SM) / (SK - SM) - Ux.D;
    Ux.Sn = (s.Sn * (SK - s.vn) - s.pt + pt_x) / (SK - SM);
    Ux.St1 = s.St1 * (SK - s.vn) / (SK - SM);
    Ux.St2 = s.St2 * (SK - s.vn) / (SK - SM);
    Ux.Bn = s.Bc;
    double den = s.D * (SK - s.vn) * (SK - SM) - s.Bn * s.Bn;
    if (fabs(den) > 1e-14) {
        double k = s.Bn * (SM - s.vn) / den;
        Ux.Bt1 = s.Bt1 - s.Bn * (s.vt1 * (SK - s.vn) + s.Bn * s.Bt1 / s.D) * k;
        Ux.Bt2 = s.Bt2 - s.Bn * (s.vt2 * (SK - s.vn) + s.Bn * s.Bt2 / s.D) * k;
    } else {
        Ux.Bt1 = s.Bt1; Ux.Bt2 = s.Bt2;
    }
    return Ux;
}
// Double-star state across the Alfven waves for one side, hlld.cpp:160-200.
// The degenerate branch at hlld.cpp:171-173 is overwritten by what follows
// (quirk Q15), so it has no effect and is not repeated here.
BH_HD CS hlld_dstar(const CS& UxL, const CS& UxR, double BnL, double SM, bool left) {
    double sq_DL = sqrt(fabs(UxL.D));
    double sq_DR = sqrt(fabs(UxR.D));
    double den = sq_DL + sq_DR;
    double sgnBn = (BnL >= 0.0) ? 1.0 : -1.0;
    double vt1xx = (sq_DL * UxL.St1 / UxL.D + sq_DR * UxR.St1 / UxR.D + sgnBn * (UxR.Bt1 - UxL.Bt1)) / den;
    double vt2xx = (sq_DL * UxL.St2 / UxL.D + sq_DR * UxR.St2 / UxR.D + sgnBn * (UxR.Bt2 - UxL.Bt2)) / den;
    double Bt1xx = (sq_DL * UxL.Bt1 / UxL.D + sq_DR * UxR.Bt1 / UxR.D +
                    sgnBn * sq_DL * sq_DR * (UxR.St1 / UxR.D - UxL.St1 / UxL.D)) / den;
    double Bt2xx = (sq_DL * UxL.Bt2 / UxL.D + sq_DR * UxR.Bt2 / UxR.D +
                    sgnBn * sq_DL * sq_DR * (UxR.St2 / UxR.D - UxL.St2 / UxL.D)) / den;
    const CS& Ux = left ? UxL : UxR;
    double sq = left ? sq_DL : sq_DR;
    CS Uxx;
    Uxx.D = Ux.D;
    Uxx.Bn = Ux.Bn;
    Uxx.St1 = Ux.D * vt1xx;
    Uxx.St2 = Ux.D * vt2xx;
    Uxx.Bt1 = Bt1xx;
    Uxx.Bt2 = Bt2xx;
    Uxx.Sn = Ux.D * SM;
    const double vt1x = Ux.St1 / Ux.D, vt2x = Ux.St2 / Ux.D;
    double dE = sq * sgnBn * (vt1xx * Bt1xx + vt2xx * Bt2xx - vt1x * Ux.Bt1 - vt2x * Ux.Bt2);
    Uxx.tau = left ? Ux.tau - dE : Ux.tau + dE;
    return Uxx;
}
// Intermediate-state flux, hlld.cpp:202-215
BH_HD CS hlld_inter(const CS& FK, double SK, const CS& Ux, const Side& s) {
    CS F;
    F.D = FK.D + SK * (Ux.D - s.D);
    F.tau = FK.tau + SK * (Ux.tau - (s.E - s.D));
    F.Sn = FK.Sn + SK * (Ux.Sn - s.Sn);
    F.St1 = FK.St1 + SK * (Ux.St1 - s.St1);
    F.St2 = FK.St2 + SK * (Ux.St2 - s.St2);
    F.Bn = FK.Bn + SK * (Ux.Bn - s.Bc);
    F.Bt1 = FK.Bt1 + SK * (Ux.Bt1 - s.Bt1);
    F.Bt2 = FK.Bt2 + SK * (Ux.Bt2 - s.Bt2);
    return F;
}
// Re-densitise by sqrt(-g) and restore component order, hlld.cpp:217-227
BH_HD Cons hlld_dens(const CS& F, double sg, int dim) {
    int t1 = (dim + 1) % 3, t2 = (dim + 2) % 3;
    Cons c;
    c.D = sg * F.D; c.tau = sg * F.tau;
    c.B[dim] = sg * F.Bn; c.B[t1] = sg * F.Bt1; c.B[t2] = sg * F.Bt2;
    c.S[dim] = sg * F.Sn; c.S[t1] = sg * F.St1; c.S[t2] = sg * F.St2;
    return c;
}
BH_HD bool cons_finite(const Cons& c) {
    return isfinite(c.D) && isfinite(c.tau) && isfinite(c.S[0]) && isfinite(c.S[1]) && isfinite(c.S[2]) &&
           isfinite(c.B[0]) && isfinite(c.B[1]) && isfinite(c.B[2]);
}

// HLLD flux at a face, hlld.cpp:230-312 (quirks Q14, Q17).  beta^dim is the
// face shift component (beta^1 = beta^2 = +0.0 in this metric).
BH_HD Cons hlld(const Prim& WL, const Prim& WR, const Cons& UL, const Cons& UR, const GPt& m, double gm, int dim,
                double& ctspeed, int& nf) {
    double sg = m.sqrtdetg;
    double btn = (dim == 0) ? m.beta0 : 0.0;
    Side L = hlld_side(WL, UL, m, gm, sg, dim, nf);
    Side R = hlld_side(WR, UR, m, gm, sg, dim, nf);
    double SL = smin(m.alpha * (L.vn - L.cf) - btn, m.alpha * (R.vn - R.cf) - btn);
    double SR = smax(m.alpha * (L.vn + L.cf) - btn, m.alpha * (R.vn + R.cf) - btn);
    CS FL = hlld_physflux(L, m.alpha, btn), FR = hlld_physflux(R, m.alpha, btn);
    ctspeed = smax(fabs(SL), fabs(SR));
    if (SL >= 0.0) return hlld_dens(FL, sg, dim);
    if (SR <= 0.0) return hlld_dens(FR, sg, dim);
    double pt_x = (SR * L.pt - SL * R.pt + SL * SR * (R.Sn - L.Sn)) / (SR - SL);
    double SM_u = SR * L.E - SL * R.E - FR.tau - FR.D + FL.tau + FL.D;
    double SM_d = SR * L.D - SL * R.D - FR.D + FL.D;
    if (fabs(SM_d) < 1e-12) return hlld_dens(hll_flux(L, R, FL, FR, SL, SR), sg, dim);
    double SM = SM_u / SM_d;
    CS UxL = hlld_star(L, SL, SM, pt_x);
    CS UxR = hlld_star(R, SR, SM, pt_x);
    double SLx = SM - fabs(L.Bn) / sqrt(fabs(UxL.D));
    double SRx = SM + fabs(R.Bn) / sqrt(fabs(UxR.D));
    if (UxL.D <= 0.0 || UxR.D <= 0.0) return hlld_dens(hll_flux(L, R, FL, FR, SL, SR), sg, dim);
    Cons F;
    if (SM >= 0.0) {
        CS FLx = hlld_inter(FL, SL, UxL, L);
        if (SLx >= 0.0) {
            F = hlld_dens(FLx, sg, dim);
        } else {
            CS UxxL = hlld_dstar(UxL, UxR, L.Bn, SM, true);
            F = hlld_dens(hlld_inter(FLx, SLx, UxxL, L), sg, dim);
        }
    } else {
        CS FRx = hlld_inter(FR, SR, UxR, R);
        if (SRx >= 0.0) {
            F = hlld_dens(FRx, sg, dim);
        } else {
            CS UxxR = hlld_dstar(UxL, UxR, L.Bn, SM, false);
            F = hlld_dens(hlld_inter(FRx, SRx, UxxR, R), sg, dim);
        }
    }
    if (!cons_finite(F)) F = hlld_dens(hll_flux(L, R, FL, FR, SL, SR), sg, dim);
    if (ctspeed > 1.0) {
        F = hlld_dens(hll_flux(L, R, FL, FR, SL, SR), sg, dim);
        ctspeed = smin(ctspeed, 1.0);
    }
    return F;
}

// ------------------------------------------------------ floor state
// amrtree::pvfs, grid.cpp:133-146: rho = 1e-7 r^-3/2, eps = 1e-5 r^-3/2,
// v = 0, B = 0 (the r-dependent factors are precomputed per cell).
BH_HD Prim floor_prim(double rho_fl, double eps_fl) {
    Prim W;
    W.rho = rho_fl; W.eps = eps_fl;
    for (int d = 0; d < 3; d++) { W.v[d] = 0.0; W.B[d] = 0.0; }
    return W;
}

}  // namespace bh
