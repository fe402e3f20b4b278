// Test infrastructure: the oracle.  CPU restatement of the reference's GRMHD
// fluid step (Fjoerde/black-hole-simulator, src/fluid), written from the
// reference's semantics for use as the checker by tests/, the smoke check
// and the CPU-baseline leg of bench.py only.  The product never links it.
//
// Every function keeps the reference's floating-point operation order (same
// association, same libm calls) so that, built by the same compiler with the
// same flags as the compiled reference (oracle/Makefile), it rounds
// identically.  Reference quirks are reproduced deliberately; each is marked
// "quirk" with its reference file:line.
#include "bhoracle.h"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <utility>
#include <vector>

namespace {

constexpr int BLK = 16, GH = 2, CPN = 20;
constexpr int NCELL = CPN * CPN * CPN;          // 8000
constexpr int NFACE = (CPN + 1) * CPN * CPN;    // 8400
constexpr int NEDGE = CPN * (CPN + 1) * (CPN + 1);  // 8820

// grid.cpp:51-89 index maps (k fastest, offset by the ghost width)
inline int lindx(int i, int j, int k) { return (i + GH) * CPN * CPN + (j + GH) * CPN + (k + GH); }
inline int ixf(int i, int j, int k) { return (i + GH) * CPN * CPN + (j + GH) * CPN + (k + GH); }          // xfc/Bfx
inline int iyf(int i, int j, int k) { return (i + GH) * (CPN + 1) * CPN + (j + GH) * CPN + (k + GH); }    // yfc/Bfy/EMFz
inline int izf(int i, int j, int k) { return (i + GH) * CPN * (CPN + 1) + (j + GH) * (CPN + 1) + (k + GH); }  // zfc/Bfz/EMFy
inline int iex(int i, int j, int k) { return (i + GH) * (CPN + 1) * (CPN + 1) + (j + GH) * (CPN + 1) + (k + GH); }

// ------------------------------------------------------------------ libm
// libquadmath (GCC runtime): binary128 log/exp, rounded once to binary64.
extern "C" __float128 logq(__float128);
extern "C" __float128 expq(__float128);
double cr_log(double x) { return (double)logq((__float128)x); }
double cr_exp(double x) { return (double)expq((__float128)x); }
struct Libm {
    double (*log)(double);
    double (*exp)(double);
};
Libm libm_for(int mode) {
    if (mode == 1) return {cr_log, cr_exp};
    return {static_cast<double (*)(double)>(::log), static_cast<double (*)(double)>(::exp)};
}

// std::min / std::max semantics (NaN in the second argument never wins)
inline double smin(double a, double b) { return (b < a) ? b : a; }
inline double smax(double a, double b) { return (a < b) ? b : a; }

// ------------------------------------------------------------- geometry
// metriccomp, metric.hpp:11-21
struct MC {
    double g[4][4], gi[4][4], G[4][4][4], sqrtdetg, gam[3][3], gami[3][3], sqrtdetgam, alpha, beta[3];
};
struct KS { double M, a, Q; };

// metric::Sigma / metric::H, spacetime.cpp:12-17
inline double ks_sigma(const KS& k, double r, double th) { return r * r + k.a * k.a * std::pow(std::cos(th), 2); }
inline double ks_H(const KS& k, double r, double th) {
    return (2 * k.M * r - k.Q * k.Q) * std::pow(ks_sigma(k, r, th), -1);
}

// metric::comp and its parts, spacetime.cpp:20-134
MC metric_comp(const KS& k, double r, double th) {
    MC m;
    const double M = k.M, a = k.a, Q = k.Q;
    const double H = ks_H(k, r, th);
    const double Sg = ks_sigma(k, r, th);
    // gcova, spacetime.cpp:20-34
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) m.g[i][j] = 0.0;
    m.g[0][0] = -1 + H;
    m.g[0][1] = m.g[1][0] = H;
    m.g[1][1] = 1 + H;
    m.g[2][2] = Sg;
    m.g[0][3] = m.g[1][3] = m.g[3][0] = m.g[3][1] = -H * a * std::pow(std::sin(th), 2);
    m.g[3][3] = (r * r + a * a + H * a * a * std::pow(std::sin(th), 2)) * std::pow(std::sin(th), 2);
    // gcntr, spacetime.cpp:35-49
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) m.gi[i][j] = 0.0;
    m.gi[0][0] = -1 - H;
    m.gi[0][1] = m.gi[1][0] = H;
    m.gi[1][1] = 1 - H;
    m.gi[2][2] = std::pow(Sg, -1);
    m.gi[1][3] = m.gi[3][1] = a * std::pow(Sg, -1);
    m.gi[3][3] = std::pow(Sg * std::sin(th) * std::sin(th), -1);
    // gdet, spacetime.cpp:52-54 (quirk Q3: sin^2 inside)
    m.sqrtdetg = (r * r + a * a * std::pow(std::sin(th), 2)) * std::sin(th);
    // christ, spacetime.cpp:57-90 (quirk Q5: the derivative table is the reference's own)
    double d[4][4][4];
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) for (int l = 0; l < 4; l++) d[i][j][l] = 0;
    const double c = std::cos(th), s = std::sin(th);
    const double S1 = r * r + a * a * std::pow(c, 2);
    d[1][0][0] = d[1][0][1] = d[1][1][0] = d[1][1][1] =
        -2 * (M * r * r - Q * Q * r - M * a * a * std::pow(c, 2)) * std::pow(S1, -2);
    d[2][0][0] = d[2][0][1] = d[2][1][0] = d[2][1][1] = 2 * a * a * (2 * M * r - Q * Q) * c * s * std::pow(S1, -1);
    d[1][2][2] = 2 * r;
    d[2][2][2] = -a * a * c * s;
    d[1][0][3] = d[1][3][0] = d[1][1][3] = d[1][3][1] =
        2 * a * a * std::pow(s, 2) * (M * r * r - Q * Q * r - M * a * a * std::pow(c, 2)) * std::pow(S1, -2);
    d[2][0][3] = d[2][3][0] = d[2][1][3] = d[2][3][1] =
        2 * a * c * s * (2 * M * r - Q * Q) * (r * r + a * a) * std::pow(S1, -2);
    d[1][3][3] = std::pow(s, 2) * (((2 * M * a * a * std::pow(s, 2)) * std::pow(S1, -1)) +
                                   ((2 * r * a * a * std::pow(s, 2) * (2 * M * r - Q * Q)) * std::pow(S1, -2)) + 2 * r);
    d[2][3][3] = std::pow(s, 2) * (((2 * std::pow(a, 4) * (2 * M * r - Q * Q) * c * std::pow(s, 3)) * std::pow(S1, -2)) +
                                   ((2 * a * a * (2 * M * r - Q * Q) * c * s) * std::pow(S1, -1))) +
                 c * s * (((a * a * std::pow(s, 2) * (2 * M * r - Q * Q)) * std::pow(S1, -1)) + r * r + a * a);
    for (int l = 0; l < 4; l++)
        for (int mm = 0; mm < 4; mm++)
            for (int n = 0; n < 4; n++)
                m.G[l][mm][n] = 0.5 * (m.gi[l][0] * (d[mm][0][n] + d[n][0][mm] - d[0][mm][n]) +
                                       m.gi[l][1] * (d[mm][1][n] + d[n][1][mm] - d[1][mm][n]) +
                                       m.gi[l][2] * (d[mm][2][n] + d[n][2][mm] - d[2][mm][n]) +
                                       m.gi[l][3] * (d[mm][3][n] + d[n][3][mm] - d[3][mm][n]));
    // lapse / shift, spacetime.cpp:93-100
    m.alpha = std::pow(1 + H, -1.0 / 2.0);
    for (int i = 0; i < 3; i++) m.beta[i] = -m.gi[0][i + 1] / m.gi[0][0];
    // spatial parts, spacetime.cpp:103-119
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) { m.gam[i][j] = m.g[i + 1][j + 1]; m.gami[i][j] = m.gi[i + 1][j + 1]; }
    m.sqrtdetgam = m.alpha * m.sqrtdetg;
    return m;
}

// ------------------------------------------------------------------ EOS
// state, thermo.cpp:10-35.  rho <= 0 throws in the reference; here it is
// counted and the arithmetic continues.
struct EOS {
    double gamma;
    std::atomic<long long>* faults;
    void chk(double rho) const { if (rho <= 0.0) faults->fetch_add(1, std::memory_order_relaxed); }
    double press(double rho, double eps) const { chk(rho); return (gamma - 1.0) * rho * eps; }
    double enth(double rho, double eps) const { chk(rho); return 1.0 + gamma * eps; }
    double temp(double rho, double p) const { chk(rho); return p / rho; }
    double cs2(double rho, double eps) const {
        chk(rho);
        double p = press(rho, eps);
        double h = enth(rho, eps);
        return (gamma * p) / (rho * h);
    }
};

// prim / cons, prim.hpp:14-27 and cons.hpp:14-59
struct Prim { double rho, eps, v[3], B[3], p, T, h, lor, b2; };
struct Cons { double D, S[3], tau, B[3]; };
inline Cons czero() { Cons c; c.D = 0.0; c.tau = 0.0; for (int i = 0; i < 3; i++) { c.S[i] = 0.0; c.B[i] = 0.0; } return c; }

// conserved::ptoc, conserved.cpp:17-73 (quirk Q7: b2 divided by W)
Cons ptoc(const KS& ks, const EOS& st, Prim& pv, double r, double th) {
    MC mc = metric_comp(ks, r, th);
    double sg = mc.sqrtdetgam;
    double v2 = 0.0;
    double vlow[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) vlow[i] += mc.gam[i][j] * pv.v[j];
        v2 += vlow[i] * pv.v[i];
    }
    v2 = smin(v2, 1.0 - 1e-10);
    pv.lor = 1.0 / std::sqrt(1.0 - v2);
    pv.p = st.press(pv.rho, pv.eps);
    pv.h = st.enth(pv.rho, pv.eps);
    double Blow[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Blow[i] += mc.gam[i][j] * pv.B[j];
    double Bv = 0.0;
    for (int i = 0; i < 3; i++) Bv += Blow[i] * pv.v[i];
    double Bsq = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Bsq += Blow[i] * mc.gam[i][j] * pv.B[j];
    pv.b2 = (Bv * Bv + Bsq) / (pv.lor);
    double bt = pv.lor * Bv / mc.alpha;
    double pstar = pv.p + pv.b2 / 2;
    double w = pv.rho * pv.h + pv.b2;
    Cons cv = czero();
    cv.D = sg * pv.rho * pv.lor;
    cv.tau = sg * (w * pv.lor * pv.lor - pstar - bt * bt) - cv.D;
    for (int i = 0; i < 3; i++) {
        double bi = (Blow[i] + mc.alpha * bt * pv.lor * vlow[i]) / pv.lor;
        cv.S[i] = sg * (w * pv.lor * pv.lor * vlow[i] - mc.alpha * bt * bi);
        cv.B[i] = sg * pv.B[i];
    }
    return cv;
}

// conserved::f_df, conserved.cpp:80-123 (quirk Q8: derivative denominator)
double f_df(const EOS& st, double xi, double Ssq, double Bsq, double SB, double Dn, double taun, double& df) {
    double Yi = SB * SB;
    double K = Bsq;
    double X = xi * xi;
    double xipK = xi + K;
    double vsq = (Ssq * X + Yi * (2.0 * xi + K)) / (X + xipK * xipK);
    if (vsq >= 1.0) { df = 0.0; return 1e30; }
    double ltz2 = 1.0 / (1.0 - vsq);
    double ltz = std::sqrt(ltz2);
    double rho = Dn / ltz;
    double G = st.gamma - 1.0;
    double p = (G / st.gamma) * (xi / ltz2 - Dn / ltz);
    double pstar = p + K / (2 * ltz2) + Yi / (2 * X);
    if (p < 0.0 || rho < 0.0) { df = 0.0; return 1e32; }
    double f = xi - taun - Dn - pstar;
    double N = Ssq * X + Yi * (2.0 * xi + K);
    double Mq = X * xipK * xipK;
    double dN = 2.0 * Ssq * xi + 2.0 * Yi;
    double dM = 2.0 * xi * xipK * xipK + 2.0 * X * X;
    double dvv = (dN * Mq - dM * N) / (Mq * Mq);
    double dltz2 = ltz2 * ltz2 * dvv;
    double dltz = dltz2 / (2.0 * ltz);
    double dp = (G / st.gamma) * ((1.0 / ltz2) - (xi * dltz2) / (ltz2 * ltz2) + Dn * dltz / ltz2);
    double dpstar = dp - (K * dltz2) / (2.0 * ltz2 * ltz2) - Yi / (xi * X);
    df = 1 - dpstar;
    return f;
}
inline double fres(const EOS& st, double xi, double Ssq, double Bsq, double SB, double Dn, double taun) {
    double dum;
    return f_df(st, xi, Ssq, Bsq, SB, Dn, taun, dum);
}

// conserved::cp_bd, conserved.cpp:232-285 (quirk Q13: bisection test; fa/fb never refreshed)
bool cp_bd(const EOS& st, double xa, double xb, double Ssq, double Bsq, double SB, double D, double tau, double& root,
           double tol) {
    double fa = fres(st, xa, Ssq, Bsq, SB, D, tau);
    double fb = fres(st, xb, Ssq, Bsq, SB, D, tau);
    if (fa * fb > 0.0) return false;
    if (std::abs(fa) < std::abs(fb)) { std::swap(xa, xb); std::swap(fa, fb); }
    double xc = xa, fc = fa;
    bool mflag = true;
    double s = 0.0, d = 0.0;
    for (int i = 0; i < 100; i++) {
        if (std::abs(xb - xa) < tol) { root = xb; return true; }
        if (fa != fc && fb != fc) {
            s = (xa * fb * fc) / ((fa - fb) - (fa - fc)) + (xb * fa * fc) / ((fb - fa) - (fb - fc)) +
                (xc * fa * fb) / ((fc - fa) - (fc - fb));
        } else {
            s = xb - fb * (xb - xa) / (fb - fa);
        }
        bool c1 = !((3 * xa + 4) / 4 < s && s < xb);
        bool c2 = mflag && std::abs(s - xb) >= std::abs(xb - xc) / 2;
        bool c3 = !mflag && std::abs(s - xb) >= std::abs(xc - d) / 2;
        if (c1 || c2 || c3) { s = (xa + xb) / 2; mflag = true; }
        else mflag = false;
        double fs = fres(st, s, Ssq, Bsq, SB, D, tau);
        d = xc; xc = xb; fc = fb;
        if (fa * fs < 0.0) xb = s;
        else xa = s;
        if (std::abs(fa) < std::abs(fb)) { std::swap(xa, xb); std::swap(fa, fb); }
    }
    return false;
}

// conserved::ctop, conserved.cpp:135-230.
// quirk Q6 + uninitialised read: inside the i-loop Bu[i]/Su[i] keep only the
// j=2 product, and for i=0,1 B[2]/S[2] are read before being assigned; the
// deterministic build of the reference zeroes them (oracle/Makefile).
// quirks Q9 (bracket expansion updates fl), Q10, Q11 (double lowering), Q12.
bool ctop(const KS& ks, const EOS& st, const Cons& cv, double r, double th, Prim& pv, int maxiter = 500,
          double tol = 1e-10) {
    MC mc = metric_comp(ks, r, th);
    double sg = mc.sqrtdetgam;
    double D = cv.D / sg;
    double tau = cv.tau / sg;
    double B[3] = {0.0, 0.0, 0.0}, S[3] = {0.0, 0.0, 0.0};
    double Bsq = 0.0, Ssq = 0.0, SB = 0.0;
    double Bu[3], Su[3];
    for (int i = 0; i < 3; i++) {
        B[i] = cv.B[i] / sg;
        S[i] = cv.S[i] / sg;
        Bu[i] = mc.gami[i][2] * B[2];
        Su[i] = mc.gami[i][2] * S[2];
        Ssq += S[i] * Su[i];
        Bsq += B[i] * Bu[i];
        SB += S[i] * Bu[i];
    }
    double xi;
    double xi_l = D;
    double xi_h = tau + D + Bsq / 2.0 + 1e3;
    double fl = fres(st, xi_l, Ssq, Bsq, SB, D, tau);
    double fh = fres(st, xi_h, Ssq, Bsq, SB, D, tau);
    while (fl * fh > 0.0 && xi_h < 1e8 * D) {
        xi_h *= 2.0;
        fl = fres(st, xi_h, Ssq, Bsq, SB, D, tau);
    }
    xi = (xi_l + xi_h) / 2.0;
    bool nr_ok = true;
    for (int it = 0; it < maxiter; it++) {
        double dF;
        double F = f_df(st, xi, Ssq, Bsq, SB, D, tau, dF);
        if (std::abs(F) < tol) break;
        if (std::abs(dF) < 1e-30) { nr_ok = false; break; }
        double xin = xi - F / dF;
        if (xin < xi_l || xin > xi_h) { nr_ok = false; break; }
        xi = xin;
    }
    if (!nr_ok) {
        double xib;
        if (!cp_bd(st, xi_l, xi_h, Ssq, Bsq, SB, D, tau, xib, tol)) return false;
        xi = xib;
    }
    double xpQ = xi + Bsq;
    double sdb2 = SB * SB;
    double vsq = (Ssq * xpQ * xpQ + sdb2 * (2.0 * xi + Bsq)) / (xpQ * xpQ * xi * xi);
    vsq = smin(vsq, 1.0 - 1e-10);
    double ltz2 = 1.0 / (1.0 - vsq);
    double ltz = std::sqrt(ltz2);
    pv.lor = ltz;
    pv.rho = D / ltz;
    double h = xi / (pv.rho * ltz2);
    pv.eps = (h - 1.0) / (st.gamma);
    pv.p = st.press(pv.rho, pv.eps);
    pv.h = st.enth(pv.rho, pv.eps);
    pv.T = st.temp(pv.rho, pv.p);
    for (int i = 0; i < 3; i++) {
        pv.v[i] = (Su[i] + (SB / xi) * Bu[i]) / xpQ;
        pv.B[i] = B[i];
    }
    double Blow[3] = {};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Blow[i] += mc.gam[i][j] * B[j];
    double Bv = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Bv += Blow[i] * mc.gam[i][j] * pv.v[j];
    pv.b2 = (Bv * Bv + Bsq) / ltz2;
    if (pv.rho <= 0.0 || pv.eps <= 0.0 || pv.p <= 0.0) return false;
    return true;
}

// ------------------------------------------------------- reconstruction
// recon.hpp:16-34 limiters
inline double minmod(double a, double b) {
    if (a * b <= 0.0) return 0.0;
    return (std::abs(a) < std::abs(b)) ? a : b;
}
inline double min3mod(double a, double b, double c) { return minmod(a, minmod(b, c)); }
inline double vanleer(double a, double b) {
    if (a * b <= 0.0) return 0.0;
    return 2.0 * a * b / (a + b);
}
inline double mclim(double a, double b) {
    if (a * b <= 0.0) return 0.0;
    double c = (a + b) / 2;
    double s = (a > 0.0) ? 1.0 : -1.0;
    return s * min3mod(std::abs(2 * a), std::abs(2 * b), std::abs(c));
}
inline double lim_(double dl, double dr, int lim) {
    switch (lim) {
        case 0: return minmod(dl, dr);
        case 1: return vanleer(dl, dr);
        case 2: return mclim(dl, dr);
        default: return minmod(dl, dr);
    }
}
// recon_scal, reconstruction.cpp:28-38
inline void recon_scal(double Wm, double Wc, double Wp, double Wpp, int lim, double& fL, double& fR) {
    double dLc = Wc - Wm, dRc = Wp - Wc;
    double cslop = lim_(dLc, dRc, lim);
    fL = Wc + cslop / 2.0;
    double dLp = Wp - Wc, dRp = Wpp - Wp;
    double pslop = lim_(dLp, dRp, lim);
    fR = Wp - pslop / 2.0;
}

// --------------------------------------------------------------- HLLD
struct Side {
    double rho, eps, p, h, vn, vt1, vt2, Bn, Bt1, Bt2, v2, ltz, ltz2, b2, pt, cf, D, tau, Sn, St1, St2, Bc, E;
};
struct CS { double D, Sn, St1, St2, tau, Bn, Bt1, Bt2; };

// fastspeed, hlld.cpp:13-18
inline double fastspeed(double rho, double h, double b2, double cs2) {
    double va2 = b2 / (rho * h + b2);
    double cf2 = cs2 + va2 * (1 - cs2);
    cf2 = smin(cf2, 1 - 1e-10);
    return std::sqrt(cf2);
}
// side_ + magcomp, hlld.cpp:20-72
Side side_(const Prim& W, const Cons& U, const MC& mc, const EOS& st, double sqrtg, int dim) {
    Side s;
    int t1 = (dim + 1) % 3, t2 = (dim + 2) % 3;
    s.rho = W.rho; s.eps = W.eps;
    s.p = st.press(W.rho, W.eps);
    s.h = st.enth(W.rho, W.eps);
    // magcomp
    s.v2 = 0.0;
    double vlow[3] = {}, Blow[3] = {};
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) {
            vlow[i] += mc.gam[i][j] * W.v[j];
            Blow[i] += mc.gam[i][j] * W.B[j];
        }
        s.v2 += vlow[i] * W.v[i];
    }
    s.v2 = smin(s.v2, 1 - 1e-10);
    s.ltz2 = 1.0 / (1.0 - s.v2);
    s.ltz = std::sqrt(s.ltz2);
    double Bv = 0.0;
    for (int i = 0; i < 3; i++) Bv += Blow[i] * W.v[i];
    double Bsq = 0.0;
    for (int i = 0; i < 3; i++) Bsq += Blow[i] * W.B[i];
    s.b2 = Bv * Bv + Bsq / s.ltz2;
    s.vn = W.v[dim]; s.vt1 = W.v[t1]; s.vt2 = W.v[t2];
    s.Bn = W.B[dim]; s.Bt1 = W.B[t1]; s.Bt2 = W.B[t2];
    s.pt = s.p + s.b2 / 2;
    double cs2 = st.cs2(W.rho, W.eps);
    s.cf = fastspeed(s.rho, s.h, s.b2, cs2);
    s.D = U.D / sqrtg; s.tau = U.tau / sqrtg;
    s.Sn = U.S[dim] / sqrtg; s.St1 = U.S[t1] / sqrtg; s.St2 = U.S[t2] / sqrtg;
    s.Bc = U.B[dim] / sqrtg;
    s.E = s.tau + s.D;
    return s;
}
// physflux, hlld.cpp:75-120 (ADM correction)
CS physflux(const Side& s, const MC& mc, int dim) {
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
    double a = mc.alpha, btn = mc.beta[dim];
    F.D = a * F.D - btn * s.D;
    F.tau = a * F.tau - btn * s.tau;
    F.Sn = a * F.Sn - btn * s.Sn;
    F.St1 = a * F.St1 - btn * s.St1;
    F.St2 = a * F.St2 - btn * s.St2;
    F.Bn = a * F.Bn - btn * s.Bn;
    F.Bt1 = a * F.Bt1 - btn * s.Bt1;
    F.Bt2 = a * F.Bt2 - btn * s.Bt2;
    return F;
}
// hll_flux, hlld.cpp:122-137 (Toro ch. 10.3)
CS hll_flux(const Side& L, const Side& R, const CS& FL, const CS& FR, double SL, double SR) {
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
// star_stt, hlld.cpp:140-158 (quirk Q16: the reference's own formulas)
CS star_stt(const Side& s, double SK, double SM, double pt_x) {
    CS Ux;
    Ux.D = s.D * (SK - s.vn) / (SK - SM);
    Ux.tau = (s.E * (SK - s.vn) - s.pt * s.vn + pt_x * SM) / (SK - SM) - Ux.D;
    Ux.Sn = (s.Sn * (SK - s.vn) - s.pt + pt_x) / (SK - SM);
    Ux.St1 = s.St1 * (SK - s.vn) / (SK - SM);
    Ux.St2 = s.St2 * (SK - s.vn) / (SK - SM);
    Ux.Bn = s.Bc;
    double den = s.D * (SK - s.vn) * (SK - SM) - s.Bn * s.Bn;
    if (std::abs(den) > 1e-14) {
        double k = s.Bn * (SM - s.vn) / den;
        Ux.Bt1 = s.Bt1 - s.Bn * (s.vt1 * (SK - s.vn) + s.Bn * s.Bt1 / s.D) * k;
        Ux.Bt2 = s.Bt2 - s.Bn * (s.vt2 * (SK - s.vn) + s.Bn * s.Bt2 / s.D) * k;
    } else {
        Ux.Bt1 = s.Bt1; Ux.Bt2 = s.Bt2;
    }
    return Ux;
}
// alfstar_stt, hlld.cpp:160-200 (quirk Q15: the degenerate branch is overwritten)
void alfstar_stt(const CS& UxL, const CS& UxR, const Side& L, double SM, CS& UxxL, CS& UxxR) {
    UxxL.D = UxL.D; UxxL.tau = UxL.tau;
    UxxR.D = UxR.D; UxxR.tau = UxR.tau;
    UxxL.Bn = UxL.Bn; UxxR.Bn = UxR.Bn;
    double sq_DL = std::sqrt(std::abs(UxL.D));
    double sq_DR = std::sqrt(std::abs(UxR.D));
    double Bn = L.Bn;
    double den = sq_DL + sq_DR;
    if (std::abs(den) < 1e-14 || std::abs(Bn) < 1e-14) { UxxL = UxL; UxxR = UxR; }
    double sgnBn = (Bn >= 0.0) ? 1.0 : -1.0;
    double vt1xx = (sq_DL * UxL.St1 / UxL.D + sq_DR * UxR.St1 / UxR.D + sgnBn * (UxR.Bt1 - UxL.Bt1)) / den;
    double vt2xx = (sq_DL * UxL.St2 / UxL.D + sq_DR * UxR.St2 / UxR.D + sgnBn * (UxR.Bt2 - UxL.Bt2)) / den;
    double Bt1xx = (sq_DL * UxL.Bt1 / UxL.D + sq_DR * UxR.Bt1 / UxR.D +
                    sgnBn * sq_DL * sq_DR * (UxR.St1 / UxR.D - UxL.St1 / UxL.D)) / den;
    double Bt2xx = (sq_DL * UxL.Bt2 / UxL.D + sq_DR * UxR.Bt2 / UxR.D +
                    sgnBn * sq_DL * sq_DR * (UxR.St2 / UxR.D - UxL.St2 / UxL.D)) / den;
    UxxL.St1 = UxL.D * vt1xx; UxxR.St1 = UxR.D * vt1xx;
    UxxL.St2 = UxL.D * vt2xx; UxxR.St2 = UxR.D * vt2xx;
    UxxL.Bt1 = Bt1xx; UxxR.Bt1 = Bt1xx;
    UxxL.Bt2 = Bt2xx; UxxR.Bt2 = Bt2xx;
    UxxL.Sn = UxL.D * SM; UxxR.Sn = UxR.D * SM;
    const double vt1xL = UxL.St1 / UxL.D, vt2xL = UxL.St2 / UxL.D;
    const double vt1xR = UxR.St1 / UxR.D, vt2xR = UxR.St2 / UxR.D;
    double dE_L = sq_DL * sgnBn * (vt1xx * Bt1xx + vt2xx * Bt2xx - vt1xL * UxL.Bt1 - vt2xL * UxL.Bt2);
    double dE_R = sq_DR * sgnBn * (vt1xx * Bt1xx + vt2xx * Bt2xx - vt1xR * UxR.Bt1 - vt2xR * UxR.Bt2);
    UxxL.tau = UxL.tau - dE_L;
    UxxR.tau = UxR.tau + dE_R;
}
// interflux, hlld.cpp:202-215
CS interflux(const CS& FK, double SK, const CS& Ux, const Side& s) {
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
// flux, hlld.cpp:217-227
Cons dens_flux(const CS& F, double sg, int dim) {
    int t1 = (dim + 1) % 3, t2 = (dim + 2) % 3;
    Cons c = czero();
    c.D = sg * F.D; c.tau = sg * F.tau;
    c.B[dim] = sg * F.Bn; c.B[t1] = sg * F.Bt1; c.B[t2] = sg * F.Bt2;
    c.S[dim] = sg * F.Sn; c.S[t1] = sg * F.St1; c.S[t2] = sg * F.St2;
    return c;
}
// hlld, hlld.cpp:230-312 (quirks Q14, Q17)
Cons hlld(const Prim& WL, const Prim& WR, const Cons& UL, const Cons& UR, const KS& ks, double r, double th,
          const EOS& st, int dim, double& ctspeed) {
    MC mc = metric_comp(ks, r, th);
    double sg = mc.sqrtdetg;
    Side L = side_(WL, UL, mc, st, sg, dim);
    Side R = side_(WR, UR, mc, st, sg, dim);
    double SL = smin(mc.alpha * (L.vn - L.cf) - mc.beta[dim], mc.alpha * (R.vn - R.cf) - mc.beta[dim]);
    double SR = smax(mc.alpha * (L.vn + L.cf) - mc.beta[dim], mc.alpha * (R.vn + R.cf) - mc.beta[dim]);
    CS FL = physflux(L, mc, dim), FR = physflux(R, mc, dim);
    ctspeed = smax(std::abs(SL), std::abs(SR));
    Cons F = czero();
    if (SR - SL < 1e-12) F = dens_flux(hll_flux(L, R, FL, FR, SL, SR), sg, dim);
    if (SL >= 0.0) return dens_flux(FL, sg, dim);
    if (SR <= 0.0) return dens_flux(FR, sg, dim);
    double pt_x = (SR * L.pt - SL * R.pt + SL * SR * (R.Sn - L.Sn)) / (SR - SL);
    double SM_u = SR * L.E - SL * R.E - FR.tau - FR.D + FL.tau + FL.D;
    double SM_d = SR * L.D - SL * R.D - FR.D + FL.D;
    if (std::abs(SM_d) < 1e-12) return dens_flux(hll_flux(L, R, FL, FR, SL, SR), sg, dim);
    double SM = SM_u / SM_d;
    CS UxL = star_stt(L, SL, SM, pt_x);
    CS UxR = star_stt(R, SR, SM, pt_x);
    double SLx = SM - std::abs(L.Bn) / std::sqrt(std::abs(UxL.D));
    double SRx = SM + std::abs(R.Bn) / std::sqrt(std::abs(UxR.D));
    if (UxL.D <= 0.0 || UxR.D <= 0.0) return dens_flux(hll_flux(L, R, FL, FR, SL, SR), sg, dim);
    if (SM >= 0.0) {
        if (SLx >= 0.0) {
            F = dens_flux(interflux(FL, SL, UxL, L), sg, dim);
        } else {
            CS UxxL, UxxR;
            alfstar_stt(UxL, UxR, L, SM, UxxL, UxxR);
            CS FLx = interflux(FL, SL, UxL, L);
            CS Fxx = interflux(FLx, SLx, UxxL, L);
            F = dens_flux(Fxx, sg, dim);
        }
    } else {
        if (SRx >= 0.0) {
            F = dens_flux(interflux(FR, SR, UxR, R), sg, dim);
        } else {
            CS UxxL, UxxR;
            alfstar_stt(UxL, UxR, L, SM, UxxL, UxxR);
            CS FRx = interflux(FR, SR, UxR, R);
            CS Fxx = interflux(FRx, SRx, UxxR, R);
            F = dens_flux(Fxx, sg, dim);
        }
    }
    if (!std::isfinite(F.D) || !std::isfinite(F.tau) || !std::isfinite(F.S[dim % 3]) ||
        !std::isfinite(F.S[(dim + 1) % 3]) || !std::isfinite(F.S[(dim + 2) % 3]) || !std::isfinite(F.B[dim % 3]) ||
        !std::isfinite(F.B[(dim + 1) % 3]) || !std::isfinite(F.B[(dim + 2) % 3])) {
        F = dens_flux(hll_flux(L, R, FL, FR, SL, SR), sg, dim);
    }
    if (ctspeed > 1.0) {
        F = dens_flux(hll_flux(L, R, FL, FR, SL, SR), sg, dim);
        ctspeed = smin(ctspeed, 1.0);
    }
    return F;
}

// ----------------------------------------------------------------- tree
struct Block {
    int lvl, ijk[3], leaf, parent, children[8];
    double xe[17], ye[17], ze[17];
    double dx() const { return xe[1] - xe[0]; }
    double dy() const { return ye[1] - ye[0]; }
    double dz() const { return ze[1] - ze[0]; }
};
struct Geo { double r, th, phi, xc, yc, zc, vol, pvol; };

struct Tree {
    KS ks;      // tree metric (M, a, Q)
    KS cell_ks; // quirk Q2: every cell carries a default metric (M=1, a=0, Q=0), metric.hpp:30
    double gamma;
    Libm lm;
    int nb;
    std::vector<Block> blk;
    std::vector<Geo> geo;            // [nb][8000]
    std::vector<double> U, W, Bf, dU;  // SoA
    std::vector<double> F;           // [nb][3][8400][8] fluxes (zero outside computed faces, quirk Q19)
    std::vector<double> E;           // [nb][3][8820] EMFs
    std::atomic<long long> faults{0};
    int limiter = 1;  // reconfp's limiter: the reference passes van Leer (grid.cpp:261,285,309)

    EOS eos() { return EOS{gamma, &faults}; }
    double& u(int b, int v, int c) { return U[((size_t)b * 8 + v) * NCELL + c]; }
    double& w(int b, int v, int c) { return W[((size_t)b * 8 + v) * NCELL + c]; }
    double& du(int b, int v, int c) { return dU[((size_t)b * 8 + v) * NCELL + c]; }
    double* bfx(int b) { return &Bf[(size_t)b * 3 * NFACE]; }
    double* bfy(int b) { return &Bf[(size_t)b * 3 * NFACE + NFACE]; }
    double* bfz(int b) { return &Bf[(size_t)b * 3 * NFACE + 2 * NFACE]; }
    double* fx(int b) { return &F[((size_t)b * 3 + 0) * NFACE * 8]; }
    double* fy(int b) { return &F[((size_t)b * 3 + 1) * NFACE * 8]; }
    double* fz(int b) { return &F[((size_t)b * 3 + 2) * NFACE * 8]; }
    double* ex(int b) { return &E[((size_t)b * 3 + 0) * NEDGE]; }
    double* ey(int b) { return &E[((size_t)b * 3 + 1) * NEDGE]; }
    double* ez(int b) { return &E[((size_t)b * 3 + 2) * NEDGE]; }
    const Geo& gc(int b, int c) const { return geo[(size_t)b * NCELL + c]; }

    Prim getW(int b, int c) {
        Prim p; std::memset(&p, 0, sizeof p);
        p.rho = w(b, 0, c); p.eps = w(b, 1, c);
        for (int d = 0; d < 3; d++) { p.v[d] = w(b, 2 + d, c); p.B[d] = w(b, 5 + d, c); }
        return p;
    }
    void setW(int b, int c, const Prim& p) {
        w(b, 0, c) = p.rho; w(b, 1, c) = p.eps;
        for (int d = 0; d < 3; d++) { w(b, 2 + d, c) = p.v[d]; w(b, 5 + d, c) = p.B[d]; }
    }
    Cons getU(int b, int c) {
        Cons q;
        q.D = u(b, 0, c);
        for (int d = 0; d < 3; d++) q.S[d] = u(b, 1 + d, c);
        q.tau = u(b, 4, c);
        for (int d = 0; d < 3; d++) q.B[d] = u(b, 5 + d, c);
        return q;
    }
    void setU(int b, int c, const Cons& q) {
        u(b, 0, c) = q.D;
        for (int d = 0; d < 3; d++) u(b, 1 + d, c) = q.S[d];
        u(b, 4, c) = q.tau;
        for (int d = 0; d < 3; d++) u(b, 5 + d, c) = q.B[d];
    }
};

inline void getF(const double* f, int idx, Cons& q) {
    const double* p = f + 8 * idx;
    q.D = p[0]; q.S[0] = p[1]; q.S[1] = p[2]; q.S[2] = p[3]; q.tau = p[4]; q.B[0] = p[5]; q.B[1] = p[6]; q.B[2] = p[7];
}
inline void setF(double* f, int idx, const Cons& q) {
    double* p = f + 8 * idx;
    p[0] = q.D; p[1] = q.S[0]; p[2] = q.S[1]; p[3] = q.S[2]; p[4] = q.tau; p[5] = q.B[0]; p[6] = q.B[1]; p[7] = q.B[2];
}

// patch::cell_init geometry, grid.cpp:149-175
void cell_init(Tree& t, int b) {
    const Block& B = t.blk[b];
    double dx = B.dx(), dy = B.dy(), dz = B.dz();
    for (int i = -GH; i < BLK + GH; i++)
        for (int j = -GH; j < BLK + GH; j++)
            for (int k = -GH; k < BLK + GH; k++) {
                Geo& g = t.geo[(size_t)b * NCELL + lindx(i, j, k)];
                g.xc = B.xe[0] + (i + 0.5) * dx;
                g.yc = B.ye[0] + (j + 0.5) * dy;
                g.zc = B.ze[0] + (k + 0.5) * dz;
                g.vol = dx * dy * dz;
                g.r = std::sqrt(g.xc * g.xc + g.yc * g.yc + g.zc * g.zc);
                if (g.r < 1e-10) g.r = 1e-10;
                g.th = std::acos(smax(-1.0, smin(1.0, g.zc / g.r)));
                g.phi = std::atan2(g.yc, g.xc);
                MC mc = metric_comp(t.cell_ks, g.r, g.th);
                g.pvol = g.vol / mc.sqrtdetg;
            }
}

// amrtree::pvfs, grid.cpp:133-146
Prim pvfs(const EOS& st, double r) {
    Prim W; std::memset(&W, 0, sizeof W);
    double rfscal = 1.0 / std::pow(std::sqrt(smax(r, 1e-10) / 1.0), 3.0);
    W.rho = 1e-7 * rfscal;
    W.eps = 1e-5 * rfscal;
    W.p = st.press(W.rho, W.eps);
    W.h = st.enth(W.rho, W.eps);
    W.T = 1e-12;
    return W;
}

// amrtree::nbhd, grid.cpp:361-386 (linear scan in quilt order; quirk Q26)
int nbhd(const Tree& t, int p, int dim, int side) {
    const Block& P = t.blk[p];
    int nb[3] = {P.ijk[0], P.ijk[1], P.ijk[2]};
    nb[dim] += side;
    for (int q = 0; q < t.nb; q++) {
        const Block& B = t.blk[q];
        if (B.lvl == P.lvl && B.ijk[0] == nb[0] && B.ijk[1] == nb[1] && B.ijk[2] == nb[2]) return q;
    }
    if (P.lvl > 0) {
        int c[3];
        for (int i = 0; i < 3; i++) c[i] = P.ijk[i] / 2;
        c[dim] += side;
        for (int q = 0; q < t.nb; q++) {
            const Block& B = t.blk[q];
            if (B.lvl == P.lvl - 1 && B.ijk[0] == c[0] && B.ijk[1] == c[1] && B.ijk[2] == c[2]) return q;
        }
    }
    return -1;
}

inline void cell_ijk(int dim, int a, int t1, int t2, int& i, int& j, int& k) {
    if (dim == 0) { i = a; j = t1; k = t2; }
    else if (dim == 1) { i = t1; j = a; k = t2; }
    else { i = t1; j = t2; k = a; }
}

// amrtree::gh_copy, grid.cpp:390-425 (quirk Q27: edges and corners skipped)
void gh_copy(Tree& t, int p, int nb, int dim, int side) {
    int pstart = (side == 1) ? BLK : -GH;
    int nbstart = (side == 1) ? 0 : BLK - GH;
    for (int g = 0; g < GH; g++)
        for (int t1 = 0; t1 < BLK; t1++)
            for (int t2 = 0; t2 < BLK; t2++) {
                int pi, pj, pk, ni, nj, nk;
                cell_ijk(dim, pstart + g, t1, t2, pi, pj, pk);
                cell_ijk(dim, nbstart + g, t1, t2, ni, nj, nk);
                int cp = lindx(pi, pj, pk), cn = lindx(ni, nj, nk);
                for (int v = 0; v < 8; v++) { t.u(p, v, cp) = t.u(nb, v, cn); t.w(p, v, cp) = t.w(nb, v, cn); }
            }
}

// amrtree::gh_prolong, grid.cpp:427-492 (quirk Q25: coarse index ignores octant and side)
void gh_prolong(Tree& t, int p, int nc, int dim, int side) {
    EOS st = t.eos();
    int pstart = (side == 1) ? BLK : -GH;
    for (int g = 0; g < GH; g++)
        for (int t1 = 0; t1 < BLK; t1++)
            for (int t2 = 0; t2 < BLK; t2++) {
                int fi, fj, fk;
                cell_ijk(dim, pstart + g, t1, t2, fi, fj, fk);
                int ci = (dim == 0) ? g / 2 : t1 / 2;
                int cj = (dim == 1) ? g / 2 : (dim == 0 ? t1 / 2 : t2 / 2);
                int ck = (dim == 2) ? g / 2 : t2 / 2;
                double offg = (g % 2 == 0) ? -0.25 : 0.25;
                double offt1 = (t1 % 2 == 0) ? -0.25 : 0.25;
                double offt2 = (t2 % 2 == 0) ? -0.25 : 0.25;
                double on, ot1, ot2;
                if (dim == 0) { on = offg; ot1 = offt1; ot2 = offt2; }
                else if (dim == 1) { on = offt1; ot1 = offg; ot2 = offt2; }
                else { on = offt1; ot1 = offt2; ot2 = offg; }
                int cc = lindx(ci, cj, ck);
                int cxm = lindx(ci - 1, cj, ck), cxp = lindx(ci + 1, cj, ck);
                int cym = lindx(ci, cj - 1, ck), cyp = lindx(ci, cj + 1, ck);
                int czm = lindx(ci, cj, ck - 1), czp = lindx(ci, cj, ck + 1);
                int cf = lindx(fi, fj, fk);
                for (int v = 0; v < 8; v++) {
                    double vc = t.u(nc, v, cc);
                    double sn = minmod(vc - t.u(nc, v, cxm), t.u(nc, v, cxp) - vc);
                    double st1 = minmod(vc - t.u(nc, v, cym), t.u(nc, v, cyp) - vc);
                    double st2 = minmod(vc - t.u(nc, v, czm), t.u(nc, v, czp) - vc);
                    t.u(p, v, cf) = vc + sn * on + st1 * ot1 + st2 * ot2;
                }
                const Geo& G = t.gc(p, cf);
                Prim pv;
                Cons U = t.getU(p, cf);
                if (ctop(t.ks, st, U, G.r, G.th, pv)) {
                    t.setW(p, cf, pv);
                } else {
                    Prim fl = pvfs(st, G.r);
                    for (int d = 0; d < 3; d++) fl.B[d] = t.w(p, 5 + d, cf);
                    t.setW(p, cf, fl);
                    t.setU(p, cf, ptoc(t.ks, st, fl, G.r, G.th));
                }
            }
}

// amrtree::gh_bndy, grid.cpp:494-522 (quirk Q33: mirror-order copy)
void gh_bndy(Tree& t, int p, int dim, int side) {
    int pstart = (side == 1) ? BLK : -GH;
    for (int g = 0; g < GH; g++)
        for (int t1 = 0; t1 < BLK; t1++)
            for (int t2 = 0; t2 < BLK; t2++) {
                int gi, gj, gk, ai, aj, ak;
                cell_ijk(dim, pstart + g, t1, t2, gi, gj, gk);
                cell_ijk(dim, (side == 1) ? BLK - 1 - g : g, t1, t2, ai, aj, ak);
                int cg = lindx(gi, gj, gk), ca = lindx(ai, aj, ak);
                for (int v = 0; v < 8; v++) { t.u(p, v, cg) = t.u(p, v, ca); t.w(p, v, cg) = t.w(p, v, ca); }
            }
}

inline bool is_ghost(int i, int j, int k) { return i < 0 || i >= BLK || j < 0 || j >= BLK || k < 0 || k >= BLK; }

// amrtree::ghosts, grid.cpp:524-571 (quirk Q29: failure keeps v; T uses eps)
void ghosts(Tree& t, int p) {
    EOS st = t.eos();
    for (int dim = 0; dim < 3; dim++)
        for (int side : {-1, 1}) {
            int nb = nbhd(t, p, dim, side);
            if (nb >= 0 && t.blk[nb].lvl == t.blk[p].lvl) gh_copy(t, p, nb, dim, side);
            else if (nb >= 0 && t.blk[nb].lvl < t.blk[p].lvl) gh_prolong(t, p, nb, dim, side);
            else if (nb < 0) gh_bndy(t, p, dim, side);
        }
    for (int i = -GH; i < BLK + GH; i++)
        for (int j = -GH; j < BLK + GH; j++)
            for (int k = -GH; k < BLK + GH; k++) {
                if (!is_ghost(i, j, k)) continue;
                int c = lindx(i, j, k);
                const Geo& G = t.gc(p, c);
                Prim pv;
                if (ctop(t.ks, st, t.getU(p, c), G.r, G.th, pv)) {
                    t.setW(p, c, pv);
                    st.press(pv.rho, pv.eps);
                    st.enth(pv.rho, pv.eps);
                    st.temp(pv.rho, pv.eps);
                } else {
                    Prim fl = pvfs(st, G.r);
                    Prim Wc = t.getW(p, c);
                    Wc.rho = fl.rho;
                    Wc.eps = fl.eps;
                    t.setW(p, c, Wc);
                    t.setU(p, c, ptoc(t.ks, st, Wc, G.r, G.th));
                }
            }
}

// c2p on every cell of every leaf, integrator.cpp:201-222 (quirk Q30: ghosts twice)
void c2p_leaves(Tree& t) {
    EOS st = t.eos();
#pragma omp parallel for schedule(dynamic)
    for (int p = 0; p < t.nb; p++) {
        if (!t.blk[p].leaf) continue;
        for (int c = 0; c < NCELL; c++) {
            const Geo& G = t.gc(p, c);
            Prim pv;
            if (ctop(t.ks, st, t.getU(p, c), G.r, G.th, pv)) {
                t.setW(p, c, pv);
            } else {
                Prim fl = pvfs(st, G.r);
                for (int d = 0; d < 3; d++) fl.B[d] = t.w(p, 5 + d, c);
                t.setW(p, c, fl);
                t.setU(p, c, ptoc(t.ks, st, fl, G.r, G.th));
            }
        }
    }
}

// reconfp, reconstruction.cpp:40-97 (causality uses the cell metric, quirk Q2)
void reconfp(Tree& t, int p, int i, int j, int k, int dim, int lim, Prim& L, Prim& R) {
    auto cidx = [&](int off) {
        return lindx(i + (dim == 0) * off, j + (dim == 1) * off, k + (dim == 2) * off);
    };
    Prim Wm = t.getW(p, cidx(-1)), Wc = t.getW(p, cidx(0)), Wp = t.getW(p, cidx(1)), Wpp = t.getW(p, cidx(2));
    std::memset(&L, 0, sizeof L); std::memset(&R, 0, sizeof R);
    double fL, fR;
    recon_scal(t.lm.log(Wm.rho), t.lm.log(Wc.rho), t.lm.log(Wp.rho), t.lm.log(Wpp.rho), lim, fL, fR);
    L.rho = t.lm.exp(fL); R.rho = t.lm.exp(fR);
    recon_scal(t.lm.log(Wm.eps), t.lm.log(Wc.eps), t.lm.log(Wp.eps), t.lm.log(Wpp.eps), lim, fL, fR);
    L.eps = t.lm.exp(fL); R.eps = t.lm.exp(fR);
    for (int d = 0; d < 3; d++) {
        recon_scal(Wm.v[d], Wc.v[d], Wp.v[d], Wpp.v[d], lim, fL, fR);
        L.v[d] = fL; R.v[d] = fR;
    }
    auto causalv = [&](Prim& Wf, int c) {
        double v2 = 0.0;
        const Geo& G = t.gc(p, c);
        MC mc = metric_comp(t.cell_ks, G.r, G.th);
        for (int m = 0; m < 3; m++)
            for (int n = 0; n < 3; n++) v2 += mc.gam[m][n] * Wf.v[m] * Wf.v[n];
        if (v2 >= 1.0) {
            double s = (1 - 1e-10) / std::sqrt(v2);
            for (int a = 0; a < 3; a++) Wf.v[a] *= s;
        }
    };
    causalv(L, cidx(0));
    causalv(R, cidx(1));
    for (int d = 0; d < 3; d++) {
        if (d == dim) {
            L.B[d] = Wc.B[d]; R.B[d] = Wp.B[d];
        } else {
            recon_scal(Wm.B[d], Wc.B[d], Wp.B[d], Wpp.B[d], lim, fL, fR);
            L.B[d] = fL; R.B[d] = fR;
        }
    }
}

// patch::fluxcomp, grid.cpp:255-329 (quirk Q1: hlld always called with dim 0)
void fluxcomp(Tree& t, int p) {
    EOS st = t.eos();
    double* Bfs[3] = {t.bfx(p), t.bfy(p), t.bfz(p)};
    double* Fs[3] = {t.fx(p), t.fy(p), t.fz(p)};
    for (int dim = 0; dim < 3; dim++) {
        int lo[3] = {0, 0, 0};
        lo[dim] = -1;
        for (int i = lo[0]; i < BLK; i++)
            for (int j = lo[1]; j < BLK; j++)
                for (int k = lo[2]; k < BLK; k++) {
                    Prim L, R;
                    reconfp(t, p, i, j, k, dim, t.limiter, L, R);
                    int bidx = dim == 0 ? ixf(i + 1, j, k) : dim == 1 ? iyf(i, j + 1, k) : izf(i, j, k + 1);
                    L.B[dim] = Bfs[dim][bidx];
                    R.B[dim] = Bfs[dim][bidx];
                    int cL = lindx(i, j, k);
                    int cR = lindx(i + (dim == 0), j + (dim == 1), k + (dim == 2));
                    double rf = (t.gc(p, cL).r + t.gc(p, cR).r) / 2;
                    double thf = (t.gc(p, cL).th + t.gc(p, cR).th) / 2;
                    Cons UL = ptoc(t.ks, st, L, rf, thf);
                    Cons UR = ptoc(t.ks, st, R, rf, thf);
                    double ct;
                    Cons Fl = hlld(L, R, UL, UR, t.ks, rf, thf, st, 0, ct);
                    int fidx = dim == 0 ? ixf(i, j, k) : dim == 1 ? iyf(i, j, k) : izf(i, j, k);
                    setF(Fs[dim], fidx, Fl);
                }
    }
}

// constrans::emfcomp / Bfupdate / f2cB, transport.cpp:11-75 (quirk Q21)
void emfcomp(Tree& t, int p) {
    const double *Fx = t.fx(p), *Fy = t.fy(p), *Fz = t.fz(p);
    double *Ex = t.ex(p), *Ey = t.ey(p), *Ez = t.ez(p);
    auto FB = [](const double* F, int idx, int d) { return F[8 * idx + 5 + d]; };
    for (int i = -GH; i < BLK + GH - 1; i++)
        for (int j = -GH; j < BLK + GH - 1; j++)
            for (int k = -GH; k < BLK + GH - 1; k++) {
                Ex[iex(i, j, k)] = (FB(Fy, iyf(i, j, k), 2) + FB(Fy, iyf(i, j, k + 1), 2) - FB(Fz, izf(i, j, k), 1) -
                                    FB(Fz, izf(i, j + 1, k), 1)) / 4;
                Ey[izf(i, j, k)] = (FB(Fz, izf(i, j, k), 0) + FB(Fz, izf(i + 1, j, k), 0) - FB(Fx, ixf(i, j, k), 2) -
                                    FB(Fx, ixf(i, j, k + 1), 2)) / 4;
                Ez[iyf(i, j, k)] = (FB(Fx, ixf(i, j, k), 1) + FB(Fx, ixf(i, j + 1, k), 1) - FB(Fy, iyf(i, j, k), 0) -
                                    FB(Fy, iyf(i + 1, j, k), 0)) / 4;
            }
}
void Bfupdate(Tree& t, int p, double dt) {
    const Block& B = t.blk[p];
    double dx = B.dx(), dy = B.dy(), dz = B.dz();
    double *Ex = t.ex(p), *Ey = t.ey(p), *Ez = t.ez(p);
    double *Bx = t.bfx(p), *By = t.bfy(p), *Bz = t.bfz(p);
    for (int i = -GH + 1; i < BLK + GH - 1; i++)
        for (int j = -GH + 1; j < BLK + GH - 1; j++)
            for (int k = -GH + 1; k < BLK + GH - 1; k++) {
                double dEz_dy = (Ez[iyf(i, j, k)] - Ez[iyf(i, j - 1, k)]) / dy;
                double dEy_dz = (Ey[izf(i, j, k)] - Ey[izf(i, j, k - 1)]) / dz;
                Bx[ixf(i, j, k)] += dt * (dEz_dy - dEy_dz);
                double dEz_dx = (Ez[iyf(i, j, k)] - Ez[iyf(i - 1, j, k)]) / dx;
                double dEx_dz = (Ex[iex(i, j, k)] - Ex[iex(i, j, k - 1)]) / dz;
                By[iyf(i, j, k)] += dt * (dEx_dz - dEz_dx);
                double dEy_dx = (Ey[izf(i, j, k)] - Ey[izf(i - 1, j, k)]) / dx;
                double dEx_dy = (Ex[iex(i, j, k)] - Ex[iex(i, j - 1, k)]) / dy;
                Bz[izf(i, j, k)] += dt * (dEy_dx - dEx_dy);
            }
}
void f2cB(Tree& t, int p) {
    double *Bx = t.bfx(p), *By = t.bfy(p), *Bz = t.bfz(p);
    for (int i = -GH + 1; i < BLK + GH; i++)
        for (int j = -GH + 1; j < BLK + GH; j++)
            for (int k = -GH + 1; k < BLK + GH; k++) {
                int c = lindx(i, j, k);
                double b0 = (Bx[ixf(i, j, k)] + Bx[ixf(i - 1, j, k)]) / 2;
                double b1 = (By[iyf(i, j, k)] + By[iyf(i, j - 1, k)]) / 2;
                double b2 = (Bz[izf(i, j, k)] + Bz[izf(i, j, k - 1)]) / 2;
                t.w(p, 5, c) = b0; t.w(p, 6, c) = b1; t.w(p, 7, c) = b2;
                const Geo& G = t.gc(p, c);
                double sg = metric_comp(t.cell_ks, G.r, G.th).sqrtdetg;
                t.u(p, 5, c) = sg * b0; t.u(p, 6, c) = sg * b1; t.u(p, 7, c) = sg * b2;
            }
}
double maxdivB(Tree& t, int p) {
    const Block& B = t.blk[p];
    double dx = B.dx(), dy = B.dy(), dz = B.dz();
    double *Bx = t.bfx(p), *By = t.bfy(p), *Bz = t.bfz(p);
    double mx = 0;
    for (int i = -GH + 1; i < BLK + GH - 1; i++)
        for (int j = -GH + 1; j < BLK + GH - 1; j++)
            for (int k = -GH + 1; k < BLK + GH - 1; k++) {
                double dv = (Bx[ixf(i, j, k)] - Bx[ixf(i - 1, j, k)]) / dx + (By[iyf(i, j, k)] - By[iyf(i, j - 1, k)]) / dy +
                            (Bz[izf(i, j, k)] - Bz[izf(i, j, k - 1)]) / dz;
                mx = smax(mx, std::abs(dv));
            }
    return mx;
}

// rk2integrator::div_flux, integrator.cpp:18-38 (cell metric, quirk Q4)
Cons div_flux(Tree& t, int p, int i, int j, int k) {
    const Block& B = t.blk[p];
    double dx = B.dx(), dy = B.dy(), dz = B.dz();
    const Geo& G = t.gc(p, lindx(i, j, k));
    double sg = metric_comp(t.cell_ks, G.r, G.th).sqrtdetg;
    Cons FxR, FxL, FyR, FyL, FzR, FzL;
    getF(t.fx(p), ixf(i, j, k), FxR); getF(t.fx(p), ixf(i - 1, j, k), FxL);
    getF(t.fy(p), iyf(i, j, k), FyR); getF(t.fy(p), iyf(i, j - 1, k), FyL);
    getF(t.fz(p), izf(i, j, k), FzR); getF(t.fz(p), izf(i, j, k - 1), FzL);
    Cons dU = czero();
    dU.D = -((FxR.D - FxL.D) / dx + (FyR.D - FyL.D) / dy + (FzR.D - FzL.D) / dz) / sg;
    dU.tau = -((FxR.tau - FxL.tau) / dx + (FyR.tau - FyL.tau) / dy + (FzR.tau - FzL.tau) / dz) / sg;
    for (int d = 0; d < 3; d++) {
        dU.S[d] = -((FxR.S[d] - FxL.S[d]) / dx + (FyR.S[d] - FyL.S[d]) / dy + (FzR.S[d] - FzL.S[d]) / dz) / sg;
        dU.B[d] = 0.0;
    }
    return dU;
}

// rk2integrator::source, integrator.cpp:41-153 (cell metric; spherical
// components land in the Cartesian slots)
Cons source(Tree& t, int p, int i, int j, int k) {
    EOS st = t.eos();
    int c = lindx(i, j, k);
    const Geo& G = t.gc(p, c);
    MC mc = metric_comp(t.cell_ks, G.r, G.th);
    Prim W = t.getW(p, c);
    double v2 = 0.0;
    double vlow[3] = {}, Blow[3] = {};
    for (int m = 0; m < 3; m++) {
        for (int n = 0; n < 3; n++) {
            vlow[m] += mc.gam[m][n] * W.v[n];
            Blow[m] += mc.gam[m][n] * W.B[n];
        }
        v2 += vlow[m] * W.v[m];
    }
    v2 = smin(v2, 1 - 1e-10);
    double ltz2 = 1.0 / (1.0 - v2);
    double ltz = std::sqrt(ltz2);
    double Bv = 0.0, Bsq = 0.0;
    for (int m = 0; m < 3; m++) { Bv += Blow[m] * W.v[m]; Bsq += Blow[m] * W.B[m]; }
    double b2 = Bv * Bv + Bsq / ltz2;
    double rho = W.rho, eps = W.eps;
    double px = st.press(rho, eps);
    double h = st.enth(rho, eps);
    double u4[4];
    u4[0] = ltz / mc.alpha;
    for (int m = 0; m < 3; m++) u4[m + 1] = ltz * (W.v[m] - mc.beta[m] / mc.alpha);
    double b4[4];
    b4[0] = ltz * Bv / mc.alpha;
    for (int m = 0; m < 3; m++) b4[m + 1] = (W.B[m] + mc.alpha * b4[0] * u4[m + 1]) / ltz;
    double T44[4][4];
    double rhohb2 = rho * h + b2;
    for (int m = 0; m < 4; m++)
        for (int n = 0; n < 4; n++) T44[m][n] = rhohb2 * u4[m] * u4[n] + px * mc.gi[m][n] - b4[m] * b4[n];
    double Gl[4][4][4] = {};
    for (int n = 0; n < 4; n++)
        for (int z = 0; z < 4; z++)
            for (int l = 0; l < 4; l++)
                for (int s = 0; s < 4; s++) Gl[n][z][l] += mc.g[n][s] * mc.G[s][z][l];
    double src[4] = {};
    for (int m = 0; m < 4; m++) {
        for (int n = 0; n < 4; n++)
            for (int l = 0; l < 4; l++) src[m] += T44[n][l] * Gl[m][n][l];
        src[m] *= mc.sqrtdetg;
    }
    Cons S = czero();
    S.D = 0.0;
    S.tau = src[0];
    S.S[0] = src[1]; S.S[1] = src[2]; S.S[2] = src[3];
    return S;
}

// rk2integrator::rkstg update phase, integrator.cpp:238-259 (quirks Q20, Q22, Q31)
void update(Tree& t, int p, double dt, int stage) {
    for (int i = 0; i < BLK; i++)
        for (int j = 0; j < BLK; j++)
            for (int k = 0; k < BLK; k++) {
                int c = lindx(i, j, k);
                Cons a = div_flux(t, p, i, j, k), s = source(t, p, i, j, k);
                Cons dU;
                dU.D = a.D + s.D; dU.tau = a.tau + s.tau;
                for (int d = 0; d < 3; d++) { dU.S[d] = a.S[d] + s.S[d]; dU.B[d] = a.B[d] + s.B[d]; }
                Cons U = t.getU(p, c);
                Cons nU;
                if (stage == 0) {
                    for (int v = 0; v < 8; v++) t.du(p, v, c) = t.u(p, v, c);
                    nU.D = U.D + dU.D * dt; nU.tau = U.tau + dU.tau * dt;
                    for (int d = 0; d < 3; d++) { nU.S[d] = U.S[d] + dU.S[d] * dt; nU.B[d] = U.B[d] + dU.B[d] * dt; }
                } else {
                    Cons A;
                    A.D = t.du(p, 0, c); A.S[0] = t.du(p, 1, c); A.S[1] = t.du(p, 2, c); A.S[2] = t.du(p, 3, c);
                    A.tau = t.du(p, 4, c); A.B[0] = t.du(p, 5, c); A.B[1] = t.du(p, 6, c); A.B[2] = t.du(p, 7, c);
                    nU.D = ((A.D + U.D) + dU.D * dt) * 0.5;
                    nU.tau = ((A.tau + U.tau) + dU.tau * dt) * 0.5;
                    for (int d = 0; d < 3; d++) {
                        nU.S[d] = ((A.S[d] + U.S[d]) + dU.S[d] * dt) * 0.5;
                        nU.B[d] = ((A.B[d] + U.B[d]) + dU.B[d] * dt) * 0.5;
                    }
                }
                const Geo& G = t.gc(p, c);
                double sg = metric_comp(t.cell_ks, G.r, G.th).sqrtdetg;
                for (int d = 0; d < 3; d++) nU.B[d] = sg * t.w(p, 5 + d, c);
                t.setU(p, c, nU);
            }
}

void ghosts_all(Tree& t) {
    for (int p = 0; p < t.nb; p++)
        if (t.blk[p].leaf) ghosts(t, p);
}

void rkstg(Tree& t, double dt, int stage) {
    ghosts_all(t);
    c2p_leaves(t);
#pragma omp parallel for schedule(dynamic)
    for (int p = 0; p < t.nb; p++) fluxcomp(t, p);
#pragma omp parallel for schedule(dynamic)
    for (int p = 0; p < t.nb; p++) { emfcomp(t, p); Bfupdate(t, p, dt); f2cB(t, p); }
#pragma omp parallel for schedule(dynamic)
    for (int p = 0; p < t.nb; p++) update(t, p, dt, stage);
}

// rk2integrator::dtcomp, integrator.cpp:156-189 (quirk Q23)
double dtcomp(Tree& t, double cfl) {
    EOS st = t.eos();
    double dt_min = std::numeric_limits<double>::max();
#pragma omp parallel for reduction(min : dt_min) schedule(dynamic)
    for (int p = 0; p < t.nb; p++) {
        if (!t.blk[p].leaf) continue;
        const Block& B = t.blk[p];
        const double ds[3] = {B.dx(), B.dy(), B.dz()};
        for (int i = 0; i < BLK; i++)
            for (int j = 0; j < BLK; j++)
                for (int k = 0; k < BLK; k++) {
                    int c = lindx(i, j, k);
                    const Geo& G = t.gc(p, c);
                    double rho = t.w(p, 0, c), eps = t.w(p, 1, c);
                    double cs = std::sqrt(st.cs2(rho, eps));
                    MC mc = metric_comp(t.cell_ks, G.r, G.th);
                    for (int d = 0; d < 3; d++) {
                        double v = t.w(p, 2 + d, c);
                        double ld_p = std::abs(mc.alpha * (v + cs) - mc.beta[d]);
                        double ld_m = std::abs(mc.alpha * (v - cs) - mc.beta[d]);
                        double vmax = smax(ld_p, ld_m);
                        if (vmax > 1e-14) dt_min = smin(dt_min, cfl * ds[d] / vmax);
                    }
                }
    }
    return dt_min;
}

// patch::floors, grid.cpp:331-358
void floors(Tree& t) {
    EOS st = t.eos();
#pragma omp parallel for schedule(dynamic)
    for (int p = 0; p < t.nb; p++) {
        for (int i = 0; i < BLK; i++)
            for (int j = 0; j < BLK; j++)
                for (int k = 0; k < BLK; k++) {
                    int c = lindx(i, j, k);
                    const Geo& G = t.gc(p, c);
                    double D = t.u(p, 0, c), tau = t.u(p, 4, c);
                    bool bhl = !std::isfinite(D) || !std::isfinite(tau) || D <= 0.0 || tau <= 0.0;
                    if (!bhl) {
                        Prim pv;
                        bhl = !ctop(t.ks, st, t.getU(p, c), G.r, G.th, pv);
                        if (!bhl) {
                            bool hl = !(pv.rho <= 0.0 || pv.eps <= 0.0);
                            if (hl) t.setW(p, c, pv);
                        }
                    }
                    Prim fl = pvfs(st, G.r);
                    if (bhl || t.w(p, 0, c) <= fl.rho) {
                        for (int d = 0; d < 3; d++) fl.B[d] = t.w(p, 5 + d, c);
                        t.setW(p, c, fl);
                        t.setU(p, c, ptoc(t.ks, st, fl, G.r, G.th));
                    }
                }
    }
}

double step(Tree& t, double cfl) {
    double dt = dtcomp(t, cfl);
    rkstg(t, dt, 0);
    rkstg(t, dt, 1);
    floors(t);
    return dt;
}

// ---------------------------------------------------------- initial data
// The set-up of main.cpp:94-159: amrtree's constructor state, fm_init, the
// global rho_max, B_pot_init on every patch, ptoc of the active zones.  The
// functions the initial data calls (initconds.cpp, and metric::comp inside it)
// come from IcLibm: mode 0 = the C library (the reference), mode 1 = correctly
// rounded (binary128 rounded once; pow's exact cases y = 2 and y = -1 are the
// IEEE operations x*x and 1/x, which are correctly rounded, as in mode 0).  The geometry the
// tree already holds (cell r, theta, the floor state, the metric inside ptoc
// and f2cB) is the C library's in both modes.
extern "C" __float128 sinq(__float128);
extern "C" __float128 cosq(__float128);
extern "C" __float128 powq(__float128, __float128);
extern "C" __float128 atan2q(__float128, __float128);
extern "C" __float128 acosq(__float128);
double crq_sin(double x) { return (double)sinq((__float128)x); }
double crq_cos(double x) { return (double)cosq((__float128)x); }
double crq_pow(double x, double y) {
    if (y == 2.0) return x * x;
    if (y == -1.0) return 1.0 / x;
    return (double)powq((__float128)x, (__float128)y);
}
// the C library's pow as the reference's compiled calls see it: GCC and
// clang fold pow(x, 2.0) into x*x and pow(x, -1.0) into 1.0/x (both exact
// rewrites) wherever the exponent is a literal, which covers every such
// exponent of the initial data
double lib_pow(double x, double y) {
    if (y == 2.0) return x * x;
    if (y == -1.0) return 1.0 / x;
    return ::pow(x, y);
}
double crq_atan2(double y, double x) { return (double)atan2q((__float128)y, (__float128)x); }
double crq_acos(double x) { return (double)acosq((__float128)x); }
struct IcLibm {
    double (*sin)(double);
    double (*cos)(double);
    double (*pow)(double, double);
    double (*atan2)(double, double);
    double (*acos)(double);
    double (*exp)(double);
    double (*log)(double);
};
IcLibm ic_libm_for(int mode) {
    typedef double (*F1)(double);
    typedef double (*F2)(double, double);
    if (mode == 1) return {crq_sin, crq_cos, crq_pow, crq_atan2, crq_acos, cr_exp, cr_log};
    return {static_cast<F1>(::sin), static_cast<F1>(::cos), lib_pow, static_cast<F2>(::atan2),
            static_cast<F1>(::acos), static_cast<F1>(::exp), static_cast<F1>(::log)};
}

// the parts of metric::comp the torus reads: g_tt, g_tphi, g_phiphi
// (gcova, spacetime.cpp:20-34), alpha (lapse, 93-95) and beta^phi (shift,
// 96-100, with g^{t phi} = 0 from gcntr, 35-49)
struct KsIc { double g00, g03, g33, alpha, beta2; };
KsIc ks_ic(const IcLibm& L, const KS& k, double r, double th) {
    const double M = k.M, a = k.a, Q = k.Q;
    auto Sigma = [&]() { return r * r + a * a * L.pow(L.cos(th), 2); };
    auto H = [&]() { return (2 * M * r - Q * Q) * L.pow(Sigma(), -1); };
    KsIc m;
    m.g00 = -1 + H();
    m.g03 = -H() * a * L.pow(L.sin(th), 2);
    m.g33 = (r * r + a * a + H() * a * a * L.pow(L.sin(th), 2)) * L.pow(L.sin(th), 2);
    m.alpha = L.pow(1 + H(), -1.0 / 2.0);
    const double gi00 = -1 - H(), gi03 = 0.0;
    m.beta2 = -gi03 / gi00;
    return m;
}
// L0_comp, initconds.cpp:121-141; false where the reference throws
bool ic_L0(const IcLibm& L, const KS& ks, double r_max, double M, double& l0) {
    const double dr = 1e-4 * M;
    const KsIc mc0 = ks_ic(L, ks, r_max, M_PI / 2.0);
    const KsIc mcp = ks_ic(L, ks, r_max + dr, M_PI / 2.0);
    const KsIc mcm = ks_ic(L, ks, r_max - dr, M_PI / 2.0);
    double gtt = mc0.g00, gtp = mc0.g03, gpp = mc0.g33;
    const double dgtt = (mcp.g00 - mcm.g00) / (2.0 * dr);
    const double dgtp = (mcp.g03 - mcm.g03) / (2.0 * dr);
    const double dgpp = (mcp.g33 - mcm.g33) / (2.0 * dr);
    double A = -gtp * (gpp * (dgtt * dgpp + 2 * dgtp * dgtp) + gtt * dgpp * dgpp) +
               std::sqrt((dgtp * dgtp - dgtt * dgpp) *
                         L.pow((gpp * (gpp * dgtt - gtt * dgpp) + 2 * gtp * gtp * dgpp - 2 * gtp * gpp * dgtp), 2)) +
               gpp * gpp * dgtt * dgtp + gtt * gpp * dgtp * dgpp + 2 * gtp * gtp * dgtp * dgpp;
    double B = dgtt * (gpp * gpp * dgtt + 4 * gtp * gtp * dgpp - 4 * gtp * gpp * dgtp) -
               2 * gtt * (gpp * (dgtt * dgpp - 2 * dgtp * dgtp) + 2 * gtp * dgtp * dgpp) + gtt * gtt * dgpp * dgpp;
    if (std::abs(B) < 1e-14) return false;
    l0 = A / B;
    return true;
}
// lnrf_, lnh_pot, W_in_, initconds.cpp:151-183
struct IcLnrf { double e2nu, e2psi, enu, epsi, omega; };
IcLnrf ic_lnrf(const KsIc& mc) {
    IcLnrf lp;
    lp.e2nu = mc.g03 * mc.g03 / mc.g33 - mc.g00;
    lp.e2psi = mc.g33;
    lp.enu = std::sqrt(lp.e2nu);
    lp.epsi = std::sqrt(lp.e2psi);
    lp.omega = -mc.g03 / mc.g33;
    return lp;
}
double ic_lnh(const IcLibm& L, const KsIc& mc, double l, double W_in) {
    const IcLnrf lp = ic_lnrf(mc);
    if (lp.e2nu <= 0.0) return -1e30;
    return -0.5 * std::sqrt((4 * l * l * lp.e2nu) / lp.e2psi + 1) +
           0.5 * L.log((std::sqrt((4 * l * l * lp.e2nu) / lp.e2psi + 1) + 1) / lp.e2nu) - l * lp.omega + W_in;
}
double ic_W_in(const IcLibm& L, const KsIc& mc, double l) {
    const IcLnrf lp = ic_lnrf(mc);
    if (lp.e2nu <= 0.0) return -1e30;
    return 0.5 * std::sqrt((4 * l * l * lp.e2nu) / lp.e2psi + 1) -
           0.5 * L.log((std::sqrt((4 * l * l * lp.e2nu) / lp.e2psi + 1) + 1) / lp.e2nu) + l * lp.omega;
}
// disctherm, initconds.cpp:186-203
struct IcTherm { double rho, p, eps; };
IcTherm ic_therm(const IcLibm& L, double lnH, double Gamma, double K) {
    IcTherm th{0.0, 0.0, 0.0};
    if (lnH <= 0.0) return th;
    const double H = L.exp(lnH);
    const double arg = (Gamma - 1) * (H - 1) / (Gamma * K);
    if (arg <= 0.0) return th;
    th.rho = L.pow(arg, 1.0 / (Gamma - 1.0));
    th.p = K * L.pow(arg, Gamma / (Gamma - 1.0));
    th.eps = th.p / ((Gamma - 1.0) * th.rho);
    return th;
}
// velcomp, initconds.cpp:206-225
void ic_vel(const KsIc& mc, const IcLnrf& lp, double l, double xc, double yc, double v[3]) {
    v[0] = v[1] = v[2] = 0.0;
    const double u_phi = std::sqrt(0.5 * (std::sqrt(1.0 + 4.0 * l * l * lp.e2nu / lp.e2psi) - 1.0));
    const double uphi = u_phi / lp.epsi + (lp.omega / lp.enu) * std::sqrt(1.0 + u_phi * u_phi);
    const double gtt = mc.g00, gtp = mc.g03, gpp = mc.g33;
    const double B = 2.0 * gtp * uphi;
    const double C = gpp * uphi * uphi + 1.0;
    const double D = B * B - 4.0 * gtt * C;
    if (D < 0.0) return;
    const double ut = (-B - std::sqrt(D)) / (2.0 * gtt);
    const double Omega = uphi / ut;
    const double vphi = Omega - mc.beta2 / mc.alpha;
    const double rcyl = std::sqrt(xc * xc + yc * yc);
    if (rcyl < 1e-10) return;
    v[0] = -vphi * yc;
    v[1] = vphi * xc;
}
// fm_init, initconds.cpp:228-591 (live code).  0 = placed, 1 = L0_comp
// throws, 2 = the reference prints an error and returns.
int ic_fm_init(Tree& t, const IcLibm& L) {
    const double M = t.ks.M, a = t.ks.a, Q = t.ks.Q, Gamma = t.gamma;
    const double r_max = 12.0 * M;
    const double r_in = 4.5 * M;
    double r_H = M + std::sqrt(M * M - a * a - Q * Q);
    double K = 0.001 * L.pow(M, Gamma - 1.0);
    double rho_tgt = 1.0;
    double l0;
    if (!ic_L0(L, t.ks, r_max, M, l0)) return 1;
    const double W_in = ic_W_in(L, ks_ic(L, t.ks, r_in, M_PI / 2.0), l0);
    const double lnh_max = ic_lnh(L, ks_ic(L, t.ks, r_max, M_PI / 2.0), l0, W_in);
    if (lnh_max <= 0.0) return 2;
    EOS st = t.eos();
    auto skip = [&](double r, double th) {
        return !std::isfinite(r) || !std::isfinite(th) || r <= 1.05 * r_H || std::abs(L.sin(th)) < 0.01;
    };
    double rho_max = 0.0;
    for (int b = 0; b < t.nb; b++)
        for (int i = 0; i < BLK; i++)
            for (int j = 0; j < BLK; j++)
                for (int k = 0; k < BLK; k++) {
                    const Geo& g = t.gc(b, lindx(i, j, k));
                    if (skip(g.r, g.th)) continue;
                    const KsIc mc = ks_ic(L, t.ks, g.r, g.th);
                    const double lnH = ic_lnh(L, mc, l0, W_in);
                    if (lnH <= 0.0 || !std::isfinite(lnH)) continue;
                    const IcTherm thm = ic_therm(L, lnH, Gamma, K);
                    if (std::isfinite(thm.rho)) rho_max = smax(rho_max, thm.rho);
                }
    if (rho_max < 1e-14) return 2;
    double rho_scal = rho_tgt / rho_max;
    for (int b = 0; b < t.nb; b++)
        for (int i = 0; i < BLK; i++)
            for (int j = 0; j < BLK; j++)
                for (int k = 0; k < BLK; k++) {
                    const int c = lindx(i, j, k);
                    const Geo& g = t.gc(b, c);
                    Prim W_p = pvfs(st, g.r);
                    if (skip(g.r, g.th)) { t.setW(b, c, W_p); continue; }
                    const KsIc mc = ks_ic(L, t.ks, g.r, g.th);
                    const double lnH = ic_lnh(L, mc, l0, W_in);
                    if (lnH <= 0.0 || !std::isfinite(lnH)) { t.setW(b, c, W_p); continue; }
                    const IcTherm thm = ic_therm(L, lnH, Gamma, K);
                    W_p.rho = rho_scal * thm.rho;
                    W_p.p = K * L.pow(W_p.rho, Gamma);
                    W_p.eps = W_p.p / ((Gamma - 1.0) * W_p.rho);
                    ic_vel(mc, ic_lnrf(mc), l0, g.xc, g.yc, W_p.v);
                    t.setW(b, c, W_p);
                }
    return 0;
}
// A_vpot, initconds.cpp:17-23
void ic_A(const IcLibm& L, double x, double y, double z, double rho, double rho_max, double A[3]) {
    double Ap = smax(rho / rho_max - 0.2, 0.0);
    double phi = L.atan2(y, x);
    double theta = L.acos(z / std::sqrt(x * x + y * y + z * z));
    A[0] = -Ap * L.sin(phi) * L.sin(theta);
    A[1] = Ap * L.cos(phi) * L.sin(theta);
    A[2] = 0.0;
}
// B_pot_init, initconds.cpp:594-668 (edge densities 25-39, edge positions
// 41-52; quirk: Bfz takes A_y on x-edges and A_x on y-edges), then f2cB
void ic_B_pot(Tree& t, const IcLibm& L, int p, double glmx_rho) {
    const Block& B = t.blk[p];
    const double dx = B.dx(), dy = B.dy(), dz = B.dz();
    double rho_max = glmx_rho;
    if (rho_max <= 1e-10) return;
    for (int i = 0; i < BLK; i++)
        for (int j = 0; j < BLK; j++)
            for (int k = 0; k < BLK; k++) rho_max = smax(rho_max, t.w(p, 0, lindx(i, j, k)));
    const int top = BLK + GH - 1;
    auto rho = [&](int i, int j, int k) { return t.w(p, 0, lindx(i, j, k)); };
    auto erx = [&](int i, int j, int k) {
        int j1 = std::min(j + 1, top), k1 = std::min(k + 1, top);
        return (rho(i, j, k) + rho(i, j1, k) + rho(i, j, k1) + rho(i, j1, k1)) / 4;
    };
    auto ery = [&](int i, int j, int k) {
        int i1 = std::min(i + 1, top), k1 = std::min(k + 1, top);
        return (rho(i, j, k) + rho(i1, j, k) + rho(i, j, k1) + rho(i1, j, k1)) / 4;
    };
    auto erz = [&](int i, int j, int k) {
        int i1 = std::min(i + 1, top), j1 = std::min(j + 1, top);
        return (rho(i, j, k) + rho(i1, j, k) + rho(i, j1, k) + rho(i1, j1, k)) / 4;
    };
    // A at an x-, y- or z-edge of cell (i, j, k), component comp
    auto Ax = [&](int comp, int i, int j, int k) {
        const Geo& g = t.gc(p, lindx(i, j, k));
        double A[3];
        ic_A(L, g.xc, g.yc + dy / 2.0, g.zc + dz / 2.0, erx(i, j, k), rho_max, A);
        return A[comp];
    };
    auto Ay = [&](int comp, int i, int j, int k) {
        const Geo& g = t.gc(p, lindx(i, j, k));
        double A[3];
        ic_A(L, g.xc + dx / 2.0, g.yc, g.zc + dz / 2.0, ery(i, j, k), rho_max, A);
        return A[comp];
    };
    auto Az = [&](int comp, int i, int j, int k) {
        const Geo& g = t.gc(p, lindx(i, j, k));
        double A[3];
        ic_A(L, g.xc + dx / 2.0, g.yc + dy / 2.0, g.zc, erz(i, j, k), rho_max, A);
        return A[comp];
    };
    double *Bx = t.bfx(p), *By = t.bfy(p), *Bz = t.bfz(p);
    for (int i = -GH + 1; i < BLK + GH; i++)
        for (int j = -GH + 1; j < BLK + GH; j++)
            for (int k = -GH + 1; k < BLK + GH; k++) {
                Bx[ixf(i, j, k)] = (Az(2, i, j, k) - Az(2, i, j - 1, k)) / dy - (Ay(1, i, j, k) - Ay(1, i, j, k - 1)) / dz;
                By[iyf(i, j, k)] = (Ax(0, i, j, k) - Ax(0, i, j, k - 1)) / dz - (Az(2, i, j, k) - Az(2, i - 1, j, k)) / dx;
                Bz[izf(i, j, k)] = (Ax(1, i, j, k) - Ax(1, i - 1, j, k)) / dx - (Ay(0, i, j, k) - Ay(0, i, j - 1, k)) / dy;
            }
    f2cB(t, p);
}
// main.cpp:94-159 after amrtree's constructor (grid.cpp:177-206).  Returns
// 0, 1 (L0_comp throws; nothing after the constructor), 2 (no torus placed),
// 3 (global rho_max below 1e-14: main returns).
int init_torus(Tree& t, int mode, double& rho_max) {
    const IcLibm L = ic_libm_for(mode);
    EOS st = t.eos();
    std::fill(t.U.begin(), t.U.end(), 0.0);
    std::fill(t.W.begin(), t.W.end(), 0.0);
    std::fill(t.Bf.begin(), t.Bf.end(), 0.0);
    for (int b = 0; b < t.nb; b++)
        for (int c = 0; c < NCELL; c++) {
            const Geo& g = t.gc(b, c);
            Prim w = pvfs(st, g.r);
            t.setW(b, c, w);
            t.setU(b, c, ptoc(t.ks, st, w, g.r, g.th));
        }
    rho_max = 0.0;
    const int fm = ic_fm_init(t, L);
    if (fm == 1) return 1;
    for (int b = 0; b < t.nb; b++)
        for (int i = 0; i < BLK; i++)
            for (int j = 0; j < BLK; j++)
                for (int k = 0; k < BLK; k++) rho_max = std::max(rho_max, t.w(b, 0, lindx(i, j, k)));
    if (rho_max < 1e-14) return 3;
    for (int b = 0; b < t.nb; b++) ic_B_pot(t, L, b, rho_max);
    for (int b = 0; b < t.nb; b++)
        for (int i = 0; i < BLK; i++)
            for (int j = 0; j < BLK; j++)
                for (int k = 0; k < BLK; k++) {
                    const int c = lindx(i, j, k);
                    const Geo& g = t.gc(b, c);
                    Prim w = t.getW(b, c);
                    t.setU(b, c, ptoc(t.ks, st, w, g.r, g.th));
                }
    return fm;
}

void put_prim(const Prim& w, double* o) {
    o[0] = w.rho; o[1] = w.eps;
    for (int d = 0; d < 3; d++) { o[2 + d] = w.v[d]; o[5 + d] = w.B[d]; }
}
void get_prim(const double* o, Prim& w) {
    std::memset(&w, 0, sizeof w);
    w.rho = o[0]; w.eps = o[1];
    for (int d = 0; d < 3; d++) { w.v[d] = o[2 + d]; w.B[d] = o[5 + d]; }
}
void put_cons(const Cons& u, double* o) {
    o[0] = u.D;
    for (int d = 0; d < 3; d++) o[1 + d] = u.S[d];
    o[4] = u.tau;
    for (int d = 0; d < 3; d++) o[5 + d] = u.B[d];
}
void get_cons(const double* o, Cons& u) {
    u.D = o[0];
    for (int d = 0; d < 3; d++) u.S[d] = o[1 + d];
    u.tau = o[4];
    for (int d = 0; d < 3; d++) u.B[d] = o[5 + d];
}

long long take_faults(Tree& t) { return t.faults.exchange(0); }

}  // namespace

extern "C" {

void* or_tree_new(int nb, const int* meta, const double* edges, double M, double a, double Q, double gm,
                  int libm_mode) {
    Tree* t = new Tree;
    t->ks = {M, a, Q};
    t->cell_ks = {1.0, 0.0, 0.0};
    t->gamma = gm;
    t->lm = libm_for(libm_mode);
    t->nb = nb;
    t->blk.resize(nb);
    for (int b = 0; b < nb; b++) {
        const int* m = meta + 14 * b;
        Block& B = t->blk[b];
        B.lvl = m[0]; B.ijk[0] = m[1]; B.ijk[1] = m[2]; B.ijk[2] = m[3];
        B.leaf = m[4]; B.parent = m[5];
        for (int c = 0; c < 8; c++) B.children[c] = m[6 + c];
        for (int i = 0; i < 17; i++) {
            B.xe[i] = edges[51 * b + i]; B.ye[i] = edges[51 * b + 17 + i]; B.ze[i] = edges[51 * b + 34 + i];
        }
    }
    t->geo.resize((size_t)nb * NCELL);
    for (int b = 0; b < nb; b++) cell_init(*t, b);
    t->U.assign((size_t)nb * 8 * NCELL, 0.0);
    t->W.assign((size_t)nb * 8 * NCELL, 0.0);
    t->dU.assign((size_t)nb * 8 * NCELL, 0.0);
    t->Bf.assign((size_t)nb * 3 * NFACE, 0.0);
    t->F.assign((size_t)nb * 3 * NFACE * 8, 0.0);
    t->E.assign((size_t)nb * 3 * NEDGE, 0.0);
    return t;
}
void or_tree_free(void* h) { delete static_cast<Tree*>(h); }
int or_init_torus(void* h, int libm_mode, double* rho_max) {
    double rm = 0.0;
    int rc = init_torus(*static_cast<Tree*>(h), libm_mode, rm);
    if (rho_max) *rho_max = rm;
    return rc;
}
void or_set_state(void* h, const double* U, const double* W, const double* Bf, const double* dU) {
    Tree* t = static_cast<Tree*>(h);
    if (U) std::memcpy(t->U.data(), U, t->U.size() * sizeof(double));
    if (W) std::memcpy(t->W.data(), W, t->W.size() * sizeof(double));
    if (Bf) std::memcpy(t->Bf.data(), Bf, t->Bf.size() * sizeof(double));
    if (dU) std::memcpy(t->dU.data(), dU, t->dU.size() * sizeof(double));
}
void or_get_state(void* h, double* U, double* W, double* Bf, double* dU) {
    Tree* t = static_cast<Tree*>(h);
    if (U) std::memcpy(U, t->U.data(), t->U.size() * sizeof(double));
    if (W) std::memcpy(W, t->W.data(), t->W.size() * sizeof(double));
    if (Bf) std::memcpy(Bf, t->Bf.data(), t->Bf.size() * sizeof(double));
    if (dU) std::memcpy(dU, t->dU.data(), t->dU.size() * sizeof(double));
}
long long or_step(void* h, double cfl, double* dt) {
    Tree* t = static_cast<Tree*>(h);
    *dt = step(*t, cfl);
    return take_faults(*t);
}
long long or_dtcomp(void* h, double cfl, double* dt) {
    Tree* t = static_cast<Tree*>(h);
    *dt = dtcomp(*t, cfl);
    return take_faults(*t);
}
long long or_rkstg(void* h, double dt, int stage) {
    Tree* t = static_cast<Tree*>(h);
    rkstg(*t, dt, stage);
    return take_faults(*t);
}
long long or_floors(void* h) {
    Tree* t = static_cast<Tree*>(h);
    floors(*t);
    return take_faults(*t);
}
long long or_ghosts_all(void* h) {
    Tree* t = static_cast<Tree*>(h);
    ghosts_all(*t);
    return take_faults(*t);
}
long long or_c2p_leaves(void* h) {
    Tree* t = static_cast<Tree*>(h);
    c2p_leaves(*t);
    return take_faults(*t);
}
long long or_flux_ct_patch(void* h, int b, double dt, double* F, double* E) {
    Tree* t = static_cast<Tree*>(h);
    fluxcomp(*t, b);
    emfcomp(*t, b);
    if (F) std::memcpy(F, t->fx(b), sizeof(double) * 3 * NFACE * 8);
    if (E) std::memcpy(E, t->ex(b), sizeof(double) * 3 * NEDGE);
    Bfupdate(*t, b, dt);
    f2cB(*t, b);
    return take_faults(*t);
}
long long or_update_terms(void* h, int b, double* dF, double* S) {
    Tree* t = static_cast<Tree*>(h);
    for (int i = 0; i < BLK; i++)
        for (int j = 0; j < BLK; j++)
            for (int k = 0; k < BLK; k++) {
                int c = (i * BLK + j) * BLK + k;
                double a[8], s[8];
                put_cons(div_flux(*t, b, i, j, k), a);
                put_cons(source(*t, b, i, j, k), s);
                for (int v = 0; v < 8; v++) { dF[v * 4096 + c] = a[v]; S[v * 4096 + c] = s[v]; }
            }
    return take_faults(*t);
}
double or_maxdivB(void* h, int b) { return maxdivB(*static_cast<Tree*>(h), b); }

// analysis::analyser quantities that depend only on the evolved state,
// src/fluid/data/quantitatives.cpp: vol_int / surf_int (:46-79) over leaf
// active zones in quilt order, summed sequentially like the reference, of
//   mass (:82-86), energy (:87-91), L_pol (:92-96), Edot (:103-107),
//   Phi_B (:114-118), thermE (:129-133), Hscal (:199-208, rs = 10 M)
// with r_horizon = M + sqrt(M^2 - a^2 - Q^2) (:18-21), val_vr / val_Br
// (:23-36) and sg = sqrt(-g) of the cell metric (a = 0, quirk Q2).
// abs[q] = the sum of |term| for q's integral (the conditioning of the sum).
void or_diagnostics(void* h, double* out, double* abs) {
    Tree& t = *static_cast<Tree*>(h);
    const double M = t.ks.M, a = t.ks.a, Q = t.ks.Q;
    const double hrz = M + std::sqrt(M * M - a * a - Q * Q);
    const double rs = 10.0 * M;
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0}, sa[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = 0; b < t.nb; b++) {
        if (!t.blk[b].leaf) continue;
        const double dx = t.blk[b].dx(), dy = t.blk[b].dy(), dz = t.blk[b].dz();
        const double shell = std::sqrt(dx * dx + dy * dy + dz * dz);
        for (int i = 0; i < BLK; i++)
            for (int j = 0; j < BLK; j++)
                for (int k = 0; k < BLK; k++) {
                    const int c = lindx(i, j, k);
                    const Geo& g = t.gc(b, c);
                    const MC mc = metric_comp(t.cell_ks, g.r, g.th);
                    const double sg = mc.sqrtdetg;
                    const double D = t.u(b, 0, c), S0 = t.u(b, 1, c), S1 = t.u(b, 2, c), tau = t.u(b, 4, c);
                    const double rho = t.w(b, 0, c), eps = t.w(b, 1, c);
                    const double vr = (g.xc * t.w(b, 2, c) + g.yc * t.w(b, 3, c) + g.zc * t.w(b, 4, c)) / g.r;
                    const double Br = (g.xc * t.w(b, 5, c) + g.yc * t.w(b, 6, c) + g.zc * t.w(b, 7, c)) / g.r;
                    double term[8];
                    term[0] = D * g.vol;
                    term[1] = (tau + D) * g.vol;
                    term[2] = (g.xc * S1 - g.yc * S0) * g.vol;
                    term[5] = rho * eps * sg * g.vol;
                    const bool on_hrz = !(std::abs(g.r - hrz) > shell);
                    const bool on_rs = !(std::abs(g.r - rs) > shell);
                    term[3] = on_hrz ? -((tau + D) * (mc.alpha * vr - (mc.beta[0] * g.xc + mc.beta[1] * g.yc +
                                                                       mc.beta[2] * g.zc) / g.r)) * g.vol
                                     : 0.0;
                    term[4] = on_hrz ? std::abs(Br) * sg * g.vol : 0.0;
                    term[6] = on_rs ? std::abs(g.zc) * rho * sg * g.vol : 0.0;
                    term[7] = on_rs ? rho * sg * g.vol : 0.0;
                    const bool in[8] = {true, true, true, on_hrz, on_hrz, true, on_rs, on_rs};
                    for (int q = 0; q < 8; q++)
                        if (in[q]) { s[q] += term[q]; sa[q] += std::abs(term[q]); }
                }
    }
    out[0] = s[0]; out[1] = s[1]; out[2] = s[2]; out[3] = s[3];
    out[4] = 0.5 * s[4];
    out[5] = s[5];
    out[6] = (s[7] > 1e-20) ? (s[6] / s[7]) / rs : 0.0;
    for (int q = 0; q < 6; q++) abs[q] = sa[q];
    abs[4] = 0.5 * sa[4];
    abs[6] = sa[6];
    abs[7] = sa[7];
    abs[5] = sa[5];
}
void or_cell_geom(void* h, int b, double* out) {
    Tree* t = static_cast<Tree*>(h);
    for (int c = 0; c < NCELL; c++) {
        const Geo& g = t->gc(b, c);
        double* o = out + 8 * c;
        o[0] = g.r; o[1] = g.th; o[2] = g.phi; o[3] = g.xc; o[4] = g.yc; o[5] = g.zc; o[6] = g.vol; o[7] = g.pvol;
    }
}
void or_set_threads(int n) { omp_set_num_threads(n); }
// the limiter reconfp applies in the step: 0 minmod, 1 van Leer, 2 MC (recon.hpp:16-34)
void or_set_limiter(void* h, int lim) { static_cast<Tree*>(h)->limiter = lim; }

void or_metric(double M, double a, double Q, double r, double th, double* o) {
    MC mc = metric_comp(KS{M, a, Q}, r, th);
    int n = 0;
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) o[n++] = mc.g[i][j];
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) o[n++] = mc.gi[i][j];
    for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) for (int k = 0; k < 4; k++) o[n++] = mc.G[i][j][k];
    o[n++] = mc.sqrtdetg;
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) o[n++] = mc.gam[i][j];
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) o[n++] = mc.gami[i][j];
    o[n++] = mc.sqrtdetgam;
    o[n++] = mc.alpha;
    for (int i = 0; i < 3; i++) o[n++] = mc.beta[i];
}
int or_ptoc(double M, double a, double Q, double gm, const double* w, double r, double th, double* u, double* der) {
    std::atomic<long long> f{0};
    EOS st{gm, &f};
    Prim pv; get_prim(w, pv);
    Cons c = ptoc(KS{M, a, Q}, st, pv, r, th);
    put_cons(c, u);
    if (der) { der[0] = pv.lor; der[1] = pv.p; der[2] = pv.h; der[3] = pv.b2; }
    return f.load() ? 1 : 0;
}
int or_ctop(double M, double a, double Q, double gm, const double* u, double r, double th, double* w, double* der,
            int libm_mode) {
    (void)libm_mode;
    std::atomic<long long> f{0};
    EOS st{gm, &f};
    Cons c; get_cons(u, c);
    Prim pv; std::memset(&pv, 0, sizeof pv);
    bool ok = ctop(KS{M, a, Q}, st, c, r, th, pv);
    put_prim(pv, w);
    if (der) { der[0] = pv.p; der[1] = pv.T; der[2] = pv.h; der[3] = pv.lor; der[4] = pv.b2; }
    if (f.load()) return 2;
    return ok ? 0 : 1;
}
int or_hlld(double M, double a, double Q, double gm, const double* WL, const double* WR, const double* UL,
            const double* UR, double r, double th, int dim, double* F, double* ct) {
    std::atomic<long long> f{0};
    EOS st{gm, &f};
    Prim wl, wr; get_prim(WL, wl); get_prim(WR, wr);
    Cons ul, ur; get_cons(UL, ul); get_cons(UR, ur);
    Cons Fl = hlld(wl, wr, ul, ur, KS{M, a, Q}, r, th, st, dim, *ct);
    put_cons(Fl, F);
    return f.load() ? 2 : 0;
}
void or_reconfp(void* h, int b, int i, int j, int k, int dim, int lim, double* L, double* R) {
    Tree* t = static_cast<Tree*>(h);
    Prim l, r;
    reconfp(*t, b, i, j, k, dim, lim, l, r);
    put_prim(l, L); put_prim(r, R);
}
double or_log(double x, int mode) { return libm_for(mode).log(x); }
double or_exp(double x, int mode) { return libm_for(mode).exp(x); }
double or_ic_fn(int op, double x, double y, int mode) {
    const IcLibm L = ic_libm_for(mode);
    switch (op) {
        case 0: return L.log(x);
        case 1: return L.exp(x);
        case 2: return L.sin(x);
        case 3: return L.cos(x);
        case 4: return L.acos(x);
        case 5: return L.pow(x, y);
        default: return L.atan2(x, y);
    }
}
}
