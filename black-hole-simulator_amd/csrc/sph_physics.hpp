// Pointwise GRMHD physics of the spherical Kerr-Schild engine ("consistent"
// preset).
//
// The reference's fluid step (src/fluid) evolves the Valencia variables
// D, S_j, tau, B^i with a PLM + HLLD + constrained-transport + RK2 scheme
// (integrator.cpp:191-275); its arithmetic carries quirks that make the
// scheme blow up within a few steps (mixed densitisation, an a = 0 cell
// metric, y/z fluxes solved as x fluxes, a non-inverse con2prim, a source
// that does not match its conserved variables).  This preset keeps the
// reference's variables and structure and makes each piece the textbook
// one (Banyuls et al. 1997; Anton et al. 2006):
//   U  = sqrt(gamma) (D, S_j, tau, B^i) with D = rho W,
//        S_j = (rho h W^2 + B^2) v_j - (B.v) B_j,
//        tau = rho h W^2 + B^2 - p - (B^2/W^2 + (B.v)^2)/2 - D     (conserved.cpp:17-73)
//   F^n = sqrt(gamma) (D V, S_j V + alpha (p* delta - b_j B^n / W),
//         tau V + alpha (p* v^n - alpha b^0 B^n / W), B^k V - B^n V^k),
//         V = alpha v^n - beta^n                                    (hlld.cpp:94-119)
//   S(S_j) = sqrt(-g) T^mn d_j g_mn / 2,
//   S(tau) = sqrt(-g) (T^m0 d_m alpha - alpha T^mn Gamma^0_mn)       (integrator.cpp:41-153)
// con2prim is a one-dimensional Newton iteration on Z = rho h W^2 as the
// reference's (conserved.cpp:80-230) with the correct residual and
// derivative, started from U alone so that the result is a function of U.
// The Riemann solver is HLL with the reference's signal-speed structure
// (alpha (v +- c) - beta, hlld.cpp:240-241) written for a general spatial
// metric; see the build's design notes for why the reference's HLLD
// intermediate states are not used here.
//
// Everything is a template on the arithmetic type R so that the same code
// gives the FP64 step and the FP32 reconstruction / Riemann / EMF path of
// the mixed-precision step.  No transcendental function is called: the
// metric comes from tables, and only sqrt and division are used.
#pragma once

#include "sph_common.hpp"

namespace sph {

template <class R>
struct Met {
    R alpha, beta[3], g[6], gi[6], sg;
};

// The Kerr-Schild (r, theta, phi) chart has a block-diagonal spatial metric:
// theta decouples, so gamma_12 = gamma_23 = 0, gamma^12 = gamma^23 = 0, and
// the shift is radial, beta^theta = beta^phi = 0.  The tables hold those words
// (exact zeros, checked where the tables are built); the physics below
// neither loads nor multiplies them.

// metric words at point p of a table with npt points per word
template <class R, class T>
SPH_HD Met<R> load_met(const T* tab, long long npt, int p) {
    Met<R> m;
    m.alpha = (R)tab[M_ALPHA * npt + p];
    m.beta[0] = (R)tab[M_B1 * npt + p];
    m.beta[1] = m.beta[2] = R(0);
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const bool zero = q == 1 || q == 4;
        m.g[q] = zero ? R(0) : (R)tab[(M_G11 + q) * npt + p];
        m.gi[q] = zero ? R(0) : (R)tab[(M_I11 + q) * npt + p];
    }
    m.sg = (R)tab[M_SG * npt + p];
    return m;
}

// lower an index with a symmetric 3x3 stored as 11 12 13 22 23 33 whose 12
// and 23 words are zero
template <class R>
SPH_HD void lower3(const R s[6], const R x[3], R y[3]) {
    y[0] = s[0] * x[0] + s[2] * x[2];
    y[1] = s[3] * x[1];
    y[2] = s[2] * x[0] + s[5] * x[2];
}
template <class R>
SPH_HD R dot3(const R a[3], const R b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

template <class R>
SPH_HD R minmod(R a, R b) {
    if (a * b <= R(0)) return R(0);
    return (a < R(0) ? -a : a) < (b < R(0) ? -b : b) ? a : b;
}
template <class R>
SPH_HD R rmax(R a, R b) { return a > b ? a : b; }
template <class R>
SPH_HD R rmin(R a, R b) { return a < b ? a : b; }
template <class R>
SPH_HD R rabs(R a) { return a < R(0) ? -a : a; }

// A primitive state at a point with everything the flux, the conserved
// variables and the signal speeds need.
template <class R>
struct Pt {
    R rho, p, u[3], B[3];   // inputs (u = W v^i)
    R W, iW, v[3], vl[3], Bl[3];
    R Bv, B2, b2, rhoh, ps;  // B.v, B^2, b^2, rho h, p + b^2/2
};

template <class R>
SPH_HD void pt_fill(const Met<R>& m, R gam, Pt<R>& s) {
    R ul[3];
    lower3(m.g, s.u, ul);
    const R uu = dot3(ul, s.u);
    s.W = sqrt(R(1) + uu);
    const R iW = R(1) / s.W;
    s.iW = iW;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        s.v[d] = s.u[d] * iW;
        s.vl[d] = ul[d] * iW;
    }
    lower3(m.g, s.B, s.Bl);
    s.Bv = dot3(s.Bl, s.v);
    s.B2 = dot3(s.Bl, s.B);
    s.b2 = s.B2 * iW * iW + s.Bv * s.Bv;
    s.rhoh = s.rho + gam / (gam - R(1)) * s.p;
    s.ps = s.p + R(0.5) * s.b2;
}

// conserved variables, undensitised: D, S_1..3, tau (the field is B itself)
template <class R>
SPH_HD void pt_cons(const Pt<R>& s, R c[5]) {
    const R Z = s.rhoh * s.W * s.W;
    c[0] = s.rho * s.W;
#pragma unroll
    for (int d = 0; d < 3; d++) c[1 + d] = (Z + s.B2) * s.vl[d] - s.Bv * s.Bl[d];
    c[4] = Z + s.B2 - s.p - R(0.5) * (s.B2 * s.iW * s.iW + s.Bv * s.Bv) - c[0];
}

// Flux in direction n, undensitised, 8 words (D, S_1..3, tau, B^1..3); the
// B^n word is zero.  c = pt_cons(s).
template <class R>
SPH_HD void pt_flux(const Met<R>& m, const Pt<R>& s, const R c[5], int n, R f[8]) {
    R V[3];  // alpha v^d - beta^d (radial shift)
    V[0] = m.alpha * s.v[0] - m.beta[0];
    V[1] = m.alpha * s.v[1];
    V[2] = m.alpha * s.v[2];
    const R Vn = V[n];
    const R ab0 = s.W * s.Bv;  // alpha b^0
    const R Bn = s.B[n];
    const R aBn = m.alpha * Bn * s.iW;
    f[0] = c[0] * Vn;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        const R bj = s.Bl[j] * s.iW + ab0 * s.vl[j];  // b_j
        f[1 + j] = c[1 + j] * Vn - bj * aBn;
    }
    f[1 + n] += m.alpha * s.ps;
    f[4] = c[4] * Vn + m.alpha * s.ps * s.v[n] - ab0 * aBn;
#pragma unroll
    for (int k = 0; k < 3; k++) f[5 + k] = s.B[k] * Vn - Bn * V[k];
    f[5 + n] = R(0);
}

// Fast magnetosonic signal speeds in direction n (coordinate speeds dx^n/dt):
// the sound-speed formula of Banyuls et al. 1997 with c_s replaced by the
// fast speed (hlld.cpp:34-39), c_f^2 = c_s^2 + v_A^2 - c_s^2 v_A^2, which
// for c_s^2 = gam p / (rho h) and v_A^2 = b^2 / (rho h + b^2) is
// (gam p + b^2) / (rho h + b^2).
template <class R>
SPH_HD void pt_speeds(const Met<R>& m, const Pt<R>& s, R gam, int n, R& lm, R& lp) {
    R cf2 = (gam * s.p + s.b2) / (s.rhoh + s.b2);
    cf2 = rmin(cf2, R(1));
    const R v2 = R(1) - s.iW * s.iW;
    const R ginn = n == 0 ? m.gi[0] : n == 1 ? m.gi[3] : m.gi[5];
    const R vn = s.v[n];
    const R den = R(1) - v2 * cf2;
    R disc = (R(1) - v2) * (ginn * den - vn * vn * (R(1) - cf2));
    disc = rmax(disc, R(0));
    const R sq = sqrt(cf2 * disc);
    const R a = vn * (R(1) - cf2);
    const R ad = m.alpha / den;
    if (n == 0) {
        lp = (a + sq) * ad - m.beta[0];
        lm = (a - sq) * ad - m.beta[0];
    } else {
        lp = (a + sq) * ad;
        lm = (a - sq) * ad;
    }
}

// One side of a face: the seven conserved words the HLL flux mixes (D,
// S_1..3, tau and the two transverse fields in cyclic order after n), their
// fluxes and the signal speeds.
template <class R>
struct Side {
    R u[7], f[7], lm, lp;
};
template <class R>
SPH_HD void face_side(const Met<R>& m, R gam, Pt<R>& s, int n, Side<R>& o) {
    pt_fill(m, gam, s);
    R c[5], f[8];
    pt_cons(s, c);
    pt_flux(m, s, c, n, f);
    pt_speeds(m, s, gam, n, o.lm, o.lp);
    const int t1 = n == 2 ? 0 : n + 1, t2 = n == 0 ? 2 : n - 1;
#pragma unroll
    for (int q = 0; q < 5; q++) {
        o.u[q] = c[q];
        o.f[q] = f[q];
    }
    o.u[5] = s.B[t1]; o.f[5] = f[5 + t1];
    o.u[6] = s.B[t2]; o.f[6] = f[5 + t2];
}

// HLL flux between face sides L and R, densitised by the face's
// sqrt(gamma); out[NFW] in the order of Side.
template <class R>
SPH_HD void hll(const Met<R>& m, const Side<R>& L, const Side<R>& Rt, R out[NFW]) {
    const R sl = rmin(rmin(L.lm, Rt.lm), R(0));
    const R sr = rmax(rmax(L.lp, Rt.lp), R(0));
    const R den = sr - sl;
    const R id = den > R(0) ? m.sg / den : R(0);
    const R slsr = sl * sr;
#pragma unroll
    for (int q = 0; q < 7; q++) out[q] = (sr * L.f[q] - sl * Rt.f[q] + slsr * (Rt.u[q] - L.u[q])) * id;
}

// Limited linear (minmod) reconstruction at the face between cells c-1 and
// c from the four cells c-2 .. c+1 along the direction (recon.hpp:16-19,
// reconstruction.cpp:28-38): left state from cell c-1, right from cell c.
template <class R>
SPH_HD void plm(R a, R b, R c, R d, R& l, R& r) {
    l = b + R(0.5) * minmod(b - a, c - b);
    r = c - R(0.5) * minmod(c - b, d - c);
}

// --------------------------------------------------------------- FP64 only

// prim -> cons, densitised; w = rho, eps, u~^1..3, B^1..3
SPH_HD void ptoc(const Met<double>& m, double gam, const double w[8], double U[8]) {
    Pt<double> s;
    s.rho = w[WRHO];
    s.p = (gam - 1.0) * w[WRHO] * w[WEPS];
    for (int d = 0; d < 3; d++) { s.u[d] = w[WU1 + d]; s.B[d] = w[WB1 + d]; }
    pt_fill(m, gam, s);
    double c[5];
    pt_cons(s, c);
    for (int q = 0; q < 5; q++) U[q] = m.sg * c[q];
    for (int d = 0; d < 3; d++) U[UB1 + d] = m.sg * s.B[d];
}

// con2prim: Newton on Z = rho h W^2 for the energy equation
//   f(Z) = Z + B^2 - p(Z) - (B^2 (1 - v^2) + (S.B)^2 / Z^2) / 2 - (tau + D)
// with v^2(Z) = (Z^2 S^2 + (S.B)^2 (2Z + B^2)) / (Z^2 (Z + B^2)^2) and
// p(Z) = (gam-1)/gam (Z (1 - v^2) - D sqrt(1 - v^2)); then bisection on a
// bracket if Newton leaves it.  Input U densitised, the field in U[5..7];
// output w (u~ and B filled).  Returns
//   0  solved (eps = 0 when the recovered pressure is not positive: the
//      floors then set it),
//   1  no root (the energy is too small for the momentum): the cold
//      inversion, u~^i = gamma^ij S_j / D, rho = D / W, eps = 0, which keeps
//      D and S (the reference floors the whole state instead, integrator.cpp:211-218),
//   2  unrecoverable (D <= 0 or non-finite input): the caller applies the
//      floor state.
constexpr int C2P_NEWTON = 60;
constexpr int C2P_BISECT = 200;

struct C2PIn {
    double D, E, S2, SB, B2;
};
SPH_HD double c2p_v2(const C2PIn& q, double Z) {
    const double ZB = Z + q.B2;
    return (Z * Z * q.S2 + q.SB * q.SB * (2.0 * Z + q.B2)) / (Z * Z * ZB * ZB);
}
SPH_HD double c2p_f(const C2PIn& q, double gam, double Z, double& df) {
    const double ZB = Z + q.B2;
    const double v2 = c2p_v2(q, Z);
    const double omv = 1.0 - v2;
    const double sq = sqrt(omv);
    const double gg = (gam - 1.0) / gam;
    const double p = gg * (Z * omv - q.D * sq);
    const double Z3 = Z * Z * Z;
    const double dv2 = -2.0 * (q.S2 / (ZB * ZB * ZB) + q.SB * q.SB * (3.0 * Z * Z + 3.0 * q.B2 * Z + q.B2 * q.B2) /
                                                          (Z3 * ZB * ZB * ZB));
    const double dp = gg * (omv - Z * dv2 + q.D * dv2 / (2.0 * sq));
    df = 1.0 - dp + 0.5 * q.B2 * dv2 + q.SB * q.SB / Z3;
    return Z + q.B2 - p - 0.5 * (q.B2 * omv + q.SB * q.SB / (Z * Z)) - q.E;
}

SPH_HD int c2p(const Met<double>& m, double gam, const double U[8], double w[8]) {
    const double is = 1.0 / m.sg;
    C2PIn q;
    double S[3], B[3], Su[3], Bl[3];
    q.D = U[UD] * is;
    for (int d = 0; d < 3; d++) { S[d] = U[US1 + d] * is; B[d] = U[UB1 + d] * is; }
    q.E = U[UTAU] * is + q.D;
    lower3(m.gi, S, Su);
    lower3(m.g, B, Bl);
    q.S2 = dot3(S, Su);
    q.SB = dot3(S, B);
    q.B2 = dot3(Bl, B);
    for (int d = 0; d < 3; d++) w[WB1 + d] = B[d];
    if (!(q.D > 0.0) || !isfinite(q.S2 + q.SB + q.B2 + q.E)) return 2;
    int rc = 0;
    double Z = 0.0;
    // bracket: v^2(Z) falls with Z and v^2 < 1 needs Z above its crossing
    // of 1 (bisection for the crossing when v^2(D) >= 1); f grows with Z
    double lo = q.D, hi = 2.0 * (rmax(q.E, 0.0) + q.B2) + q.D;
    for (int it = 0; it < 64 && c2p_v2(q, hi) >= 1.0; it++) hi = 2.0 * hi;
    if (c2p_v2(q, lo) >= 1.0) {
        double a = lo, b = hi;
        for (int it = 0; it < 200; it++) {
            const double mid = 0.5 * (a + b);
            if (!(mid > a) || !(mid < b)) break;
            if (c2p_v2(q, mid) >= 1.0) a = mid; else b = mid;
        }
        lo = b;
    }
    if (!(q.E > 0.0) || !(c2p_v2(q, lo) < 1.0) || !(lo < hi)) {
        rc = 1;
    } else {
        Z = q.D + gam * rmax(q.E - 0.5 * q.B2 - q.D, 0.0);
        if (!(Z > lo)) Z = lo;
        if (!(Z < hi)) Z = 0.5 * (lo + hi);
        bool ok = false;
        for (int it = 0; it < C2P_NEWTON; it++) {
            double df;
            const double f = c2p_f(q, gam, Z, df);
            if (!isfinite(f) || !(df > 0.0)) break;
            const double Zn = Z - f / df;
            if (!(Zn > lo) || !(Zn < hi)) break;
            const double dz = Zn - Z;
            Z = Zn;
            if (rabs(dz) <= 1e-14 * Z) { ok = true; break; }
        }
        if (!ok) {
            double d;
            const double flo = c2p_f(q, gam, lo, d);
            double fhi = c2p_f(q, gam, hi, d);
            for (int it = 0; it < 64 && fhi < 0.0; it++) { hi = 2.0 * hi; fhi = c2p_f(q, gam, hi, d); }
            if (!(flo <= 0.0) || !(fhi >= 0.0)) {
                rc = 1;
            } else {
                for (int it = 0; it < C2P_BISECT; it++) {
                    const double mid = 0.5 * (lo + hi);
                    if (!(mid > lo) || !(mid < hi)) break;
                    const double fm = c2p_f(q, gam, mid, d);
                    if (fm < 0.0) lo = mid; else hi = mid;
                }
                Z = 0.5 * (lo + hi);
            }
        }
    }
    if (rc == 0) {
        const double v2 = c2p_v2(q, Z);
        if (!(v2 < 1.0)) {
            rc = 1;
        } else {
            const double W = 1.0 / sqrt(1.0 - v2);
            const double rho = q.D / W;
            const double p = (gam - 1.0) / gam * (Z / (W * W) - rho);
            if (!(rho > 0.0) || !isfinite(p)) {
                rc = 1;
            } else {
                w[WRHO] = rho;
                w[WEPS] = p > 0.0 ? p / ((gam - 1.0) * rho) : 0.0;
                double vl[3], vu[3];
                const double iZB = 1.0 / (Z + q.B2);
                for (int d = 0; d < 3; d++) vl[d] = (S[d] + q.SB * Bl[d] / Z) * iZB;
                lower3(m.gi, vl, vu);
                for (int d = 0; d < 3; d++) w[WU1 + d] = W * vu[d];
                return 0;
            }
        }
    }
    // cold inversion: S_j = D u~_j
    const double iD = 1.0 / q.D;
    for (int d = 0; d < 3; d++) w[WU1 + d] = Su[d] * iD;
    const double W = sqrt(1.0 + q.S2 * iD * iD);
    w[WRHO] = q.D / W;
    w[WEPS] = 0.0;
    return 1;
}

// Floors and ceilings on a primitive state at a cell (after con2prim):
// rho >= rho_fl, eps >= eps_fl, Lorentz factor <= w_max, b^2/rho <= sigma_max
// (grid.cpp:331-358, grid.hpp:140-143).  A density or energy within a
// relative 1e-10 below its floor counts as at the floor (a zone left at the
// floor comes back from con2prim a rounding error away from it).  Returns 1
// if w changed.
SPH_HD int apply_floors(const Met<double>& m, double gam, double rfl, double efl, double sigma_max, double w_max,
                        double w[8]) {
    int ch = 0;
    double ul[3];
    lower3(m.g, w + WU1, ul);
    const double uu = dot3(ul, w + WU1);
    const double wm2 = w_max * w_max - 1.0;
    if (uu > wm2) {
        const double s = sqrt(wm2 / uu);
        for (int d = 0; d < 3; d++) w[WU1 + d] *= s;
        ch = 1;
    }
    if (w[WRHO] < rfl * (1.0 - 1e-10)) { w[WRHO] = rfl; ch = 1; }
    if (w[WEPS] < efl * (1.0 - 1e-10)) { w[WEPS] = efl; ch = 1; }
    // magnetisation ceiling
    double Bl[3];
    lower3(m.g, w + WB1, Bl);
    lower3(m.g, w + WU1, ul);
    const double W2 = 1.0 + dot3(ul, w + WU1);
    const double Bu = dot3(Bl, w + WU1);  // W (B.v)
    const double b2 = (dot3(Bl, w + WB1) + Bu * Bu) / W2;
    if (b2 > sigma_max * w[WRHO]) { w[WRHO] = b2 / sigma_max; ch = 1; }
    return ch;
}

// The floor state of a cell whose con2prim failed: rho_fl, eps_fl, at rest
// with respect to the normal observer, the cell's field kept.
SPH_HD void floor_state(double rfl, double efl, double w[8]) {
    w[WRHO] = rfl;
    w[WEPS] = efl;
    w[WU1] = w[WU2] = w[WU3] = 0.0;
}

// Geometric source terms at a cell centre, densitised by sqrt(-g); out[5]
// for D, S_1..3, tau.  gc = the cell-centre table, npt its word stride, p the
// point.
template <class Tb>
SPH_HD void source(const Tb* gc, long long npt, int p, const Met<double>& m, double gam, const double w[8],
                   double out[5]) {
    Pt<double> s;
    s.rho = w[WRHO];
    s.p = (gam - 1.0) * w[WRHO] * w[WEPS];
#pragma unroll
    for (int d = 0; d < 3; d++) { s.u[d] = w[WU1 + d]; s.B[d] = w[WB1 + d]; }
    pt_fill(m, gam, s);
    const double ia = 1.0 / m.alpha;
    // four-velocity and magnetic four-vector, contravariant (radial shift)
    double u[4], b[4];
    u[0] = s.W * ia;
    u[1] = s.W * (s.v[0] - m.beta[0] * ia);
    u[2] = s.W * s.v[1];
    u[3] = s.W * s.v[2];
    b[0] = s.W * s.Bv * ia;
#pragma unroll
    for (int d = 0; d < 3; d++) b[1 + d] = (s.B[d] + m.alpha * b[0] * u[1 + d]) / s.W;
    // inverse four-metric g^00 = -1/alpha^2, g^0i = beta^i/alpha^2,
    // g^ij = gamma^ij - beta^i beta^j/alpha^2: the pairs 02, 03, 12, 23 vanish
    const double ia2 = ia * ia;
    const double g00 = -ia2, g01 = m.beta[0] * ia2, g11 = m.gi[0] - m.beta[0] * m.beta[0] * ia2;
    const double g13 = m.gi[2], g22 = m.gi[3], g33 = m.gi[5];
    const double wt = s.rhoh + s.b2, ps = s.ps;
    // T^mn over the 10 independent pairs 00 01 02 03 11 12 13 22 23 33
    double Tm[10];
    Tm[0] = wt * u[0] * u[0] + ps * g00 - b[0] * b[0];
    Tm[1] = wt * u[0] * u[1] + ps * g01 - b[0] * b[1];
    Tm[2] = wt * u[0] * u[2] - b[0] * b[2];
    Tm[3] = wt * u[0] * u[3] - b[0] * b[3];
    Tm[4] = wt * u[1] * u[1] + ps * g11 - b[1] * b[1];
    Tm[5] = wt * u[1] * u[2] - b[1] * b[2];
    Tm[6] = wt * u[1] * u[3] + ps * g13 - b[1] * b[3];
    Tm[7] = wt * u[2] * u[2] + ps * g22 - b[2] * b[2];
    Tm[8] = wt * u[2] * u[3] - b[2] * b[3];
    Tm[9] = wt * u[3] * u[3] + ps * g33 - b[3] * b[3];
    // contractions, off-diagonal pairs twice; d_j g_mn vanishes for the pairs
    // 02, 12, 23 (g_t theta = g_r theta = g_theta phi = 0)
    double s1 = 0.0, s2 = 0.0, sg0 = 0.0;
#pragma unroll
    for (int q = 0; q < 10; q++) {
        const double wq = (q == 0 || q == 4 || q == 7 || q == 9) ? 0.5 : 1.0;
        if (q != 2 && q != 5 && q != 8) {
            s1 += wq * Tm[q] * (double)gc[(C_DG1 + q) * npt + p];
            s2 += wq * Tm[q] * (double)gc[(C_DG2 + q) * npt + p];
        }
        sg0 += (2.0 * wq) * Tm[q] * (double)gc[(C_GAM0 + q) * npt + p];
    }
    const double sqg = m.alpha * m.sg;
    out[0] = 0.0;
    out[1] = sqg * s1;
    out[2] = sqg * s2;
    out[3] = 0.0;
    out[4] = sqg * (Tm[1] * (double)gc[C_DA1 * npt + p] + Tm[2] * (double)gc[C_DA2 * npt + p] - m.alpha * sg0);
}

// CFL time step of one cell: cfl * min over the directions (phi skipped for
// an axisymmetric run) of dx^d / max|lambda^d|.
SPH_HD double cell_dt(const Met<double>& m, double gam, const double w[8], const double dx[3], int ndir) {
    Pt<double> s;
    s.rho = w[WRHO];
    s.p = (gam - 1.0) * w[WRHO] * w[WEPS];
    for (int d = 0; d < 3; d++) { s.u[d] = w[WU1 + d]; s.B[d] = w[WB1 + d]; }
    pt_fill(m, gam, s);
    double dt = 1e300;
    // unrolled so that the direction index is a constant in each copy (a
    // runtime index into the point's arrays would place them in memory)
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (d >= ndir) break;
        double lm, lp;
        pt_speeds(m, s, gam, d, lm, lp);
        const double c = rmax(rabs(lm), rabs(lp));
        if (c > 0.0) dt = rmin(dt, dx[d] / c);
    }
    return dt;
}

// Per-zone terms of the analyser quantities (analysis::analyser,
// src/fluid/data/quantitatives.cpp:109-193) in the consistent preset, from
// the four-velocity u and the field four-vector b (as in source()):
//   out[0] rho eps              thermal energy density      (thermE)
//   out[1] b^2 / 2              magnetic energy density     (magE)
//   out[2] p                    gas pressure                (P_beta's numerator)
//   out[3] p + b^2 / 2          total pressure, rho >= rho_min   (alpha_SS)
//   out[4] rho h u^r u_phi      Reynolds stress, rho >= rho_min  (alpha_SS, maxreyn)
//   out[5] -b^r b_phi           Maxwell stress, rho >= rho_min   (alpha_SS, maxreyn)
// and, on the zone row at the horizon (horizon = true; else 0):
//   out[6] -rho u^r             rest-mass inflow            (Mdot)
//   out[7] -T^r_phi             angular-momentum inflow     (Ldot)
//   out[8] -(b^2 u^r u_t - b^r b_t)   electromagnetic energy outflow (L_BZ)
// The caller multiplies by sqrt(-g) dr dtheta dphi (volume terms) or
// sqrt(-g) dtheta dphi (horizon terms).  The reference's versions combine
// coordinate components of v and B instead (Ldot subtracts B^r and B^phi
// separately, quantitatives.cpp:109-113); these are the covariant ones.
constexpr int NANA = 9;
SPH_HD void analysis_terms(const Met<double>& m, double gam, const double w[8], double rho_min, bool horizon,
                           double out[NANA]) {
    Pt<double> s;
    s.rho = w[WRHO];
    s.p = (gam - 1.0) * w[WRHO] * w[WEPS];
#pragma unroll
    for (int d = 0; d < 3; d++) { s.u[d] = w[WU1 + d]; s.B[d] = w[WB1 + d]; }
    pt_fill(m, gam, s);
    const double ia = 1.0 / m.alpha;
    double u[4], b[4];
    u[0] = s.W * ia;
    u[1] = s.W * (s.v[0] - m.beta[0] * ia);
    u[2] = s.W * s.v[1];
    u[3] = s.W * s.v[2];
    b[0] = s.W * s.Bv * ia;
#pragma unroll
    for (int d = 0; d < 3; d++) b[1 + d] = (s.B[d] + m.alpha * b[0] * u[1 + d]) / s.W;
    // lowering with g_tt = -alpha^2 + beta_r beta^r, g_ti = beta_i = gamma_ij beta^j
    const double br_ = m.g[0] * m.beta[0], bph_ = m.g[2] * m.beta[0];
    const double gtt = -m.alpha * m.alpha + br_ * m.beta[0];
    auto low_t = [&](const double x[4]) { return gtt * x[0] + br_ * x[1] + bph_ * x[3]; };
    auto low_ph = [&](const double x[4]) { return bph_ * x[0] + m.g[2] * x[1] + m.g[5] * x[3]; };
    const double uph = low_ph(u), bph = low_ph(b), ut = low_t(u), bt = low_t(b);
    const bool dense = s.rho >= rho_min;
    out[0] = s.rho * w[WEPS];
    out[1] = 0.5 * s.b2;
    out[2] = s.p;
    out[3] = dense ? s.p + 0.5 * s.b2 : 0.0;
    out[4] = dense ? s.rhoh * u[1] * uph : 0.0;
    out[5] = dense ? -b[1] * bph : 0.0;
    out[6] = horizon ? -s.rho * u[1] : 0.0;
    out[7] = horizon ? -((s.rhoh + s.b2) * u[1] * uph - b[1] * bph) : 0.0;
    out[8] = horizon ? -(s.b2 * u[1] * ut - b[1] * bt) : 0.0;
}
// the quantities from the run's sums S[NANA + 1] of the terms times their
// measures (S[NANA] = Phi_B): thermE, magE, Ldot, L_BZ, eta_BZ, phi_BH,
// P_beta, alpha_SS, Maxwell/Reynolds, Mdot (the reference's guards against
// vanishing denominators, quantitatives.cpp:129-193)
SPH_HD void analysis_combine(const double S[NANA + 1], double out[10]) {
    const double mdot = S[6], amd = mdot < 0.0 ? -mdot : mdot;
    out[0] = S[0];
    out[1] = S[1];
    out[2] = S[7];
    out[3] = S[8];
    out[4] = amd > 1e-25 ? S[8] / amd : 0.0;
    out[5] = amd > 1e-30 ? S[NANA] / sqrt(amd) : 0.0;
    out[6] = S[1] > 1e-20 ? S[2] / S[1] : 1e30;
    out[7] = S[3] > 1e-25 ? (S[4] + S[5]) / S[3] : 0.0;
    out[8] = (S[4] < 0.0 ? -S[4] : S[4]) > 1e-30 ? S[5] / S[4] : 0.0;
    out[9] = mdot;
}

}  // namespace sph
