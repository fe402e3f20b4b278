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
// The Riemann solver is HLLD (hlld below): the reference's five-wave
// structure (hlld.cpp:230-312: fast waves, Alfven waves, contact, HLL
// fallback) with the relativistic intermediate states of Mignone, Ugliano &
// Bodo (2009), solved in the orthonormal frame of the normal observer at the
// face; HLL (hll), with the signal speeds alpha (v +- c) - beta of
// hlld.cpp:240-241 written for a general spatial metric, is its fallback and
// a selectable solver of its own.
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

// The con2prim energy residual in Z = rho h W^2 (its definition is with
// c2p below), for any arithmetic type: the FP64 con2prim and the HLLD
// solver's initial total pressure both use it.
template <class R>
struct C2PInT {
    R D, E, S2, SB, B2;
};
using C2PIn = C2PInT<double>;
template <class R>
SPH_HD R c2p_v2(const C2PInT<R>& q, R Z) {
    const R ZB = Z + q.B2;
    return (Z * Z * q.S2 + q.SB * q.SB * (R(2) * Z + q.B2)) / (Z * Z * ZB * ZB);
}
template <class R>
SPH_HD R c2p_f(const C2PInT<R>& q, R gam, R Z, R& df) {
    // 1/Z and 1/(Z + B^2) from one division
    const R ZB = Z + q.B2;
    const R r = R(1) / (Z * ZB);
    const R iZ = ZB * r, iZB = Z * r;
    const R iZ2 = iZ * iZ, iZB2 = iZB * iZB;
    const R SB2 = q.SB * q.SB;
    const R v2 = (q.S2 + SB2 * (R(2) * Z + q.B2) * iZ2) * iZB2;
    const R omv = R(1) - v2;
    const R sq = sqrt(omv);
    const R gg = (gam - R(1)) / gam;
    const R p = gg * (Z * omv - q.D * sq);
    const R iZ3 = iZ2 * iZ, iZB3 = iZB2 * iZB;
    const R dv2 = R(-2) * (q.S2 * iZB3 + SB2 * (R(3) * Z * Z + R(3) * q.B2 * Z + q.B2 * q.B2) * iZ3 * iZB3);
    const R dp = gg * (omv - Z * dv2 + q.D * dv2 / (R(2) * sq));
    df = R(1) - dp + R(0.5) * q.B2 * dv2 + SB2 * iZ3;
    return Z + q.B2 - p - R(0.5) * (q.B2 * omv + SB2 * iZ2) - q.E;
}

// ------------------------------------------------------------------- HLLD
// The five-wave relativistic HLLD solver of Mignone, Ugliano & Bodo (2009,
// MNRAS 393, 1141; "MUB09") in the orthonormal frame of the normal observer
// at the face, where special-relativistic formulas hold (Pons et al. 1998;
// White, Stone & Gammie 2016).  This is the wave structure of the
// reference's hlld (hlld.cpp:230-312: outer fast waves, Alfven waves around
// a contact, HLL fallback) with intermediate states that satisfy the
// relativistic jump conditions.
//
// Frame: x along the unit normal n_i = delta_i^n / sqrt(gamma^nn), y and z
// orthonormal in the face, cyclic after n (n = r: theta, phi; n = theta:
// phi, r_perp; n = phi: r, theta).  The Kerr-Schild chart's block-diagonal
// metric gives the triads in closed form:
//   n = r:     e_x = gamma^{k r} / sqrt(gamma^rr), e_y = e_theta / sqrt(g_thth),
//              e_z = e_phi / sqrt(g_phph)
//   n = theta: e_x = e_theta / sqrt(g_thth), e_y = e_phi / sqrt(g_phph),
//              e_z = gamma^{k r} / sqrt(gamma^rr)
//   n = phi:   e_x = gamma^{k phi} / sqrt(gamma^phph), e_y = e_r / sqrt(g_rr),
//              e_z = e_theta / sqrt(g_thth)
// Frame components: X^(a) = e_a^k X_k.  The coordinate surface x^n = const
// moves through the frame with normal speed v_s = beta^n / (alpha sqrt(gamma^nn)),
// and the coordinate flux is F^n = sqrt(gamma) alpha sqrt(gamma^nn) M (F^ - v_s U^)
// (M maps frame components back: S_j with the lowered triad, B^k with the
// triad itself), plus sqrt(gamma) beta^k B^n in the B^k flux (k != n), the
// part of B^k V^n - B^n V^k that the normal observer's frame does not see.

// the frame scalars of a face of direction n
template <class R>
struct Frame {
    R rn, irn;    // sqrt(gamma^nn) and its inverse
    R s1, is1;    // sqrt of the first transverse metric word and its inverse
    R s2, is2;    // the second (n = theta: sqrt(gamma^rr), the r_perp axis)
    R g13, gi13;  // gamma_r phi, gamma^r phi
    R vs;         // normal speed of the coordinate surface in the frame
    R c;          // sqrt(gamma) alpha sqrt(gamma^nn)
};
template <class R>
SPH_HD Frame<R> make_frame(const Met<R>& m, R gi13, int n) {
    Frame<R> f;
    f.rn = sqrt(n == 0 ? m.gi[0] : n == 1 ? m.gi[3] : m.gi[5]);
    f.irn = R(1) / f.rn;
    f.s1 = sqrt(n == 0 ? m.g[3] : n == 1 ? m.g[5] : m.g[0]);
    f.s2 = sqrt(n == 0 ? m.g[5] : n == 1 ? m.gi[0] : m.g[3]);
    f.is1 = R(1) / f.s1;
    f.is2 = R(1) / f.s2;
    f.g13 = m.g[2];
    f.gi13 = gi13;
    f.vs = n == 0 ? m.beta[0] * f.irn / m.alpha : R(0);
    f.c = m.sg * m.alpha * f.rn;
    return f;
}
// frame components (x, y, z) of a vector given contravariant X and lowered Xl
template <class R>
SPH_HD void to_frame(const Frame<R>& f, int n, const R X[3], const R Xl[3], R o[3]) {
    if (n == 0) {
        o[0] = X[0] * f.irn; o[1] = Xl[1] * f.is1; o[2] = Xl[2] * f.is2;
    } else if (n == 1) {
        o[0] = X[1] * f.irn; o[1] = Xl[2] * f.is1; o[2] = X[0] * f.is2;
    } else {
        o[0] = X[2] * f.irn; o[1] = Xl[0] * f.is1; o[2] = Xl[1] * f.is2;
    }
}

// One side of a face in the frame: velocity, field (x = the common normal
// field), the seven conserved words D, m_x, m_y, m_z, E (= tau + D), B_y,
// B_z, their fluxes, and the fast signal speeds (MUB09 sec. 2).
template <class R>
struct FSide {
    R v[3], B[3], iW2, vB, pt;
    R u[7], f[7], lm, lp;
};
template <class R>
SPH_HD void frame_side(const Met<R>& m, const Frame<R>& fr, R gam, Pt<R>& s, int n, R bx, FSide<R>& o) {
    pt_fill(m, gam, s);
    to_frame(fr, n, s.v, s.vl, o.v);
    R Bf[3];
    to_frame(fr, n, s.B, s.Bl, Bf);
    o.B[0] = bx;
    o.B[1] = Bf[1];
    o.B[2] = Bf[2];
    o.iW2 = s.iW * s.iW;
    o.vB = s.Bv;
    o.pt = s.ps;
    const R E = s.rhoh * s.W * s.W + s.B2 - s.ps;
    o.u[0] = s.rho * s.W;
#pragma unroll
    for (int a = 0; a < 3; a++) o.u[1 + a] = (E + s.ps) * o.v[a] - s.Bv * o.B[a];
    o.u[4] = E;
    o.u[5] = o.B[1];
    o.u[6] = o.B[2];
    const R vx = o.v[0];
    o.f[0] = o.u[0] * vx;
#pragma unroll
    for (int a = 0; a < 3; a++) o.f[1 + a] = o.u[1 + a] * vx - bx * (o.B[a] * o.iW2 + s.Bv * o.v[a]);
    o.f[1] += s.ps;
    o.f[4] = o.u[1];
    o.f[5] = o.B[1] * vx - bx * o.v[1];
    o.f[6] = o.B[2] * vx - bx * o.v[2];
    // fast speeds: the coordinate formula of pt_speeds with gamma^nn = 1,
    // alpha = 1, beta = 0
    R cf2 = (gam * s.p + s.b2) / (s.rhoh + s.b2);
    cf2 = rmin(cf2, R(1));
    const R v2 = R(1) - o.iW2;
    const R den = R(1) - v2 * cf2;
    R disc = (R(1) - v2) * (den - vx * vx * (R(1) - cf2));
    disc = rmax(disc, R(0));
    const R sq = sqrt(cf2 * disc);
    const R a = vx * (R(1) - cf2);
    const R id = R(1) / den;
    o.lp = (a + sq) * id;
    o.lm = (a - sq) * id;
}

// The state behind a fast wave of speed lam with total pressure p, from
// Rv = lam U - F(U) of the outer state (MUB09 eqs. 21-31): velocity, field,
// total enthalpy w = rho h + b^2, eta = sgn sqrt(w) and the vector K whose
// x component is the Alfven speed (MUB09 eq. 43).
template <class R>
struct AState {
    R v[3], B[3], vB, w, eta, K[3];
};
template <class R>
SPH_HD void fast_state(const R Rv[7], R lam, R bx, R p, R sgn, AState<R>& a) {
    const R Rmx = Rv[1], Rmy = Rv[2], Rmz = Rv[3], RE = Rv[4], RBy = Rv[5], RBz = Rv[6];
    const R l2 = R(1) - lam * lam;
    const R A = Rmx - lam * RE + p * l2;
    const R G = RBy * RBy + RBz * RBz;
    const R C = RBy * Rmy + RBz * Rmz;
    const R Q = -A - G + bx * bx * l2;
    const R X = bx * (A * lam * bx + C) - (A + G) * (lam * p + RE);
    // v = N / X; lambda - v_x = D2 / X
    const R Nx = bx * (A * bx + lam * C) - (A + G) * (p + Rmx);
    const R cc = C + bx * (lam * Rmx - RE);
    const R Ny = Q * Rmy + RBy * cc, Nz = Q * Rmz + RBz * cc;
    const R D2 = lam * X - Nx;
    R iX, iD2;
    if constexpr (sizeof(R) == 8) {
        // both reciprocals from one division (FP64's range holds X D2)
        const R t = R(1) / (X * D2);
        iX = D2 * t;
        iD2 = X * t;
    } else {
        iX = R(1) / X;
        iD2 = R(1) / D2;
    }
    a.v[0] = Nx * iX;
    a.v[1] = Ny * iX;
    a.v[2] = Nz * iX;
    // (R_B - B_x v) / (lambda - v_x) and p + (R_E - v.R_m) / (lambda - v_x)
    a.B[0] = bx;
    a.B[1] = (RBy * X - bx * Ny) * iD2;
    a.B[2] = (RBz * X - bx * Nz) * iD2;
    a.vB = a.v[0] * bx + a.v[1] * a.B[1] + a.v[2] * a.B[2];
    a.w = p + (RE * X - (Nx * Rmx + Ny * Rmy + Nz * Rmz)) * iD2;
    a.eta = sgn * sqrt(rmax(a.w, R(0)));
    const R iden = R(1) / (lam * p + RE + a.eta * bx);
    a.K[0] = (Rmx + p + a.eta * lam * bx) * iden;
    a.K[1] = (Rmy + a.eta * RBy) * iden;
    a.K[2] = (Rmz + a.eta * RBz) * iden;
}
// the normal velocity behind a fast wave at total pressure p as N_x / X
// (fast_state's first step)
template <class R>
SPH_HD void fast_vx(const R Rv[7], R lam, R bx, R p, R& Nx, R& X) {
    const R Rmx = Rv[1], Rmy = Rv[2], Rmz = Rv[3], RE = Rv[4], RBy = Rv[5], RBz = Rv[6];
    const R l2 = R(1) - lam * lam;
    const R A = Rmx - lam * RE + p * l2;
    const R G = RBy * RBy + RBz * RBz;
    const R C = RBy * Rmy + RBz * Rmz;
    X = bx * (A * lam * bx + C) - (A + G) * (lam * p + RE);
    Nx = bx * (A * bx + lam * C) - (A + G) * (p + Rmx);
}
// conserved words (D, m, E, B_y, B_z) of a state from its velocity v, field
// B, v.B, total pressure p, D and E
template <class R>
SPH_HD void fan_cons(const R v[3], const R B[3], R vB, R p, R D, R E, R u[7]) {
    u[0] = D;
#pragma unroll
    for (int a = 0; a < 3; a++) u[1 + a] = (E + p) * v[a] - vB * B[a];
    u[4] = E;
    u[5] = B[1];
    u[6] = B[2];
}
// Next to the contact: the field between the Alfven waves is the HLL
// average of the induction equation over [K_x^L, K_x^R] (MUB09 eq. 45),
// B^c_t = N_t / dK with dK = K_x^R - K_x^L, and the velocity on the side of
// Alfven state a is v = K - B^c (1 - K^2) / (eta - K.B^c) (MUB09 eq. 47).
// Both are formed here without dividing by dK: with Yh = (1 - K^2) /
// (dK (eta - K_x B_x) - K_t.N_t), v_x = K_x - dK Yh B_x and v_t = K_t - Yh N_t.
// The normal velocity jump across the contact is then dK (1 - B_x (Yh_R -
// Yh_L)); its factor dK vanishes where the two Alfven waves coincide, which
// is no fan, so the contact condition solved is g = 1 - B_x (Yh_R - Yh_L) = 0.
template <class R>
SPH_HD R contact_side(const AState<R>& a, R dK, R bx, const R N[3], R vc[3]) {
    const R K2 = a.K[0] * a.K[0] + a.K[1] * a.K[1] + a.K[2] * a.K[2];
    const R Yh = (R(1) - K2) / (dK * (a.eta - a.K[0] * bx) - (a.K[1] * N[1] + a.K[2] * N[2]));
    vc[0] = a.K[0] - dK * Yh * bx;
    vc[1] = a.K[1] - Yh * N[1];
    vc[2] = a.K[2] - Yh * N[2];
    return Yh;
}

template <class R>
struct HlldTol {};
// Secant tolerance on the pressure step (tol: the secant converges
// superlinearly, so the pressure is then good to about tol^1.6) and on the
// residual, which is dimensionless (ftol), and its first step per precision;
// bdeg: a normal field with B_x^2 < bdeg p is degenerate (Alfven waves
// merged with the contact).  The threshold is the same in both precisions, so that FP32 and FP64 solve the
// same fan, and large enough that the five-wave states, whose contact field
// divides by K_x^R - K_x^L ~ B_x, stay well conditioned in FP32 (on the
// axisymmetric torus over 200 steps, FP32 against FP64: L1(rho) 1.2e-5 at
// 1e-6, 4.9e-7 at 1e-4)
template <>
struct HlldTol<double> {
    static constexpr double tol = 1e-8, ftol = 1e-13, bdeg = 1e-4, dp = 1e-4;
};
template <>
struct HlldTol<float> {
    static constexpr float tol = 1e-4f, ftol = 1e-6f, bdeg = 1e-4f, dp = 1e-3f;
};
constexpr int HLLD_ITER = 24;

// residual of the contact condition at total pressure p (g above; with a
// degenerate normal field the Alfven waves merge with the contact and the
// residual is the normal velocity jump between the two fast-wave states);
// also the velocities next to the contact and Nc = (dK, N_y, N_z), from
// which the contact field is B^c_t = N_t / dK (formed once, after the
// iteration)
template <class R>
SPH_HD R contact_residual(const R RL[7], const R RR[7], R lL, R lR, R bx, R p, bool deg, AState<R>& aL,
                          AState<R>& aR, R Nc[3], R vcL[3], R vcR[3]) {
    const R sb = bx < R(0) ? R(-1) : R(1);
    fast_state(RL, lL, bx, p, -sb, aL);
    fast_state(RR, lR, bx, p, sb, aR);
    if (deg) return aR.v[0] - aL.v[0];
    const R dK = aR.K[0] - aL.K[0];
    R N[3];
    N[0] = R(0);
#pragma unroll
    for (int a = 1; a < 3; a++)
        N[a] = (aR.B[a] * (aR.K[0] - aR.v[0]) + bx * aR.v[a]) - (aL.B[a] * (aL.K[0] - aL.v[0]) + bx * aL.v[a]);
    const R YL = contact_side(aL, dK, bx, N, vcL);
    const R YR = contact_side(aR, dK, bx, N, vcR);
    Nc[0] = dK;
    Nc[1] = N[1];
    Nc[2] = N[2];
    return R(1) - bx * (YR - YL);
}

// The residual of contact_residual alone, with fewer divisions, for the
// trial pressures of the iteration: degenerate, (N_x^R X^L - N_x^L X^R) /
// (X^R X^L); else g with the two sides' Yh over one denominator.
template <class R>
SPH_HD R contact_g(const R RL[7], const R RR[7], R lL, R lR, R bx, R p, bool deg) {
    const R sb = bx < R(0) ? R(-1) : R(1);
    if (deg) {
        R NL, XL, NR, XR;
        fast_vx(RL, lL, bx, p, NL, XL);
        fast_vx(RR, lR, bx, p, NR, XR);
        if constexpr (sizeof(R) == 8) return (NR * XL - NL * XR) / (XR * XL);
        else return NR / XR - NL / XL;
    }
    AState<R> aL, aR;
    fast_state(RL, lL, bx, p, -sb, aL);
    fast_state(RR, lR, bx, p, sb, aR);
    const R dK = aR.K[0] - aL.K[0];
    R N[3];
#pragma unroll
    for (int a = 1; a < 3; a++)
        N[a] = (aR.B[a] * (aR.K[0] - aR.v[0]) + bx * aR.v[a]) - (aL.B[a] * (aL.K[0] - aL.v[0]) + bx * aL.v[a]);
    auto nd = [&](const AState<R>& a, R& n, R& d) {
        n = R(1) - (a.K[0] * a.K[0] + a.K[1] * a.K[1] + a.K[2] * a.K[2]);
        d = dK * (a.eta - a.K[0] * bx) - (a.K[1] * N[1] + a.K[2] * N[2]);
    };
    R nL, dL, nR, dR;
    nd(aL, nL, dL);
    nd(aR, nR, dR);
    return R(1) - bx * (nR * dL - nL * dR) / (dR * dL);
}

// Total pressure p + b^2/2 of the HLL state U = (RR - RL) / (lR - lL) in the
// frame: a Newton solve of the con2prim residual from the initial Z of c2p;
// returns -1 where it does not converge (the caller keeps its estimate).
template <class R>
SPH_HD R hll_total_pressure(const R RL[7], const R RR[7], R lL, R lR, R bx, R gam) {
    const R idl = R(1) / (lR - lL);
    R U[7];
#pragma unroll
    for (int q = 0; q < 7; q++) U[q] = (RR[q] - RL[q]) * idl;
    C2PInT<R> q;
    q.D = U[0];
    q.E = U[4];
    q.S2 = U[1] * U[1] + U[2] * U[2] + U[3] * U[3];
    q.SB = U[1] * bx + U[2] * U[5] + U[3] * U[6];
    q.B2 = bx * bx + U[5] * U[5] + U[6] * U[6];
    if (!(q.D > R(0)) || !(q.E > R(0))) return R(-1);
    R Z = q.D + gam * rmax(q.E - R(0.5) * q.B2 - q.D, R(0));
    for (int it = 0; it < 30; it++) {
        R df;
        const R f = c2p_f(q, gam, Z, df);
        const R Zn = Z - f / df;
        if (!(Zn > R(0)) || !isfinite(Zn)) return R(-1);
        const R dz = Zn - Z;
        Z = Zn;
        if (rabs(dz) <= HlldTol<R>::tol * Z) {
            const R v2 = c2p_v2(q, Z);
            if (!(v2 < R(1))) return R(-1);
            const R W2i = R(1) - v2;
            const R p = (gam - R(1)) / gam * (Z * W2i - q.D * sqrt(W2i));
            const R vB = q.SB / Z;
            return rmax(p, R(0)) + R(0.5) * (q.B2 * W2i + vB * vB);
        }
    }
    return R(-1);
}

// The Riemann fan of a solved face, for checking the intermediate states:
// speeds lL <= laL <= lc <= laR <= lR and the conserved words (D, m, E,
// B_y, B_z) of the regions aL, cL, cR, aR, at total pressure p (deg: the
// Alfven waves merged with the contact, no c regions).
template <class R>
struct HlldFan {
    R lL, laL, lc, laR, lR, p;
    R UaL[7], UcL[7], UcR[7], UaR[7];
    int deg, hll;
    int why;  // 0 solved; fallback: 1 no convergence, 2 fast-wave states, 3 fan ordering, 4 non-finite flux
};

// Frame HLLD flux through a surface moving with normal speed vs: out[8] =
// D, m_x, m_y, m_z, E, B_y, B_z, B_x words of F - vs U of the Riemann fan at
// x/t = vs.  Returns 0 for an HLLD flux, 1 where the solver fell back to HLL
// (no converged total pressure, or intermediate states outside the fan's
// wave ordering or non-finite: MUB09 sec. 3.4).  hll_only takes the HLL
// branch directly; fan, if given, receives the fan (both for checking).
// the HLLD part of hlld_frame for a surface inside the fan (lL < vs < lR):
// true with out[0..6] set, false where the solver falls back to HLL.  RL,
// RR = lambda U - F of the outer states; pmk = the reference's HLL estimate
// of the total pressure (hlld.cpp:263), pmin half the smaller outer total
// pressure.
template <class R>
SPH_HD bool hlld_solve(const R RL[7], const R RR[7], R pmk, R pmin, R bx, R vs, R gam, R lL, R lR, R out[8],
                       HlldFan<R>* fan) {
    const R idl = R(1) / (lR - lL);
    // initial total pressure of the secant iteration: the total pressure of
    // the HLL state (MUB09 sec. 3.4; exact for an isolated contact, and on
    // the admissible branch of the contact condition, which has spurious
    // roots), else the reference's estimate
    R p0 = pmk;
    if (!(p0 > R(0)) || !isfinite(p0)) p0 = pmin;
    const bool deg = bx * bx < HlldTol<R>::bdeg * p0;
    if (deg) {
        // the contact of the two-wave problem on the HLL state (Mignone &
        // Bodo 2006): with B_x = 0 the normal velocities behind the fast
        // waves are (p + R_mx) / (lambda p + R_E) and their equality is a
        // quadratic in p, here through the contact speed's root (the
        // solution at B_x = 0, the start of the iteration below otherwise);
        // no positive root (a vacuum between the waves) takes the fallback
        const R UE = (RR[4] - RL[4]) * idl, Um = (RR[1] - RL[1]) * idl;
        const R FE = (lL * RR[4] - lR * RL[4]) * idl, Fm = (lL * RR[1] - lR * RL[1]) * idl;
        const R b = UE + Fm;
        const R disc = rmax(b * b - R(4) * FE * Um, R(0));
        const R ls = R(2) * Um / (b + sqrt(disc));
        p0 = Fm - ls * FE;
        if (!(p0 > R(0)) || !isfinite(p0)) {
            if (fan) fan->why = 1;
            return false;
        }
    } else {
        const R ph = hll_total_pressure(RL, RR, lL, lR, bx, gam);
        if (ph > R(0) && isfinite(ph)) p0 = ph;
    }
    AState<R> aL, aR;
    R Nc[3], vcL[3], vcR[3];
    R pa = p0;
    R fa = contact_residual(RL, RR, lL, lR, bx, pa, deg, aL, aR, Nc, vcL, vcR);
    R pb = p0, fb = fa;
    // with no normal field the two-wave contact is the root itself
    bool ok = rabs(fa) <= HlldTol<R>::ftol || (deg && bx == R(0));
    if (!ok) {
        // the iteration needs only the residual: the states of each trial
        // pressure are temporaries, and those of the last one are formed
        // again after it (fewer registers live across the loop)
        auto g = [&](R p) { return contact_g(RL, RR, lL, lR, bx, p, deg); };
        pb = p0 * (R(1) + HlldTol<R>::dp);
        fb = g(pb);
        ok = rabs(fb) <= HlldTol<R>::ftol;
        for (int it = 0; it < HLLD_ITER && !ok; it++) {
            if (!isfinite(fa) || !isfinite(fb)) break;
            if (fb == fa) {
                // no change of the residual: converged to rounding, unless
                // the step was still large
                ok = fb == R(0) || rabs(pb - pa) <= R(10) * HlldTol<R>::tol * pb;
                break;
            }
            R pn = pb - fb * (pb - pa) / (fb - fa);
            if (!(pn > R(0))) pn = R(0.5) * pb;
            pa = pb;
            fa = fb;
            pb = pn;
            fb = g(pb);
            if (rabs(pb - pa) <= HlldTol<R>::tol * pb || rabs(fb) <= HlldTol<R>::ftol) ok = true;
        }
        fb = contact_residual(RL, RR, lL, lR, bx, pb, deg, aL, aR, Nc, vcL, vcR);
    }
    // admissibility at the converged pressure (aL, aR, Nc, vcL, vcR are
    // those of pb): positive total enthalpy, sub-luminal velocities and the
    // wave ordering of the fan
    R Bc[3] = {bx, R(0), R(0)};
    if (!deg) {
        const R idK = R(1) / Nc[0];
        Bc[1] = Nc[1] * idK;
        Bc[2] = Nc[2] * idK;
    }
    auto sub = [](const R v[3]) { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2] < R(1); };
    bool good = ok && isfinite(fb) && aL.w > R(0) && aR.w > R(0) && sub(aL.v) && sub(aR.v);
    if (fan) fan->why = !(ok && isfinite(fb)) ? 1 : !good ? 2 : 0;
    const R p = pb;
    R lc, laL = lL, laR = lR;
    if (deg) {
        lc = R(0.5) * (aL.v[0] + aR.v[0]);
        good = good && lL <= lc && lc <= lR;
    } else {
        laL = aL.K[0];
        laR = aR.K[0];
        lc = R(0.5) * (vcL[0] + vcR[0]);
        // a fan whose Alfven waves coincide is not one
        good = good && lL <= laL && laL < laR && laL <= lc && lc <= laR && laR <= lR && sub(vcL) && sub(vcR);
    }
    if (fan) {
        if (fan->why == 0 && !good) fan->why = 3;
        fan->laL = laL; fan->lc = lc; fan->laR = laR; fan->p = p; fan->deg = deg ? 1 : 0;
        for (int side = 0; side < 2; side++) {
            const AState<R>& a = side ? aR : aL;
            const R* Rv = side ? RR : RL;
            const R lf = side ? lR : lL, la = side ? laR : laL;
            R* Ua = side ? fan->UaR : fan->UaL;
            R* Uc = side ? fan->UcR : fan->UcL;
            const R ia = R(1) / (lf - a.v[0]);
            fan_cons(a.v, a.B, a.vB, p, Rv[0] * ia, (Rv[4] + p * a.v[0] - a.vB * bx) * ia, Ua);
            for (int q = 0; q < 7; q++) Uc[q] = Ua[q];
            if (!deg) {
                const R* vc = side ? vcR : vcL;
                const R vBc = vc[0] * Bc[0] + vc[1] * Bc[1] + vc[2] * Bc[2];
                const R ic = R(1) / (la - vc[0]);
                fan_cons(vc, Bc, vBc, p, Ua[0] * (la - a.v[0]) * ic, (la * Ua[4] - Ua[1] + p * vc[0] - vBc * bx) * ic,
                         Uc);
            }
        }
    }
    if (!good) return false;
    // the region of the fan that contains x/t = vs: the fast-wave state of
    // its side (F = lambda U - R), then across the Alfven wave if inside it
    const bool left = vs <= lc;
    const AState<R> a = left ? aL : aR;
    R Rv[7];
#pragma unroll
    for (int q = 0; q < 7; q++) Rv[q] = left ? RL[q] : RR[q];
    const R lf = left ? lL : lR;
    const R la = left ? laL : laR;
    R Ux[7], Fx[7];
    {
        const R ia = R(1) / (lf - a.v[0]);
        fan_cons(a.v, a.B, a.vB, p, Rv[0] * ia, (Rv[4] + p * a.v[0] - a.vB * bx) * ia, Ux);
    }
#pragma unroll
    for (int q = 0; q < 7; q++) Fx[q] = lf * Ux[q] - Rv[q];
    const bool inner = !deg && (left ? vs > laL : vs <= laR);
    if (inner) {
        R vc[3];
#pragma unroll
        for (int q = 0; q < 3; q++) vc[q] = left ? vcL[q] : vcR[q];
        const R vBc = vc[0] * Bc[0] + vc[1] * Bc[1] + vc[2] * Bc[2];
        const R ic = R(1) / (la - vc[0]);
        const R Dc = Ux[0] * (la - a.v[0]) * ic;
        const R Ec = (la * Ux[4] - Ux[1] + p * vc[0] - vBc * bx) * ic;
        R Uc[7];
        fan_cons(vc, Bc, vBc, p, Dc, Ec, Uc);
#pragma unroll
        for (int q = 0; q < 7; q++) {
            Fx[q] += la * (Uc[q] - Ux[q]);
            Ux[q] = Uc[q];
        }
    }
#pragma unroll
    for (int q = 0; q < 7; q++) {
        out[q] = Fx[q] - vs * Ux[q];
        good = good && isfinite(out[q]);
    }
    if (fan) {
        fan->hll = good ? 0 : 1;
        if (!good) fan->why = 4;
    }
    return good;
}

// Frame flux through a surface moving with normal speed vs: out[8] = D,
// m_x, m_y, m_z, E, B_y, B_z, B_x words of F - vs U of the Riemann fan at
// x/t = vs.  Returns 0 for the upwind or HLLD flux, 1 where the HLLD solver
// fell back to HLL (no converged total pressure, or intermediate states
// outside the fan's wave ordering or non-finite: MUB09 sec. 3.4).  hll_only
// takes the HLL flux directly; fan, if given, receives the fan (both for
// checking).
template <class R>
SPH_HD int hlld_frame(const FSide<R>& L, const FSide<R>& Rr, R bx, R vs, R gam, R out[8], bool hll_only = false,
                      HlldFan<R>* fan = nullptr) {
    const R lL = rmin(L.lm, Rr.lm), lR = rmax(L.lp, Rr.lp);
    out[7] = -vs * bx;
    if (fan) {
        fan->lL = lL;
        fan->lR = lR;
        fan->hll = 1;
    }
    // the surface outside the fan: the upwind state's flux (HLL's too)
    if (!(vs > lL)) {
#pragma unroll
        for (int q = 0; q < 7; q++) out[q] = L.f[q] - vs * L.u[q];
        if (fan) fan->hll = 0;
        return hll_only ? 1 : 0;
    }
    if (!(vs < lR)) {
#pragma unroll
        for (int q = 0; q < 7; q++) out[q] = Rr.f[q] - vs * Rr.u[q];
        if (fan) fan->hll = 0;
        return hll_only ? 1 : 0;
    }
    // from here on the outer states enter only through R = lambda U - F
    R RL[7], RR[7];
#pragma unroll
    for (int q = 0; q < 7; q++) {
        RL[q] = lL * L.u[q] - L.f[q];
        RR[q] = lR * Rr.u[q] - Rr.f[q];
    }
    const R idl = R(1) / (lR - lL);
    if (!hll_only) {
        const R pmk = (lR * L.pt - lL * Rr.pt + lL * lR * (Rr.u[1] - L.u[1])) * idl;
        if (hlld_solve(RL, RR, pmk, R(0.5) * rmin(L.pt, Rr.pt), bx, vs, gam, lL, lR, out, fan)) return 0;
    }
    // HLL: F_hll - vs U_hll = ((lL - vs) RR - (lR - vs) RL) / (lR - lL)
#pragma unroll
    for (int q = 0; q < 7; q++) out[q] = ((lL - vs) * RR[q] - (lR - vs) * RL[q]) * idl;
    return 1;
}

// map a frame flux (hlld_frame's out) to the coordinate flux words of hll
// (D, S_1..3, tau, the two transverse B^k), densitised; bn = the coordinate
// normal field B^n
template <class R>
SPH_HD void frame_to_coord(const Met<R>& m, const Frame<R>& f, int n, const R F[8], R bn, R out[NFW]) {
    const R c = f.c;
    out[0] = c * F[0];
    out[4] = c * (F[4] - F[0]);
    const R ext = m.sg * m.beta[0] * bn;  // sqrt(gamma) beta^r B^n in the B^r flux (n != r)
    if (n == 0) {
        out[1] = c * (F[1] * f.irn + f.g13 * f.is2 * F[3]);
        out[2] = c * (f.s1 * F[2]);
        out[3] = c * (f.s2 * F[3]);
        out[5] = c * (F[5] * f.is1);                          // B^theta
        out[6] = c * (F[6] * f.is2 + f.gi13 * f.irn * F[7]);  // B^phi
    } else if (n == 1) {
        out[1] = c * (f.g13 * f.is1 * F[2] + F[3] * f.is2);
        out[2] = c * (F[1] * f.irn);
        out[3] = c * (f.s1 * F[2]);
        out[5] = c * (F[5] * f.is1 + f.gi13 * f.is2 * F[6]);  // B^phi
        out[6] = c * (f.s2 * F[6]) + ext;                     // B^r
    } else {
        out[1] = c * (f.s1 * F[2]);
        out[2] = c * (f.s2 * F[3]);
        out[3] = c * (F[1] * f.irn + f.g13 * f.is1 * F[2]);
        out[5] = c * (F[5] * f.is1 + f.gi13 * f.irn * F[7]) + ext;  // B^r
        out[6] = c * (F[6] * f.is2);                                // B^theta
    }
}

// HLLD flux at a face (the coordinate words of hll); gi13 = gamma^r phi at
// the face.  Axis faces (sqrt(gamma) = 0) carry no flux.  Returns 1 where
// the HLL fallback was taken.  remet(g) returns the face metric again (and
// gamma^r phi in g) for the map back to coordinates: the frame is formed a
// second time from it after the solve, with the same operations, so that
// neither the metric nor the frame has to be held through the solve (a
// caller that has the metric at hand returns it).
template <class R, class MetFn>
SPH_HD int hlld_m(const Met<R>& m, R gi13, R gam, Pt<R>& PL, Pt<R>& PR, int n, R out[NFW], MetFn remet,
                  bool hll_only = false, HlldFan<R>* fan = nullptr) {
    if (!(m.sg > R(0))) {
#pragma unroll
        for (int q = 0; q < NFW; q++) out[q] = R(0);
        return 0;
    }
    const R bn = PL.B[n];
    R F[8];
    int fb;
    {
        const Frame<R> fr = make_frame(m, gi13, n);
        const R bx = bn * fr.irn;
        FSide<R> L, Rr;
        frame_side(m, fr, gam, PL, n, bx, L);
        frame_side(m, fr, gam, PR, n, bx, Rr);
        fb = hlld_frame(L, Rr, bx, fr.vs, gam, F, hll_only, fan);
    }
    R g13;
    const Met<R> m2 = remet(g13);
    const Frame<R> f2 = make_frame(m2, g13, n);
    frame_to_coord(m2, f2, n, F, bn, out);
    return fb;
}
template <class R>
SPH_HD int hlld(const Met<R>& m, R gi13, R gam, Pt<R>& PL, Pt<R>& PR, int n, R out[NFW], bool hll_only = false,
                HlldFan<R>* fan = nullptr) {
    return hlld_m(m, gi13, gam, PL, PR, n, out, [&](R& g) { g = gi13; return m; }, hll_only, fan);
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
