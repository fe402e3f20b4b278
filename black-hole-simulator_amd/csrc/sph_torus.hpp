// Fishbone-Moncrief torus of the spherical Kerr-Schild engine, pointwise.
//
// The reference builds its torus in torus::init::fm_init
// (src/fluid/solver/initconds.cpp:121-225, 431-589): constant specific
// angular momentum from the Keplerian orbit at r_max (L0_comp), the
// enthalpy potential (lnh_pot) and a polytrope.  This is the same torus
// (Fishbone & Moncrief 1976) with its velocity written in the Kerr-Schild
// basis (the reference puts Cartesian velocities into spherical slots,
// initconds.cpp:224):
//   l    = u^t u_phi of the prograde circular orbit at r_max
//        = (r^3/2 + a)(r^2 - 2 a r^1/2 + a^2) / (r^3/2 (r^3/2 - 3 r^1/2 + 2a))
//   ln h = 1/2 ln[(1 + sqrt(1 + 4 l^2 Sigma^2 Delta / (A^2 sin^2))) A / (Sigma Delta)]
//          - 1/2 sqrt(1 + 4 l^2 Sigma^2 Delta / (A^2 sin^2)) - 2 a r l / A  - (same at r_in, pi/2)
//   rho  = ((h - 1)(gam - 1) / (kappa gam))^(1/(gam-1)),  p = rho (h - 1)(gam - 1)/gam
//   u^phi from the same potential, u^r = u^theta = 0 (Boyer-Lindquist and
//   Kerr-Schild agree on u^r = 0 orbits), u^t from u.u = -1, and
//   u~^r = u^t beta^r, u~^phi = u^phi (u~^i = u^i + u^t beta^i).
// The functions the torus calls (exp, log, pow) come from the policy M so
// that the device uses its correctly rounded implementations and the oracle
// its quad-precision ones; sqrt and division are IEEE on both.
#pragma once

#include "sph_common.hpp"

namespace sph {

struct TorusPar {
    double a, gam, r_in, r_max, kappa;
    double l, lnh_in;  // angular momentum and ln h bracket at (r_in, pi/2): host scalars
};

inline double fm_l(double a, double r) {
    const double sr = std::sqrt(r), r32 = r * sr;
    return (r32 + a) * (r * r - 2.0 * a * sr + a * a) / (r32 * (r32 - 3.0 * sr + 2.0 * a));
}

// the bracket of ln h before the r_in term is subtracted
template <class M>
SPH_HD double fm_lnh_raw(const TorusPar& T, double r, double s, double c) {
    const double a = T.a, a2 = a * a, l = T.l;
    const double Sig = r * r + a2 * c * c;
    const double Del = r * r - 2.0 * r + a2;
    const double Ar = (r * r + a2) * (r * r + a2) - Del * a2 * s * s;
    const double q = std::sqrt(1.0 + 4.0 * l * l * Sig * Sig * Del / (Ar * Ar * s * s));
    return 0.5 * M::log((1.0 + q) * Ar / (Sig * Del)) - 0.5 * q - 2.0 * a * r * l / Ar;
}

// Torus state at a cell centre (r, sin, cos) with Kerr-Schild g_tt, g_tphi,
// g_phph and beta^r of the cell; returns 0 outside the torus (then nothing
// is written), else rho and p (unnormalised), u~^r, u~^phi.
template <class M>
SPH_HD int fm_cell(const TorusPar& T, double r, double s, double c, double gtt, double gtp, double gpp,
                   double beta_r, double& rho, double& p, double& ur, double& uph) {
    if (!(r >= T.r_in)) return 0;
    const double lnh = fm_lnh_raw<M>(T, r, s, c) - T.lnh_in;
    if (!(lnh > 0.0)) return 0;
    const double h = M::exp(lnh);
    const double gm1 = T.gam - 1.0;
    rho = M::pow((h - 1.0) * gm1 / (T.kappa * T.gam), 1.0 / gm1);
    p = rho * (h - 1.0) * gm1 / T.gam;
    const double a = T.a, a2 = a * a, l = T.l;
    const double Sig = r * r + a2 * c * c;
    const double Del = r * r - 2.0 * r + a2;
    const double Ar = (r * r + a2) * (r * r + a2) - Del * a2 * s * s;
    const double em2chi = Sig * Sig * Del / (Ar * Ar * s * s);
    const double up1 = std::sqrt((-1.0 + std::sqrt(1.0 + 4.0 * l * l * em2chi)) * 0.5);
    uph = 2.0 * a * r * std::sqrt(1.0 + up1 * up1) / std::sqrt(Ar * Sig * Del) + std::sqrt(Sig / Ar) * up1 / s;
    // u^t: gtt (u^t)^2 + 2 gtp u^t u^phi + gpp (u^phi)^2 = -1, positive root
    const double B = 2.0 * gtp * uph, C = gpp * uph * uph + 1.0;
    const double ut = (B + std::sqrt(B * B - 4.0 * gtt * C)) / (-2.0 * gtt);
    ur = ut * beta_r;
    return 1;
}

// numpy's PCG64 (128-bit LCG, XSL-RR output) advanced to output n of a stream
// whose state and increment are given; returns random() of that output,
// (next_uint64 >> 11) * 2^-53.  Output n is produced from the state after
// n + 1 steps.
struct U128 {
    unsigned long long hi, lo;
};
SPH_HD U128 u128_mul(U128 a, U128 b) {
#if defined(__HIP_DEVICE_COMPILE__)
    const unsigned long long lo = a.lo * b.lo;
    const unsigned long long hi = __umul64hi(a.lo, b.lo) + a.hi * b.lo + a.lo * b.hi;
    return {hi, lo};
#else
    const unsigned __int128 x = ((unsigned __int128)a.hi << 64) | a.lo;
    const unsigned __int128 y = ((unsigned __int128)b.hi << 64) | b.lo;
    const unsigned __int128 z = x * y;
    return {(unsigned long long)(z >> 64), (unsigned long long)z};
#endif
}
SPH_HD U128 u128_add(U128 a, U128 b) {
    const unsigned long long lo = a.lo + b.lo;
    return {a.hi + b.hi + (lo < a.lo ? 1ull : 0ull), lo};
}
SPH_HD double pcg64_uniform(U128 state, U128 inc, unsigned long long n) {
    const U128 mult = {0x2360ED051FC65DA4ull, 0x4385DF649FCCF645ull};
    // advance by n + 1 steps (O'Neill's pcg_advance_lcg_128)
    unsigned long long delta = n + 1;
    U128 acc_mult = {0, 1}, acc_plus = {0, 0}, cur_mult = mult, cur_plus = inc;
    while (delta > 0) {
        if (delta & 1) {
            acc_mult = u128_mul(acc_mult, cur_mult);
            acc_plus = u128_add(u128_mul(acc_plus, cur_mult), cur_plus);
        }
        cur_plus = u128_mul(u128_add(cur_mult, U128{0, 1}), cur_plus);
        cur_mult = u128_mul(cur_mult, cur_mult);
        delta >>= 1;
    }
    const U128 s = u128_add(u128_mul(acc_mult, state), acc_plus);
    const unsigned long long x = s.hi ^ s.lo;
    const unsigned rot = (unsigned)(s.hi >> 58);
    const unsigned long long out = (x >> rot) | (x << ((64u - rot) & 63u));
    return (double)(out >> 11) * (1.0 / 9007199254740992.0);
}

}  // namespace sph
