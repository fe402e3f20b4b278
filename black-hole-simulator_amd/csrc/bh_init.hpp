// Fishbone-Moncrief torus and vector-potential magnetic field (the initial
// data of src/fluid/main.cpp:98-159) for the device initialisation, with the
// correctly rounded functions of bh_crtrig.hpp in place of the C library's.
// The arithmetic of every expression follows the reference operation for
// operation; quantities the tree's geometry tables already hold (cell r and
// theta, the floor state, the cell and tree metric) are read from them.
#pragma once

#include "bh_crtrig.hpp"
#include "bh_physics.hpp"

namespace bh {

// g_tt, g_tphi, g_phiphi, alpha and beta^phi of metric::comp
// (src/fluid/physics/spacetime.cpp:20-34, 93-100); pow(x, 2) = x*x and
// pow(x, -1) = 1/x are the correctly rounded pow.
struct KsParts { double gtt, gtp, gpp, alpha, beta2; };
BH_HD KsParts ks_parts(double M, double a, double Q, double r, double th) {
    KsParts k;
    const double Sg = r * r + a * a * cr_pow(cr_cos(th), 2);
    const double H = (2 * M * r - Q * Q) * cr_pow(Sg, -1);
    const double s2 = cr_pow(cr_sin(th), 2);
    k.gtt = -1 + H;
    k.gtp = -H * a * s2;
    k.gpp = (r * r + a * a + H * a * a * s2) * s2;
    k.alpha = cr_pow(1 + H, -1.0 / 2.0);
    const double gi00 = -1 - H;
    const double gi03 = 0.0;  // g^{t phi} is never set (spacetime.cpp:35-49)
    k.beta2 = -gi03 / gi00;
    return k;
}

// LNRF quantities, initconds.cpp:151-172
struct LnrfParts { double e2nu, e2psi, enu, epsi, omega; };
BH_HD LnrfParts lnrf_parts(const KsParts& mc) {
    LnrfParts lp;
    lp.e2nu = mc.gtp * mc.gtp / mc.gpp - mc.gtt;
    lp.e2psi = mc.gpp;
    lp.enu = sqrt(lp.e2nu);
    lp.epsi = sqrt(lp.e2psi);
    lp.omega = -mc.gtp / mc.gpp;
    return lp;
}
// ln h of the torus potential, initconds.cpp:174-178
BH_HD double fm_lnh(const KsParts& mc, double l, double W_in) {
    const LnrfParts lp = lnrf_parts(mc);
    if (lp.e2nu <= 0.0) return -1e30;
    return -0.5 * sqrt((4 * l * l * lp.e2nu) / lp.e2psi + 1) +
           0.5 * cr_log((sqrt((4 * l * l * lp.e2nu) / lp.e2psi + 1) + 1) / lp.e2nu) - l * lp.omega + W_in;
}
// potential at the inner edge, initconds.cpp:179-183
BH_HD double fm_W_in(const KsParts& mc, double l) {
    const LnrfParts lp = lnrf_parts(mc);
    if (lp.e2nu <= 0.0) return -1e30;
    return 0.5 * sqrt((4 * l * l * lp.e2nu) / lp.e2psi + 1) -
           0.5 * cr_log((sqrt((4 * l * l * lp.e2nu) / lp.e2psi + 1) + 1) / lp.e2nu) + l * lp.omega;
}
// density of disctherm, initconds.cpp:186-203 (pressure and energy are not used)
BH_HD double fm_rho(double lnH, double Gamma, double K) {
    if (lnH <= 0.0) return 0.0;
    const double H = cr_exp(lnH);
    const double arg = (Gamma - 1) * (H - 1) / (Gamma * K);
    if (arg <= 0.0) return 0.0;
    return cr_pow(arg, 1.0 / (Gamma - 1.0));
}
// Cartesian components of the azimuthal velocity, initconds.cpp:206-225
// (quirk: the azimuthal angular velocity placed in Cartesian slots)
BH_HD void fm_velocity(const KsParts& mc, double l, double xc, double yc, double v[3]) {
    const LnrfParts lp = lnrf_parts(mc);
    v[0] = 0.0; v[1] = 0.0; v[2] = 0.0;
    const double u_phi = sqrt(0.5 * (sqrt(1.0 + 4.0 * l * l * lp.e2nu / lp.e2psi) - 1.0));
    const double uphi = u_phi / lp.epsi + (lp.omega / lp.enu) * sqrt(1.0 + u_phi * u_phi);
    const double B = 2.0 * mc.gtp * uphi;
    const double C = mc.gpp * uphi * uphi + 1.0;
    const double D = B * B - 4.0 * mc.gtt * C;
    if (D < 0.0) return;
    const double ut = (-B - sqrt(D)) / (2.0 * mc.gtt);
    const double Omega = uphi / ut;
    const double vphi = Omega - mc.beta2 / mc.alpha;
    const double rcyl = sqrt(xc * xc + yc * yc);
    if (rcyl < 1e-10) return;
    v[0] = -vphi * yc;
    v[1] = vphi * xc;
}

// Scalars of fm_init (initconds.cpp:228-440): horizon, polytropic constant,
// angular momentum l0 from the LNRF condition at r_max (L0_comp,
// initconds.cpp:121-141), inner potential and ln h at r_max.
struct FmScalars { double r_H, K, l0, W_in, lnh_max; };
// 0 = torus can be placed; 1 = L0_comp's denominator vanishes (the reference
// divides by it); 2 = ln h <= 0 at r_max (the reference prints and returns)
inline int fm_scalars(double M, double a, double Q, double Gamma, FmScalars& s) {
    const double r_max = 12.0 * M;
    const double r_in = 4.5 * M;
    s.r_H = M + sqrt(M * M - a * a - Q * Q);
    s.K = 0.001 * cr_pow(M, Gamma - 1.0);
    const double dr = 1e-4 * M;
    const double th = M_PI / 2.0;
    const KsParts mc0 = ks_parts(M, a, Q, r_max, th);
    const KsParts mcp = ks_parts(M, a, Q, r_max + dr, th);
    const KsParts mcm = ks_parts(M, a, Q, r_max - dr, th);
    const double gtt = mc0.gtt, gtp = mc0.gtp, gpp = mc0.gpp;
    const double dgtt = (mcp.gtt - mcm.gtt) / (2.0 * dr);
    const double dgtp = (mcp.gtp - mcm.gtp) / (2.0 * dr);
    const double dgpp = (mcp.gpp - mcm.gpp) / (2.0 * dr);
    const double A = -gtp * (gpp * (dgtt * dgpp + 2 * dgtp * dgtp) + gtt * dgpp * dgpp) +
                     sqrt((dgtp * dgtp - dgtt * dgpp) *
                          cr_pow((gpp * (gpp * dgtt - gtt * dgpp) + 2 * gtp * gtp * dgpp - 2 * gtp * gpp * dgtp), 2)) +
                     gpp * gpp * dgtt * dgtp + gtt * gpp * dgtp * dgpp + 2 * gtp * gtp * dgtp * dgpp;
    const double B = dgtt * (gpp * gpp * dgtt + 4 * gtp * gtp * dgpp - 4 * gtp * gpp * dgtp) -
                     2 * gtt * (gpp * (dgtt * dgpp - 2 * dgtp * dgtp) + 2 * gtp * dgtp * dgpp) + gtt * gtt * dgpp * dgpp;
    if (fabs(B) < 1e-14) return 1;
    s.l0 = A / B;
    s.W_in = fm_W_in(ks_parts(M, a, Q, r_in, th), s.l0);
    s.lnh_max = fm_lnh(ks_parts(M, a, Q, r_max, th), s.l0, s.W_in);
    if (s.lnh_max <= 0.0) return 2;
    return 0;
}

// A_phi of the single poloidal loop in Cartesian components,
// initconds.cpp:17-23: A_x = -A_phi sin(phi) sin(theta),
// A_y = A_phi cos(phi) sin(theta), A_z = 0, A_phi = max(rho/rho_max - 0.2, 0)
BH_HD void fm_A_xy(double x, double y, double z, double rh, double rho_max, double& Ax, double& Ay) {
    const double Ap = smax(rh / rho_max - 0.2, 0.0);
    const double phi = cr_atan2(y, x);
    const double theta = cr_acos(z / sqrt(x * x + y * y + z * z));
    const double st = cr_sin(theta);
    Ax = -Ap * cr_sin(phi) * st;
    Ay = Ap * cr_cos(phi) * st;
}
BH_HD double fm_A(int comp, double x, double y, double z, double rh, double rho_max) {
    if (comp == 2) return 0.0;
    const double Ap = smax(rh / rho_max - 0.2, 0.0);
    const double phi = cr_atan2(y, x);
    const double theta = cr_acos(z / sqrt(x * x + y * y + z * z));
    if (comp == 0) return -Ap * cr_sin(phi) * cr_sin(theta);
    return Ap * cr_cos(phi) * cr_sin(theta);
}

}  // namespace bh
