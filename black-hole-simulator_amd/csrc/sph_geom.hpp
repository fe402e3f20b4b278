// Host-side geometry tables of the spherical Kerr-Schild engine.
//
// The Kerr(-Newman) metric in Kerr-Schild coordinates (t, r, theta, phi) with
// M = 1 and charge Q, the same spacetime as the reference's metric::comp
// (src/fluid/physics/spacetime.cpp:20-119), written from the textbook form
// g = eta + H l l with H = (2r - Q^2) / Sigma:
//   g_tt = -(1 - H), g_tr = H, g_tphi = -a H sin^2, g_rr = 1 + H,
//   g_rphi = -a (1 + H) sin^2, g_thth = Sigma,
//   g_phph = sin^2 (Sigma + a^2 (1 + H) sin^2),
//   alpha = (1 + H)^-1/2, beta^r = H / (1 + H), beta^theta = beta^phi = 0.
// gamma^ij is the inverse of the spatial block (theta decouples; the r-phi
// block is inverted directly), sqrt(gamma) = Sigma sin(theta) / alpha.  The
// first derivatives are analytic; Gamma^0_mn is formed from them and g^0m.
// These tables are the engine's only source of geometry; the device never
// evaluates a transcendental function in the step.
#pragma once

#include <cmath>
#include <stdexcept>
#include <vector>

#include "sph_common.hpp"

namespace sph {

struct KS {
    double g[4][4];      // g_mn
    double gu0[4];       // g^0m
    double dg[2][10];    // d_r g, d_theta g in the 10-pair order
    double alpha, beta[3], gam[6], igam[6], sg;
    double da[2];        // d_r alpha, d_theta alpha
    double gam0[10];     // Gamma^0_mn
};

// pole: the point lies on the polar axis (sin theta taken as exactly 0)
inline KS ks_metric(double a, double Q, double r, double s, double c, bool pole) {
    KS k{};
    const double s2 = s * s, c2 = c * c, a2 = a * a;
    const double z = 2.0 * r - Q * Q;
    const double Sig = r * r + a2 * c2;
    const double H = z / Sig;
    const double Sig_r = 2.0 * r, Sig_t = -2.0 * a2 * c * s;
    const double H_r = (2.0 * Sig - z * Sig_r) / (Sig * Sig);
    const double H_t = -z * Sig_t / (Sig * Sig);
    for (int m = 0; m < 4; m++)
        for (int n = 0; n < 4; n++) k.g[m][n] = 0.0;
    k.g[0][0] = H - 1.0;
    k.g[0][1] = k.g[1][0] = H;
    k.g[0][3] = k.g[3][0] = -a * H * s2;
    k.g[1][1] = 1.0 + H;
    k.g[1][3] = k.g[3][1] = -a * (1.0 + H) * s2;
    k.g[2][2] = Sig;
    k.g[3][3] = s2 * (Sig + a2 * (1.0 + H) * s2);
    // derivatives in the pair order 00 01 02 03 11 12 13 22 23 33
    double (&dr)[10] = k.dg[0];
    double (&dt)[10] = k.dg[1];
    for (int q = 0; q < 10; q++) dr[q] = dt[q] = 0.0;
    dr[0] = H_r;                       dt[0] = H_t;
    dr[1] = H_r;                       dt[1] = H_t;
    dr[3] = -a * s2 * H_r;             dt[3] = -a * (H_t * s2 + 2.0 * H * s * c);
    dr[4] = H_r;                       dt[4] = H_t;
    dr[6] = -a * s2 * H_r;             dt[6] = -a * (H_t * s2 + 2.0 * (1.0 + H) * s * c);
    dr[7] = Sig_r;                     dt[7] = Sig_t;
    dr[9] = s2 * (Sig_r + a2 * s2 * H_r);
    dt[9] = 2.0 * s * c * (Sig + a2 * (1.0 + H) * s2) + s2 * (Sig_t + a2 * (H_t * s2 + 2.0 * (1.0 + H) * s * c));
    // lapse and shift from g^tt = -(1 + H), g^tr = H
    k.gu0[0] = -(1.0 + H);
    k.gu0[1] = H;
    k.gu0[2] = 0.0;
    k.gu0[3] = 0.0;
    k.alpha = 1.0 / std::sqrt(1.0 + H);
    k.beta[0] = H / (1.0 + H);
    k.beta[1] = 0.0;
    k.beta[2] = 0.0;
    const double ia3 = 1.0 / ((1.0 + H) * std::sqrt(1.0 + H));
    k.da[0] = -0.5 * ia3 * H_r;
    k.da[1] = -0.5 * ia3 * H_t;
    // spatial metric and its inverse
    k.gam[0] = k.g[1][1]; k.gam[1] = k.g[1][2]; k.gam[2] = k.g[1][3];
    k.gam[3] = k.g[2][2]; k.gam[4] = k.g[2][3]; k.gam[5] = k.g[3][3];
    const double det = k.gam[0] * k.gam[5] - k.gam[2] * k.gam[2];
    k.igam[1] = 0.0;
    k.igam[4] = 0.0;
    k.igam[3] = 1.0 / k.gam[3];
    if (pole) {
        // gamma^rr of the axis limit; gamma^phph diverges and is never read
        // at a theta face (only gamma^thth is), so it is stored as 0
        k.igam[0] = 1.0 / k.gam[0];
        k.igam[2] = 0.0;
        k.igam[5] = 0.0;
    } else {
        k.igam[0] = k.gam[5] / det;
        k.igam[2] = -k.gam[2] / det;
        k.igam[5] = k.gam[0] / det;
    }
    k.sg = pole ? 0.0 : Sig * s / k.alpha;
    // Gamma^0_mn = g^0l (d_m g_ln + d_n g_lm - d_l g_mn) / 2, d_t = d_phi = 0
    auto D = [&](int l, int m, int n) -> double {  // d_l g_mn
        if (l != 1 && l != 2) return 0.0;
        static const int pidx[4][4] = {{0, 1, 2, 3}, {1, 4, 5, 6}, {2, 5, 7, 8}, {3, 6, 8, 9}};
        return k.dg[l - 1][pidx[m][n]];
    };
    const int pm[10] = {0, 0, 0, 0, 1, 1, 1, 2, 2, 3};
    const int pn[10] = {0, 1, 2, 3, 1, 2, 3, 2, 3, 3};
    for (int q = 0; q < 10; q++) {
        const int m = pm[q], n = pn[q];
        double acc = 0.0;
        for (int l = 0; l < 4; l++) acc += k.gu0[l] * (D(m, l, n) + D(n, l, m) - D(l, m, n));
        k.gam0[q] = 0.5 * acc;
    }
    return k;
}

inline void put_met(const KS& k, double* tab, long long npt, int p) {
    // the physics skips these words (a block-diagonal spatial metric and a
    // radial shift); they must be exact zeros
    if (k.beta[1] != 0.0 || k.beta[2] != 0.0 || k.gam[1] != 0.0 || k.gam[4] != 0.0 || k.igam[1] != 0.0 ||
        k.igam[4] != 0.0)
        throw std::logic_error("metric table: gamma_12, gamma_23 or beta^theta, beta^phi not zero");
    tab[M_ALPHA * npt + p] = k.alpha;
    for (int d = 0; d < 3; d++) tab[(M_B1 + d) * npt + p] = k.beta[d];
    for (int q = 0; q < 6; q++) {
        tab[(M_G11 + q) * npt + p] = k.gam[q];
        tab[(M_I11 + q) * npt + p] = k.igam[q];
    }
    tab[M_SG * npt + p] = k.sg;
}

// Tables of one level: nr x nth zones with r faces at exp(lnr0 + i dlnr),
// theta faces at j dth (0 .. pi).  gc: NGC x npt, gf: 2 x NMET x npt (r
// faces, then theta faces), dr: nr + 5.  Theta ghost points mirror the
// interior across the axis; r ghost points continue the logarithmic grid.
inline void build_level_tables(double a, double Q, double rho_fl, double eps_fl, int nr, int nth, double lnr0,
                               double dlnr, std::vector<double>& gc, std::vector<double>& gf,
                               std::vector<double>& dr) {
    const int npt = (nr + 5) * (nth + 5);
    const double dth = M_PI / nth;
    gc.assign((size_t)NGC * npt, 0.0);
    gf.assign((size_t)2 * NMET * npt, 0.0);
    dr.assign(nr + 5, 0.0);
    for (int ii = -2; ii <= nr + 2; ii++)
        dr[ii + 2] = std::exp(lnr0 + (ii + 1) * dlnr) - std::exp(lnr0 + ii * dlnr);
    auto centre_j = [&](int jj) { return jj < 0 ? -1 - jj : (jj >= nth ? 2 * nth - 1 - jj : jj); };
    auto face_j = [&](int jj) { return jj < 0 ? -jj : (jj > nth ? 2 * nth - jj : jj); };
    for (int jj = -2; jj <= nth + 2; jj++)
        for (int ii = -2; ii <= nr + 2; ii++) {
            const int p = (jj + 2) * (nr + 5) + (ii + 2);
            const double rc = std::exp(lnr0 + (ii + 0.5) * dlnr);
            const double rf = std::exp(lnr0 + ii * dlnr);
            // cell centre (mirrored rows for theta ghosts)
            {
                const int jm = centre_j(jj);
                const double th = (jm + 0.5) * dth;
                const double s = std::sin(th), c = std::cos(th);
                const KS k = ks_metric(a, Q, rc, s, c, false);
                put_met(k, gc.data(), npt, p);
                for (int e = 0; e < 2; e++)
                    if (k.dg[e][2] != 0.0 || k.dg[e][5] != 0.0 || k.dg[e][8] != 0.0)
                        throw std::logic_error("metric table: derivative of g_t theta, g_r theta or g_theta phi not zero");
                for (int q = 0; q < 10; q++) {
                    gc[(size_t)(C_DG1 + q) * npt + p] = k.dg[0][q];
                    gc[(size_t)(C_DG2 + q) * npt + p] = k.dg[1][q];
                    gc[(size_t)(C_GAM0 + q) * npt + p] = k.gam0[q];
                }
                gc[(size_t)C_DA1 * npt + p] = k.da[0];
                gc[(size_t)C_DA2 * npt + p] = k.da[1];
                const double rr = rc * std::sqrt(rc);
                gc[(size_t)C_RFL * npt + p] = rho_fl / rr;
                gc[(size_t)C_EFL * npt + p] = eps_fl / rr;
                gc[(size_t)C_R * npt + p] = rc;
                gc[(size_t)C_TH * npt + p] = th;
                gc[(size_t)C_SIN * npt + p] = s;
                gc[(size_t)C_COS * npt + p] = c;
            }
            // r face at (r_{ii-1/2}, theta_jj)
            {
                const int jm = centre_j(jj);
                const double th = (jm + 0.5) * dth;
                put_met(ks_metric(a, Q, rf, std::sin(th), std::cos(th), false), gf.data(), npt, p);
            }
            // theta face at (r_ii, theta_{jj-1/2})
            {
                const int jm = face_j(jj);
                const bool pole = jm == 0 || jm == nth;
                const double th = jm * dth;
                const double s = pole ? 0.0 : std::sin(th);
                const double c = jm == 0 ? 1.0 : (jm == nth ? -1.0 : std::cos(th));
                put_met(ks_metric(a, Q, rc, s, c, pole), gf.data() + (size_t)NMET * npt, npt, p);
            }
        }
}

}  // namespace sph
