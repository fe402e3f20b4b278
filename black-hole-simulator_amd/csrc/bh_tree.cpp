// Host-side block tree: construction, Fishbone-Moncrief initial data,
// vector-potential magnetic field and regrid, restated from the reference
// (src/fluid/grid/grid.cpp, src/fluid/solver/initconds.cpp, main.cpp:98-168).
// Built by the host compiler with contraction off; libm calls are kept as the
// reference writes them so the initial data and geometry round identically.
#include "bh_tree.hpp"

#include <map>
#include <stdexcept>
#include <string>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <sstream>
#include <stdexcept>

namespace bh {

// ------------------------------------------------------------------ metric
// metric::comp and parts, spacetime.cpp:12-134 (quirks Q3, Q5 kept).
Metric metric_comp(double M, double a, double Q, double r, double th) {
    Metric m;
    const double Sg = r * r + a * a * std::pow(std::cos(th), 2);
    const double H = (2 * M * r - Q * Q) * std::pow(Sg, -1);
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) { m.g[i][j] = 0.0; m.gi[i][j] = 0.0; }
    m.g[0][0] = -1 + H;
    m.g[0][1] = m.g[1][0] = H;
    m.g[1][1] = 1 + H;
    m.g[2][2] = Sg;
    m.g[0][3] = m.g[1][3] = m.g[3][0] = m.g[3][1] = -H * a * std::pow(std::sin(th), 2);
    m.g[3][3] = (r * r + a * a + H * a * a * std::pow(std::sin(th), 2)) * std::pow(std::sin(th), 2);
    m.gi[0][0] = -1 - H;
    m.gi[0][1] = m.gi[1][0] = H;
    m.gi[1][1] = 1 - H;
    m.gi[2][2] = std::pow(Sg, -1);
    m.gi[1][3] = m.gi[3][1] = a * std::pow(Sg, -1);
    m.gi[3][3] = std::pow(Sg * std::sin(th) * std::sin(th), -1);
    m.sqrtdetg = (r * r + a * a * std::pow(std::sin(th), 2)) * std::sin(th);
    const double c = std::cos(th), s = std::sin(th);
    const double S1 = r * r + a * a * std::pow(c, 2);
    double d[4][4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            for (int l = 0; l < 4; l++) d[i][j][l] = 0;
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
    m.alpha = std::pow(1 + H, -1.0 / 2.0);
    for (int i = 0; i < 3; i++) m.beta[i] = -m.gi[0][i + 1] / m.gi[0][0];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) { m.gam[i][j] = m.g[i + 1][j + 1]; m.gami[i][j] = m.gi[i + 1][j + 1]; }
    m.sqrtdetgam = m.alpha * m.sqrtdetg;
    return m;
}

GPt gpt_from(const Metric& m) {
    GPt p;
    p.sqrtdetg = m.sqrtdetg; p.sqrtdetgam = m.sqrtdetgam; p.alpha = m.alpha; p.beta0 = m.beta[0];
    p.g11 = m.gam[0][0]; p.g13 = m.gam[0][2]; p.g22 = m.gam[1][1]; p.g33 = m.gam[2][2];
    p.gi13 = m.gami[0][2]; p.gi33 = m.gami[2][2];
    return p;
}
GCell gcell_from(const Metric& m) {
    GCell c;
    c.sqrtdetg = m.sqrtdetg; c.alpha = m.alpha; c.beta0 = m.beta[0];
    c.g11 = m.gam[0][0]; c.g22 = m.gam[1][1]; c.g33 = m.gam[2][2];
    c.gi00 = m.gi[0][0]; c.gi01 = m.gi[0][1]; c.gi11 = m.gi[1][1]; c.gi22 = m.gi[2][2]; c.gi33 = m.gi[3][3];
    return c;
}

// ------------------------------------------------------------------ patch
// patch::patch + patch::cell_init, grid.cpp:19-48 and 149-175.  Fresh cells
// hold zeros (the reference's std::vector<cell> value-initialises them).
Patch::Patch(int lvl_, std::array<int, 3> ijk_, std::array<double, 3> c1, std::array<double, 3> c2, Patch* par)
    : lvl(lvl_), ijk(ijk_), parent(par) {
    children.fill(nullptr);
    for (int i = 0; i <= BLK; i++) {
        double k = double(i) / BLK;
        xe[i] = c1[0] + k * (c2[0] - c1[0]);
        ye[i] = c1[1] + k * (c2[1] - c1[1]);
        ze[i] = c1[2] + k * (c2[2] - c1[2]);
    }
    U.assign((size_t)8 * NCELL, 0.0);
    W.assign((size_t)8 * NCELL, 0.0);
    Bf.assign((size_t)3 * NFACE, 0.0);
    r.resize(NCELL); th.resize(NCELL); xc.resize(NCELL); yc.resize(NCELL); zc.resize(NCELL); vol.resize(NCELL);
    double ddx = dx(), ddy = dy(), ddz = dz();
    for (int i = -GH; i < BLK + GH; i++)
        for (int j = -GH; j < BLK + GH; j++)
            for (int k = -GH; k < BLK + GH; k++) {
                int c = lindx(i, j, k);
                xc[c] = xe[0] + (i + 0.5) * ddx;
                yc[c] = ye[0] + (j + 0.5) * ddy;
                zc[c] = ze[0] + (k + 0.5) * ddz;
                vol[c] = ddx * ddy * ddz;
                double rr = std::sqrt(xc[c] * xc[c] + yc[c] * yc[c] + zc[c] * zc[c]);
                if (rr < 1e-10) rr = 1e-10;
                r[c] = rr;
                th[c] = std::acos(smax(-1.0, smin(1.0, zc[c] / rr)));
            }
}
Prim Patch::getW(int c) const {
    Prim p;
    p.rho = W[(size_t)WRHO * NCELL + c]; p.eps = W[(size_t)WEPS * NCELL + c];
    for (int d = 0; d < 3; d++) { p.v[d] = W[(size_t)(WV0 + d) * NCELL + c]; p.B[d] = W[(size_t)(WB0 + d) * NCELL + c]; }
    return p;
}
void Patch::setW(int c, const Prim& p) {
    W[(size_t)WRHO * NCELL + c] = p.rho; W[(size_t)WEPS * NCELL + c] = p.eps;
    for (int d = 0; d < 3; d++) { W[(size_t)(WV0 + d) * NCELL + c] = p.v[d]; W[(size_t)(WB0 + d) * NCELL + c] = p.B[d]; }
}
Cons Patch::getU(int c) const {
    Cons q;
    q.D = U[(size_t)UD * NCELL + c];
    for (int d = 0; d < 3; d++) q.S[d] = U[(size_t)(US0 + d) * NCELL + c];
    q.tau = U[(size_t)UTAU * NCELL + c];
    for (int d = 0; d < 3; d++) q.B[d] = U[(size_t)(UB0 + d) * NCELL + c];
    return q;
}
void Patch::setU(int c, const Cons& q) {
    U[(size_t)UD * NCELL + c] = q.D;
    for (int d = 0; d < 3; d++) U[(size_t)(US0 + d) * NCELL + c] = q.S[d];
    U[(size_t)UTAU * NCELL + c] = q.tau;
    for (int d = 0; d < 3; d++) U[(size_t)(UB0 + d) * NCELL + c] = q.B[d];
}

// patch::B_init, grid.cpp:208-251 (quirk Q28: the z sweep tests i, not k)
void Patch::B_init() {
    auto WB = [&](int d, int i, int j, int k) { return W[(size_t)(WB0 + d) * NCELL + lindx(i, j, k)]; };
    for (int i = -GH; i < BLK + GH + 1; i++)
        for (int j = -GH; j < BLK + GH; j++)
            for (int k = -GH; k < BLK + GH; k++) {
                double v;
                if (i > -GH && i < BLK + GH) v = (WB(0, i - 1, j, k) + WB(0, i, j, k)) / 2;
                else if (i == -GH) v = WB(0, i, j, k);
                else v = WB(0, i - 1, j, k);
                bfx()[ixf(i, j, k)] = v;
            }
    for (int i = -GH; i < BLK + GH; i++)
        for (int j = -GH; j < BLK + GH + 1; j++)
            for (int k = -GH; k < BLK + GH; k++) {
                double v;
                if (j > -GH && j < BLK + GH) v = (WB(1, i, j - 1, k) + WB(1, i, j, k)) / 2;
                else if (j == -GH) v = WB(1, i, j, k);
                else v = WB(1, i, j - 1, k);
                bfy()[iyf(i, j, k)] = v;
            }
    for (int i = -GH; i < BLK + GH; i++)
        for (int j = -GH; j < BLK + GH; j++)
            for (int k = -GH; k < BLK + GH + 1; k++) {
                double v;
                if (k > -GH && k < BLK + GH) v = (WB(2, i, j, k - 1) + WB(2, i, j, k)) / 2;
                else if (i == -GH) v = WB(2, i, j, k);
                else v = WB(2, i, j, k - 1);
                bfz()[izf(i, j, k)] = v;
            }
}

// -------------------------------------------------------------- tree
GPt HostTree::tree_point(double r, double th) const {
    return gpt_from(metric_comp(prm.M, prm.a, prm.Q, r, th));
}

// amrtree::pvfs, grid.cpp:133-146
Prim HostTree::pvfs(double r) const {
    double rfscal = 1.0 / std::pow(std::sqrt(smax(r, 1e-10) / 1.0), 3.0);
    return floor_prim(1e-7 * rfscal, 1e-5 * rfscal);
}

// amrtree ctor, grid.cpp:177-206: roots in (i, j, k) order, then the floor
// state and its conserved form in every stored cell.
HostTree::HostTree(const double lo[3], const double hi[3], int nqlt, double M, double a, double Q, double gm) {
    prm = {M, a, Q, gm};
    double px = (hi[0] - lo[0]) / nqlt, py = (hi[1] - lo[1]) / nqlt, pz = (hi[2] - lo[2]) / nqlt;
    for (int i = 0; i < nqlt; i++)
        for (int j = 0; j < nqlt; j++)
            for (int k = 0; k < nqlt; k++) {
                std::array<double, 3> c1 = {lo[0] + i * px, lo[1] + j * py, lo[2] + k * pz};
                std::array<double, 3> c2 = {c1[0] + px, c1[1] + py, c1[2] + pz};
                quilt.push_back(std::make_unique<Patch>(0, std::array<int, 3>{i, j, k}, c1, c2, nullptr));
            }
    int nf = 0;
    for (auto& p : quilt)
        for (int c = 0; c < NCELL; c++) {
            Prim w = pvfs(p->r[c]);
            p->setW(c, w);
            p->setU(c, ptoc(tree_point(p->r[c], p->th[c]), prm.gamma, w, nf));
        }
    faults += nf;
}

// Replace the block table by a caller's, meta[nb][14] in the layout of
// export_blocks (level, i, j, k, leaf, parent, children[8]) and in quilt
// order.  Roots are patches of the constructor's root grid (their boxes are
// taken from it); every other block is a child of an earlier block, in the
// octant its ijk names, with the box refine() would give it, so the
// geometry is the reference's bit for bit.  Every zone starts in the floor
// state; the caller sets the state next.
void HostTree::set_blocks(int nb, const int* meta) {
    if (nb <= 0 || !meta) throw std::invalid_argument("set_blocks: empty block table");
    std::map<std::array<int, 3>, const Patch*> roots;
    for (auto& pp : quilt)
        if (pp->lvl == 0) roots[pp->ijk] = pp.get();
    auto bad = [](int b, const char* what) {
        throw std::invalid_argument("set_blocks: block " + std::to_string(b) + ": " + what);
    };
    std::vector<std::unique_ptr<Patch>> nq(nb);
    for (int b = 0; b < nb; b++) {
        const int* m = meta + 14 * (size_t)b;
        const std::array<int, 3> ijk = {m[1], m[2], m[3]};
        const int par = m[5];
        if (par < 0) {
            auto it = roots.find(ijk);
            if (m[0] != 0 || it == roots.end()) bad(b, "a block without parent must be a root of the constructor grid");
            const Patch* r = it->second;
            nq[b] = std::make_unique<Patch>(0, ijk, std::array<double, 3>{r->xe[0], r->ye[0], r->ze[0]},
                                            std::array<double, 3>{r->xe[BLK], r->ye[BLK], r->ze[BLK]}, nullptr);
        } else {
            if (par >= b) bad(b, "its parent must come earlier in the table");
            const Patch* P = nq[par].get();
            const int* pm = meta + 14 * (size_t)par;
            int oct[3];
            for (int e = 0; e < 3; e++) {
                oct[e] = ijk[e] - 2 * P->ijk[e];
                if (oct[e] != 0 && oct[e] != 1) bad(b, "ijk is not inside its parent");
            }
            const int h = 4 * oct[0] + 2 * oct[1] + oct[2];
            if (m[0] != P->lvl + 1 || pm[6 + h] != b) bad(b, "parent and child entries disagree");
            const double lo[3] = {P->xe[0], P->ye[0], P->ze[0]}, hi[3] = {P->xe[BLK], P->ye[BLK], P->ze[BLK]};
            std::array<double, 3> a, c;
            for (int e = 0; e < 3; e++) {
                const double mid = (lo[e] + hi[e]) / 2;
                a[e] = oct[e] ? mid : lo[e];
                c[e] = oct[e] ? hi[e] : mid;
            }
            nq[b] = std::make_unique<Patch>(m[0], ijk, a, c, nq[par].get());
        }
    }
    for (int b = 0; b < nb; b++) {
        const int* m = meta + 14 * (size_t)b;
        bool any = false;
        for (int h = 0; h < 8; h++) {
            const int c = m[6 + h];
            if (c < 0) continue;
            if (c >= nb || meta[14 * (size_t)c + 5] != b) bad(b, "child entry does not point back");
            nq[b]->children[h] = nq[c].get();
            any = true;
        }
        if ((m[4] != 0) == any) bad(b, "leaf flag and children disagree");
        nq[b]->leaf = !any;
    }
    int nf = 0;
    for (auto& p : nq)
        for (int c = 0; c < NCELL; c++) {
            Prim w = pvfs(p->r[c]);
            p->setW(c, w);
            p->setU(c, ptoc(tree_point(p->r[c], p->th[c]), prm.gamma, w, nf));
        }
    faults += nf;
    quilt = std::move(nq);
    erased_.clear();
}

int HostTree::index_of(const Patch* p) const {
    if (!p) return -1;
    for (size_t i = 0; i < quilt.size(); i++)
        if (quilt[i].get() == p) return (int)i;
    return -2;
}

// ------------------------------------------------------- initial torus
// Fishbone-Moncrief helpers, initconds.cpp:121-225.
namespace {
struct Lnrf { double e2nu, e2psi, enu, epsi, omega; };
Lnrf lnrf_of(const Metric& mc) {
    const double gtt = mc.g[0][0], gtp = mc.g[0][3], gpp = mc.g[3][3];
    Lnrf lp;
    lp.e2nu = gtp * gtp / gpp - gtt;
    lp.e2psi = gpp;
    lp.enu = std::sqrt(lp.e2nu);
    lp.epsi = std::sqrt(lp.e2psi);
    lp.omega = -gtp / gpp;
    return lp;
}
double lnh_pot(const Metric& mc, double l, double W_in) {
    const Lnrf lp = lnrf_of(mc);
    if (lp.e2nu <= 0.0) return -1e30;
    return -0.5 * std::sqrt((4 * l * l * lp.e2nu) / lp.e2psi + 1) +
           0.5 * std::log((std::sqrt((4 * l * l * lp.e2nu) / lp.e2psi + 1) + 1) / lp.e2nu) - l * lp.omega + W_in;
}
double W_in_of(const Metric& mc, double l) {
    const Lnrf lp = lnrf_of(mc);
    if (lp.e2nu <= 0.0) return -1e30;
    return 0.5 * std::sqrt((4 * l * l * lp.e2nu) / lp.e2psi + 1) -
           0.5 * std::log((std::sqrt((4 * l * l * lp.e2nu) / lp.e2psi + 1) + 1) / lp.e2nu) + l * lp.omega;
}
struct Therm { double rho, p, eps; };
Therm disctherm(double lnH, double Gamma, double K) {
    Therm t{0.0, 0.0, 0.0};
    if (lnH <= 0.0) return t;
    const double H = std::exp(lnH);
    const double arg = (Gamma - 1) * (H - 1) / (Gamma * K);
    if (arg <= 0.0) return t;
    t.rho = std::pow(arg, 1.0 / (Gamma - 1.0));
    t.p = K * std::pow(arg, Gamma / (Gamma - 1.0));
    t.eps = t.p / ((Gamma - 1.0) * t.rho);
    return t;
}
// quirk Q38: Cartesian components of the azimuthal velocity
std::array<double, 3> velcomp(const Metric& mc, const Lnrf& lp, double l, double xc, double yc) {
    const double u_phi = std::sqrt(0.5 * (std::sqrt(1.0 + 4.0 * l * l * lp.e2nu / lp.e2psi) - 1.0));
    const double uphi = u_phi / lp.epsi + (lp.omega / lp.enu) * std::sqrt(1.0 + u_phi * u_phi);
    const double gtt = mc.g[0][0], gtp = mc.g[0][3], gpp = mc.g[3][3];
    const double B = 2.0 * gtp * uphi;
    const double C = gpp * uphi * uphi + 1.0;
    const double D = B * B - 4.0 * gtt * C;
    if (D < 0.0) return {0.0, 0.0, 0.0};
    const double ut = (-B - std::sqrt(D)) / (2.0 * gtt);
    const double Omega = uphi / ut;
    const double vphi = Omega - mc.beta[2] / mc.alpha;
    const double rcyl = std::sqrt(xc * xc + yc * yc);
    if (rcyl < 1e-10) return {0.0, 0.0, 0.0};
    return {-vphi * yc, vphi * xc, 0.0};
}
// LNRF angular momentum at r_max, initconds.cpp:121-141
bool L0_comp(const TreeParams& P, double r_max, double M, double& out) {
    const double dr = 1e-4 * M;
    const Metric mc0 = metric_comp(P.M, P.a, P.Q, r_max, M_PI / 2.0);
    const Metric mcp = metric_comp(P.M, P.a, P.Q, r_max + dr, M_PI / 2.0);
    const Metric mcm = metric_comp(P.M, P.a, P.Q, r_max - dr, M_PI / 2.0);
    double gtt = mc0.g[0][0], gtp = mc0.g[0][3], gpp = mc0.g[3][3];
    const double dgtt = (mcp.g[0][0] - mcm.g[0][0]) / (2.0 * dr);
    const double dgtp = (mcp.g[0][3] - mcm.g[0][3]) / (2.0 * dr);
    const double dgpp = (mcp.g[3][3] - mcm.g[3][3]) / (2.0 * dr);
    double A = -gtp * (gpp * (dgtt * dgpp + 2 * dgtp * dgtp) + gtt * dgpp * dgpp) +
               std::sqrt((dgtp * dgtp - dgtt * dgpp) *
                         std::pow((gpp * (gpp * dgtt - gtt * dgpp) + 2 * gtp * gtp * dgpp - 2 * gtp * gpp * dgtp), 2)) +
               gpp * gpp * dgtt * dgtp + gtt * gpp * dgtp * dgpp + 2 * gtp * gtp * dgtp * dgpp;
    double B = dgtt * (gpp * gpp * dgtt + 4 * gtp * gtp * dgpp - 4 * gtp * gpp * dgtp) -
               2 * gtt * (gpp * (dgtt * dgpp - 2 * dgtp * dgtp) + 2 * gtp * dgtp * dgpp) + gtt * gtt * dgpp * dgpp;
    if (std::abs(B) < 1e-14) return false;
    out = A / B;
    return true;
}
}  // namespace

// torus::init::fm_init, initconds.cpp:228-591 (live code; quirk Q34).
// Returns 0 on success, 1 when the reference would throw, 2 when it would
// print an error and leave the tree unchanged.
int HostTree::fm_init() {
    const double M = prm.M, a = prm.a, Q = prm.Q, Gamma = prm.gamma;
    const double r_max = 12.0 * M;
    const double r_in = 4.5 * M;
    double r_H = M + std::sqrt(M * M - a * a - Q * Q);
    double K = 0.001 * std::pow(M, Gamma - 1.0);
    double rho_tgt = 1.0;
    double l0;
    if (!L0_comp(prm, r_max, M, l0)) {
        last_msg = "LNRF angular momentum: denominator is zero (r_max inside the ergosphere)";
        return 1;
    }
    const Metric mc_in = metric_comp(M, a, Q, r_in, M_PI / 2.0);
    const double W_in = W_in_of(mc_in, l0);
    const Metric mc_max = metric_comp(M, a, Q, r_max, M_PI / 2.0);
    const double lnh_max = lnh_pot(mc_max, l0, W_in);
    if (lnh_max <= 0.0) {
        last_msg = "ln(h) <= 0 at r = r_max: the torus cannot be initialised";
        return 2;
    }
    double rho_max = 0.0;
    for (auto& p : quilt)
        for (int i = 0; i < BLK; i++)
            for (int j = 0; j < BLK; j++)
                for (int k = 0; k < BLK; k++) {
                    int c = lindx(i, j, k);
                    double r = p->r[c], th = p->th[c];
                    if (!std::isfinite(r) || !std::isfinite(th) || r <= 1.05 * r_H || std::abs(std::sin(th)) < 0.01)
                        continue;
                    Metric mc = metric_comp(M, a, Q, r, th);
                    const double lnH = lnh_pot(mc, l0, W_in);
                    if (lnH <= 0.0 || !std::isfinite(lnH)) continue;
                    const Therm thm = disctherm(lnH, Gamma, K);
                    if (std::isfinite(thm.rho)) rho_max = smax(rho_max, thm.rho);
                }
    if (rho_max < 1e-14) {
        last_msg = "no torus cells found";
        return 2;
    }
    double rho_scal = rho_tgt / rho_max;
    for (auto& p : quilt)
        for (int i = 0; i < BLK; i++)
            for (int j = 0; j < BLK; j++)
                for (int k = 0; k < BLK; k++) {
                    int c = lindx(i, j, k);
                    double r = p->r[c], th = p->th[c];
                    Prim W_p = pvfs(r);
                    if (!std::isfinite(r) || !std::isfinite(th) || r <= 1.05 * r_H || std::abs(std::sin(th)) < 0.01) {
                        p->setW(c, W_p);
                        continue;
                    }
                    Metric mc = metric_comp(M, a, Q, r, th);
                    const double lnH = lnh_pot(mc, l0, W_in);
                    if (lnH <= 0.0 || !std::isfinite(lnH)) {
                        p->setW(c, W_p);
                        continue;
                    }
                    const Therm thm = disctherm(lnH, Gamma, K);
                    W_p.rho = rho_scal * thm.rho;
                    double pp = K * std::pow(W_p.rho, Gamma);
                    W_p.eps = pp / ((Gamma - 1.0) * W_p.rho);
                    const Lnrf lp = lnrf_of(mc);
                    const auto vel = velcomp(mc, lp, l0, p->xc[c], p->yc[c]);
                    W_p.v[0] = vel[0]; W_p.v[1] = vel[1]; W_p.v[2] = vel[2];
                    p->setW(c, W_p);
                }
    std::ostringstream os;
    os.precision(17);
    os << "L_0=" << l0 << " W_in=" << W_in << " ln(h)_max=" << lnh_max << " rho_max=" << rho_max;
    last_msg = os.str();
    return 0;
}

// global maximum density over active zones, main.cpp:101-111
double HostTree::rho_max() const {
    double rm = 0.0;
    for (const auto& p : quilt)
        for (int i = 0; i < BLK; i++)
            for (int j = 0; j < BLK; j++)
                for (int k = 0; k < BLK; k++) rm = smax(rm, p->W[(size_t)WRHO * NCELL + lindx(i, j, k)]);
    return rm;
}

// torus::init::B_pot_init, initconds.cpp:17-52 and 594-668 (quirks Q35, Q39)
void HostTree::B_pot_init(Patch& p, double glmx_rho) {
    const double dx = p.dx(), dy = p.dy(), dz = p.dz();
    double rho_max = glmx_rho;
    if (rho_max <= 1e-10) return;
    auto rho = [&](int i, int j, int k) { return p.W[(size_t)WRHO * NCELL + lindx(i, j, k)]; };
    for (int i = 0; i < BLK; i++)
        for (int j = 0; j < BLK; j++)
            for (int k = 0; k < BLK; k++) rho_max = smax(rho_max, rho(i, j, k));
    const int top = BLK + GH - 1;
    auto edge_rhox = [&](int i, int j, int k) {
        int j1 = std::min(j + 1, top), k1 = std::min(k + 1, top);
        return (rho(i, j, k) + rho(i, j1, k) + rho(i, j, k1) + rho(i, j1, k1)) / 4;
    };
    auto edge_rhoy = [&](int i, int j, int k) {
        int i1 = std::min(i + 1, top), k1 = std::min(k + 1, top);
        return (rho(i, j, k) + rho(i1, j, k) + rho(i, j, k1) + rho(i1, j, k1)) / 4;
    };
    auto edge_rhoz = [&](int i, int j, int k) {
        int i1 = std::min(i + 1, top), j1 = std::min(j + 1, top);
        return (rho(i, j, k) + rho(i1, j, k) + rho(i, j1, k) + rho(i1, j1, k)) / 4;
    };
    // A_vpot, initconds.cpp:17-23: only A_phi is non-zero
    auto A = [&](int comp, double x, double y, double z, double rh) {
        double Ap = smax(rh / rho_max - 0.2, 0.0);
        double phi = std::atan2(y, x);
        double theta = std::acos(z / std::sqrt(x * x + y * y + z * z));
        if (comp == 0) return -Ap * std::sin(phi) * std::sin(theta);
        if (comp == 1) return Ap * std::cos(phi) * std::sin(theta);
        return 0.0;
    };
    // edge positions, initconds.cpp:41-52: x-edge (xc, yc+dy/2, zc+dz/2),
    // y-edge (xc+dx/2, yc, zc+dz/2), z-edge (xc+dx/2, yc+dy/2, zc)
    auto A_xedge = [&](int comp, int i, int j, int k) {
        int c = lindx(i, j, k);
        return A(comp, p.xc[c], p.yc[c] + dy / 2.0, p.zc[c] + dz / 2.0, edge_rhox(i, j, k));
    };
    auto A_yedge = [&](int comp, int i, int j, int k) {
        int c = lindx(i, j, k);
        return A(comp, p.xc[c] + dx / 2.0, p.yc[c], p.zc[c] + dz / 2.0, edge_rhoy(i, j, k));
    };
    auto A_zedge = [&](int comp, int i, int j, int k) {
        int c = lindx(i, j, k);
        return A(comp, p.xc[c] + dx / 2.0, p.yc[c] + dy / 2.0, p.zc[c], edge_rhoz(i, j, k));
    };
    for (int i = -GH + 1; i < BLK + GH; i++)
        for (int j = -GH + 1; j < BLK + GH; j++)
            for (int k = -GH + 1; k < BLK + GH; k++) {
                // x-face: A_y on y-edges (k, k-1), A_z on z-edges (j, j-1)
                double Ay_kp = A_yedge(1, i, j, k), Ay_km = A_yedge(1, i, j, k - 1);
                double Az_jp = A_zedge(2, i, j, k), Az_jm = A_zedge(2, i, j - 1, k);
                p.bfx()[ixf(i, j, k)] = (Az_jp - Az_jm) / dy - (Ay_kp - Ay_km) / dz;
                // y-face: A_x on x-edges (k, k-1), A_z on z-edges (i, i-1)
                double Ax_kp = A_xedge(0, i, j, k), Ax_km = A_xedge(0, i, j, k - 1);
                double Az_ip = A_zedge(2, i, j, k), Az_im = A_zedge(2, i - 1, j, k);
                p.bfy()[iyf(i, j, k)] = (Ax_kp - Ax_km) / dz - (Az_ip - Az_im) / dx;
                // z-face: A_y on x-edges (i, i-1), A_x on y-edges (j, j-1) (quirk Q39)
                double Ay_ip = A_xedge(1, i, j, k), Ay_im = A_xedge(1, i - 1, j, k);
                double Ax_jp = A_yedge(0, i, j, k), Ax_jm = A_yedge(0, i, j - 1, k);
                p.bfz()[izf(i, j, k)] = (Ay_ip - Ay_im) / dx - (Ax_jp - Ax_jm) / dy;
            }
    // constrans::f2cB, transport.cpp:60-75 (cell metric)
    for (int i = -GH + 1; i < BLK + GH; i++)
        for (int j = -GH + 1; j < BLK + GH; j++)
            for (int k = -GH + 1; k < BLK + GH; k++) {
                int c = lindx(i, j, k);
                double b0 = (p.bfx()[ixf(i, j, k)] + p.bfx()[ixf(i - 1, j, k)]) / 2;
                double b1 = (p.bfy()[iyf(i, j, k)] + p.bfy()[iyf(i, j - 1, k)]) / 2;
                double b2 = (p.bfz()[izf(i, j, k)] + p.bfz()[izf(i, j, k - 1)]) / 2;
                p.w(WB0, c) = b0; p.w(WB1, c) = b1; p.w(WB2, c) = b2;
                double sg = metric_comp(1.0, 0.0, 0.0, p.r[c], p.th[c]).sqrtdetg;
                p.u(UB0, c) = sg * b0; p.u(UB1, c) = sg * b1; p.u(UB2, c) = sg * b2;
            }
}

// U = ptoc(W) on active zones, main.cpp:148-159
void HostTree::ptoc_interior() {
    int nf = 0;
    for (auto& p : quilt)
        for (int i = 0; i < BLK; i++)
            for (int j = 0; j < BLK; j++)
                for (int k = 0; k < BLK; k++) {
                    int c = lindx(i, j, k);
                    p->setU(c, ptoc(tree_point(p->r[c], p->th[c]), prm.gamma, p->getW(c), nf));
                }
    faults += nf;
}

// ctop, else floor state with the cell's old B and its conserved form
// (the recovery used by refine and ccrstr, grid.cpp:641-648 and 705-712)
void HostTree::floor_or_keep(Patch& p, int c, const Cons& q) {
    int nf = 0;
    GPt g = tree_point(p.r[c], p.th[c]);
    Prim pv;
    if (ctop(g, prm.gamma, q, pv, nf)) {
        p.setW(c, pv);
    } else {
        Prim fl = pvfs(p.r[c]);
        for (int d = 0; d < 3; d++) fl.B[d] = p.w(WB0 + d, c);
        p.setW(c, fl);
        p.setU(c, ptoc(g, prm.gamma, fl, nf));
    }
    faults += nf;
}

// ------------------------------------------------------------- regrid
// The eight children of a patch, appended to the quilt in octant order
// h = 4 ix + 2 iy + iz (the order amrtree::refine creates them in,
// grid.cpp:574-600), each with the half-open box of its octant.
void HostTree::append_children(Patch* p) {
    const double lo[3] = {p->xe[0], p->ye[0], p->ze[0]}, hi[3] = {p->xe[BLK], p->ye[BLK], p->ze[BLK]};
    const double mid[3] = {(lo[0] + hi[0]) / 2, (lo[1] + hi[1]) / 2, (lo[2] + hi[2]) / 2};
    for (int h = 0; h < 8; h++) {
        const int oct[3] = {(h >> 2) & 1, (h >> 1) & 1, h & 1};
        std::array<double, 3> a, b;
        std::array<int, 3> ijk;
        for (int e = 0; e < 3; e++) {
            a[e] = oct[e] ? mid[e] : lo[e];
            b[e] = oct[e] ? hi[e] : mid[e];
            ijk[e] = 2 * p->ijk[e] + oct[e];
        }
        quilt.push_back(std::make_unique<Patch>(p->lvl + 1, ijk, a, b, p));
        p->children[h] = quilt.back().get();
    }
}

// amrtree::refine, grid.cpp:574-655 (quirk Q24: +/- neighbours swapped)
void HostTree::refine(Patch* p) {
    if (!p->leaf) return;
    append_children(p);
    for (int h = 0; h < 8; h++) {
        Patch* ch = p->children[h];
        int ix = (h >> 2) & 1, iy = (h >> 1) & 1, iz = h & 1;
        for (int i = 0; i < BLK; i++)
            for (int j = 0; j < BLK; j++)
                for (int k = 0; k < BLK; k++) {
                    int pi = (i + ix * BLK) / 2, pj = (j + iy * BLK) / 2, pk = (k + iz * BLK) / 2;
                    double ox = (i % 2 == 0) ? -0.25 : 0.25;
                    double oy = (j % 2 == 0) ? -0.25 : 0.25;
                    double oz = (k % 2 == 0) ? -0.25 : 0.25;
                    int pc = lindx(pi, pj, pk);
                    int ppx = lindx(pi + 1, pj, pk), pmx = lindx(pi - 1, pj, pk);
                    int ppy = lindx(pi, pj + 1, pk), pmy = lindx(pi, pj - 1, pk);
                    int ppz = lindx(pi, pj, pk + 1), pmz = lindx(pi, pj, pk - 1);
                    int cc = lindx(i, j, k);
                    for (int v = 0; v < 8; v++) {
                        double vc = p->u(v, pc);
                        double sx = minmod(vc - p->u(v, ppx), p->u(v, pmx) - vc);
                        double sy = minmod(vc - p->u(v, ppy), p->u(v, pmy) - vc);
                        double sz = minmod(vc - p->u(v, ppz), p->u(v, pmz) - vc);
                        ch->u(v, cc) = vc + sx * ox + sy * oy + sz * oz;
                    }
                    floor_or_keep(*ch, cc, ch->getU(cc));
                }
        ch->B_init();
    }
    p->leaf = false;
}

// amrtree::ccrstr, grid.cpp:657-718: volume-weighted restriction
void HostTree::ccrstr(Patch* p) {
    for (int pi = 0; pi < BLK; pi++)
        for (int pj = 0; pj < BLK; pj++)
            for (int pk = 0; pk < BLK; pk++) {
                int pc = lindx(pi, pj, pk);
                double acc[8];
                for (int v = 0; v < 8; v++) acc[v] = 0.0;
                double volt = 0.0;
                for (int h = 0; h < 8; h++) {
                    Patch* ch = p->children[h];
                    if (!ch) continue;
                    int ix = (h >> 2) & 1, iy = (h >> 1) & 1, iz = h & 1;
                    for (int di = 0; di < 2; di++)
                        for (int dj = 0; dj < 2; dj++)
                            for (int dk = 0; dk < 2; dk++) {
                                int fi = 2 * (pi - ix * (BLK / 2)) + di;
                                int fj = 2 * (pj - iy * (BLK / 2)) + dj;
                                int fk = 2 * (pk - iz * (BLK / 2)) + dk;
                                if (fi < 0 || fi >= BLK || fj < 0 || fj >= BLK || fk < 0 || fk >= BLK) continue;
                                int fc = lindx(fi, fj, fk);
                                double vol = ch->vol[fc];
                                volt += vol;
                                for (int v = 0; v < 8; v++) acc[v] += vol * ch->u(v, fc);
                            }
                }
                if (volt > 0.0) {
                    double inv = 1.0 / volt;
                    for (int v = 0; v < 8; v++) acc[v] *= inv;
                }
                for (int v = 0; v < 8; v++) p->u(v, pc) = acc[v];
                floor_or_keep(*p, pc, p->getU(pc));
            }
    p->B_init();
}

// amrtree::flatten, grid.cpp:720-736 (children removed, quilt order kept)
void HostTree::flatten(Patch* p) {
    if (p->leaf) return;
    for (int h = 0; h < 8; h++)
        if (p->children[h] && !p->children[h]->leaf) return;
    ccrstr(p);
    for (int h = 0; h < 8; h++) {
        Patch* ch = p->children[h];
        if (!ch) continue;
        erased_.push_back(ch);
        quilt.erase(std::remove_if(quilt.begin(), quilt.end(),
                                   [&](const std::unique_ptr<Patch>& q) { return q.get() == ch; }),
                    quilt.end());
        p->children[h] = nullptr;
    }
    p->leaf = true;
}

// The refinement verdict of one leaf (the flag pass of amrtree::regrid,
// grid.cpp:744-795): it asks to be refined where a zone lies inside r < 3
// or its density or field varies by more than half its value across a zone,
// and may be erased only if no zone's density varies by more than a tenth.
// The operations are the reference's, so that the verdicts, and with them
// the block tables, agree bit for bit: the density contrast is the length
// of the centred-difference gradient times dx (also for dy, dz) over rho;
// the field's is the sum over the axes of each component's own centred
// difference times the spacing, over |B|.
HostTree::LeafVerdict HostTree::leaf_verdict(Patch* p) const {
    constexpr double kRefine = 0.5, kKeep = 0.1, kHoleZone = 3.0;
    constexpr int kDeepest = 4;
    const bool deeper = p->lvl < kDeepest;
    const double h[3] = {p->dx(), p->dy(), p->dz()};
    LeafVerdict v{false, true};
    auto centred = [&](int var, int i, int j, int k, int axis) {
        const int up = axis == 0 ? lindx(i + 1, j, k) : axis == 1 ? lindx(i, j + 1, k) : lindx(i, j, k + 1);
        const int dn = axis == 0 ? lindx(i - 1, j, k) : axis == 1 ? lindx(i, j - 1, k) : lindx(i, j, k - 1);
        return std::abs(p->w(var, up) - p->w(var, dn)) / (2 * h[axis]);
    };
    for (int i = 0; i < BLK; i++)
        for (int j = 0; j < BLK; j++)
            for (int k = 0; k < BLK; k++) {
                const int c = lindx(i, j, k);
                if (p->r[c] < kHoleZone && deeper) {
                    v.refine = true;
                    v.erasable = false;
                    continue;
                }
                const double rho = p->w(WRHO, c);
                if (rho > 0.0) {
                    const double gx = centred(WRHO, i, j, k, 0), gy = centred(WRHO, i, j, k, 1),
                                 gz = centred(WRHO, i, j, k, 2);
                    const double contrast = std::sqrt(gx * gx + gy * gy + gz * gz) * h[0] / rho;
                    if (contrast > kRefine && deeper) { v.refine = true; v.erasable = false; }
                    if (contrast > kKeep) v.erasable = false;
                }
                const double b0 = p->w(WB0, c), b1 = p->w(WB1, c), b2 = p->w(WB2, c);
                const double bmag = std::sqrt(b0 * b0 + b1 * b1 + b2 * b2);
                if (bmag > 1e-14) {
                    const double var = (centred(WB0, i, j, k, 0) * h[0] + centred(WB1, i, j, k, 1) * h[1] +
                                        centred(WB2, i, j, k, 2) * h[2]) / bmag;
                    if (var > kRefine && deeper) { v.refine = true; v.erasable = false; }
                }
            }
    if (!deeper) v.refine = false;
    if (p->lvl == 0) v.erasable = false;
    return v;
}

// whether a parent's children can be merged back into it: every child is
// a leaf and none is about to be refined (grid.cpp:817-830)
static bool children_mergeable(const Patch* parent) {
    for (int h = 0; h < 8; h++) {
        const Patch* c = parent->children[h];
        if (c && (!c->leaf || c->refine)) return false;
    }
    return true;
}

// amrtree::regrid, grid.cpp:739-835: verdicts for every leaf, then the
// leaves that asked are refined in quilt order, then for each leaf that may
// be erased (in the quilt order of before the refinement) its parent is
// flattened when all its children can be merged.  The refine/erase marks
// start false and are cleared when acted on; a leaf removed by an earlier
// flatten of the same pass is skipped (the reference reads it after release).
void HostTree::regrid() {
    for (auto& pp : quilt) {
        if (!pp->leaf) continue;
        const LeafVerdict v = leaf_verdict(pp.get());
        if (v.refine) pp->refine = true;
        if (v.erasable) pp->erase = true;
    }
    std::vector<Patch*> grow, shrink;
    for (auto& pp : quilt) {
        if (pp->leaf && pp->refine) grow.push_back(pp.get());
        if (pp->leaf && pp->erase) shrink.push_back(pp.get());
    }
    for (Patch* p : grow) {
        refine(p);
        p->refine = false;
    }
    auto gone = [&](const Patch* p) { return std::find(erased_.begin(), erased_.end(), p) != erased_.end(); };
    for (Patch* p : shrink) {
        if (gone(p)) continue;
        if (p->parent && !p->parent->leaf && children_mergeable(p->parent)) flatten(p->parent);
        if (!gone(p)) p->erase = false;
    }
    erased_.clear();
}

// amrtree::regrid's tree changes (grid.cpp:739-835) with the flags of the
// flagging loop given: the refine pass creates the children in the same
// quilt positions as refine() (no data), the flatten pass removes the same
// children as flatten() (no restriction) and logs what the data work needs.
void HostTree::regrid_structure(const std::vector<unsigned char>& leaf_flags, RegridLog& log) {
    size_t n = 0;
    for (auto& pp : quilt) {
        if (!pp->leaf) continue;
        if (n >= leaf_flags.size()) throw std::runtime_error("regrid flags: fewer flags than leaves");
        const unsigned char f = leaf_flags[n++];
        if (f & 1) pp->refine = true;
        if (f & 2) pp->erase = true;
    }
    if (n != leaf_flags.size()) throw std::runtime_error("regrid flags: more flags than leaves");
    std::vector<Patch*> grow, shrink;
    for (auto& pp : quilt) {
        if (pp->leaf && pp->refine) grow.push_back(pp.get());
        if (pp->leaf && pp->erase) shrink.push_back(pp.get());
    }
    for (Patch* p : grow) {
        append_children(p);
        p->leaf = false;
        p->refine = false;
        log.refined.push_back(p);
    }
    auto gone = [&](const Patch* p) { return std::find(erased_.begin(), erased_.end(), p) != erased_.end(); };
    for (Patch* p : shrink) {
        if (gone(p)) continue;
        Patch* par = p->parent;
        if (par && !par->leaf && children_mergeable(par)) {
            std::array<int, 8> slots;
            for (int h = 0; h < 8; h++) {
                Patch* ch = par->children[h];
                slots[h] = ch ? ch->dev : -1;
                if (!ch) continue;
                erased_.push_back(ch);
                quilt.erase(std::remove_if(quilt.begin(), quilt.end(),
                                           [&](const std::unique_ptr<Patch>& q) { return q.get() == ch; }),
                            quilt.end());
                par->children[h] = nullptr;
            }
            par->leaf = true;
            log.flattened.push_back({par, slots});
        }
        if (!gone(p)) p->erase = false;
    }
    erased_.clear();
}

// ------------------------------------------------------- state exchange
void HostTree::export_blocks(int* meta, double* edges) const {
    for (size_t b = 0; b < quilt.size(); b++) {
        const Patch* p = quilt[b].get();
        int* m = meta + 14 * b;
        m[0] = p->lvl; m[1] = p->ijk[0]; m[2] = p->ijk[1]; m[3] = p->ijk[2];
        m[4] = p->leaf ? 1 : 0;
        m[5] = index_of(p->parent);
        for (int h = 0; h < 8; h++) m[6 + h] = index_of(p->children[h]);
        double* e = edges + 51 * b;
        for (int i = 0; i <= BLK; i++) { e[i] = p->xe[i]; e[17 + i] = p->ye[i]; e[34 + i] = p->ze[i]; }
    }
}
void HostTree::export_state(double* U, double* W, double* Bf) const {
    for (size_t b = 0; b < quilt.size(); b++) {
        const Patch* p = quilt[b].get();
        if (U) std::memcpy(U + b * 8 * NCELL, p->U.data(), sizeof(double) * 8 * NCELL);
        if (W) std::memcpy(W + b * 8 * NCELL, p->W.data(), sizeof(double) * 8 * NCELL);
        if (Bf) std::memcpy(Bf + b * 3 * NFACE, p->Bf.data(), sizeof(double) * 3 * NFACE);
    }
}
void HostTree::import_state(const double* U, const double* W, const double* Bf) {
    for (size_t b = 0; b < quilt.size(); b++) {
        Patch* p = quilt[b].get();
        if (U) std::memcpy(p->U.data(), U + b * 8 * NCELL, sizeof(double) * 8 * NCELL);
        if (W) std::memcpy(p->W.data(), W + b * 8 * NCELL, sizeof(double) * 8 * NCELL);
        if (Bf) std::memcpy(p->Bf.data(), Bf + b * 3 * NFACE, sizeof(double) * 3 * NFACE);
    }
}

}  // namespace bh

// ------------------------------------------------------------ geometry
namespace bh {

void build_geometry_blocks(const HostTree& t, const std::vector<int>& blocks, GeomHost& g) {
    const int nb = (int)blocks.size();
    const TreeParams P = t.prm;
    g.cell_tree.assign((size_t)nb * NGCT * NCELL, 0.0);
    g.cell_cell.assign((size_t)nb * NGCC * NCELL, 0.0);
    g.floors.assign((size_t)nb * 2 * NCELL, 0.0);
    g.face.assign((size_t)nb * 3 * NGFC * NFC, 0.0);
    g.chr.assign((size_t)nb * NCHR * NINT, 0.0);
    g.dxyz.assign((size_t)nb * 3, 0.0);
    g.th.assign((size_t)nb * NCELL, 0.0);
    int bad_chr = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : bad_chr)
    for (int b = 0; b < nb; b++) {
        const Patch& p = *t.quilt[blocks[b]];
        g.dxyz[3 * b] = p.dx(); g.dxyz[3 * b + 1] = p.dy(); g.dxyz[3 * b + 2] = p.dz();
        double* ct = &g.cell_tree[(size_t)b * NGCT * NCELL];
        double* cc = &g.cell_cell[(size_t)b * NGCC * NCELL];
        double* fl = &g.floors[(size_t)b * 2 * NCELL];
        for (int c = 0; c < NCELL; c++) {
            GPt m = gpt_from(metric_comp(P.M, P.a, P.Q, p.r[c], p.th[c]));
            const double vt[NGCT] = {m.sqrtdetgam, m.alpha, m.g11, m.g13, m.g22, m.g33, m.gi13, m.gi33};
            for (int v = 0; v < NGCT; v++) ct[(size_t)v * NCELL + c] = vt[v];
            GCell q = gcell_from(metric_comp(1.0, 0.0, 0.0, p.r[c], p.th[c]));
            const double vc[NGCC] = {q.sqrtdetg, q.alpha, q.beta0, q.g11, q.g22, q.g33,
                                     q.gi00, q.gi01, q.gi11, q.gi22, q.gi33};
            for (int v = 0; v < NGCC; v++) cc[(size_t)v * NCELL + c] = vc[v];
            Prim f = t.pvfs(p.r[c]);
            fl[c] = f.rho;
            fl[NCELL + c] = f.eps;
            g.th[(size_t)b * NCELL + c] = p.th[c];
        }
        // faces: (r, theta) of a face is the mean of the two cells' (grid.cpp:266-269)
        for (int dir = 0; dir < 3; dir++) {
            double* fc = &g.face[((size_t)b * 3 + dir) * NGFC * NFC];
            for (int f = 0; f < NFC; f++) {
                int i, j, k;
                if (dir == 0) { i = f / 256 - 1; j = (f / 16) % 16; k = f % 16; }
                else if (dir == 1) { i = f / 272; j = (f / 16) % 17 - 1; k = f % 16; }
                else { i = f / 272; j = (f / 17) % 16; k = f % 17 - 1; }
                int cL = lindx(i, j, k);
                int cR = lindx(i + (dir == 0), j + (dir == 1), k + (dir == 2));
                double rf = (p.r[cL] + p.r[cR]) / 2;
                double thf = (p.th[cL] + p.th[cR]) / 2;
                GPt m = gpt_from(metric_comp(P.M, P.a, P.Q, rf, thf));
                const double vf[NGFC] = {m.sqrtdetg, m.sqrtdetgam, m.alpha, m.beta0, m.g11, m.g13, m.g22, m.g33};
                for (int v = 0; v < NGFC; v++) fc[(size_t)v * NFC + f] = vf[v];
            }
        }
        // lowered Christoffels of the cell metric on active zones
        // (rk2integrator::source, integrator.cpp:129-137)
        double* ch = &g.chr[(size_t)b * NCHR * NINT];
        for (int i = 0; i < BLK; i++)
            for (int j = 0; j < BLK; j++)
                for (int k = 0; k < BLK; k++) {
                    int c = lindx(i, j, k);
                    int a = (i * BLK + j) * BLK + k;
                    Metric m = metric_comp(1.0, 0.0, 0.0, p.r[c], p.th[c]);
                    for (int mm = 0; mm < 4; mm++)
                        for (int n = 0; n < 4; n++)
                            for (int l = n; l < 4; l++) {
                                double s = 0.0;
                                for (int q = 0; q < 4; q++) s += m.g[mm][q] * m.G[q][n][l];
                                int slot = -1;
                                for (int z = 0; z < NCHR; z++)
                                    if (kChrIdx[z][0] == mm && kChrIdx[z][1] == n && kChrIdx[z][2] == l) slot = z;
                                if (slot >= 0) {
                                    ch[(size_t)slot * NINT + a] = s;
                                } else if (s != 0.0 || std::signbit(s)) {
                                    bad_chr = 1;
                                }
                            }
                }
    }
    if (bad_chr) throw std::runtime_error("a lowered Christoffel of the cell metric expected to be +0.0 is not");
}

void build_geometry(const HostTree& t, GeomHost& g) {
    std::vector<int> all(t.nblocks());
    for (int b = 0; b < t.nblocks(); b++) all[b] = b;
    build_geometry_blocks(t, all, g);
}

void build_ghost_plan(const HostTree& t, GhostPlan& g) {
    const int nb = t.nblocks();
    std::vector<long long> keys(nb);
    auto key = [](int lvl, int i, int j, int k) {
        return ((long long)lvl << 54) ^ ((long long)(i & 0x3ffff) << 36) ^ ((long long)(j & 0x3ffff) << 18) ^
               (long long)(k & 0x3ffff);
    };
    // first block in quilt order with a given (level, ijk), as the linear scan finds it
    std::vector<std::pair<long long, int>> idx;
    idx.reserve(nb);
    for (int b = 0; b < nb; b++) {
        const Patch& p = *t.quilt[b];
        idx.push_back({key(p.lvl, p.ijk[0], p.ijk[1], p.ijk[2]), b});
    }
    std::stable_sort(idx.begin(), idx.end(), [](auto& x, auto& y) { return x.first < y.first; });
    auto find = [&](int lvl, int i, int j, int k) -> int {
        long long kk = key(lvl, i, j, k);
        auto it = std::lower_bound(idx.begin(), idx.end(), std::make_pair(kk, -1),
                                   [](auto& x, auto& y) { return x.first < y.first; });
        if (it == idx.end() || it->first != kk) return -1;
        const Patch& q = *t.quilt[it->second];
        if (q.lvl != lvl || q.ijk[0] != i || q.ijk[1] != j || q.ijk[2] != k) return -1;
        return it->second;
    };
    g.leaves.clear();
    g.tab.assign((size_t)nb * 12, -1);
    std::vector<std::vector<int>> pred(nb);
    for (int b = 0; b < nb; b++) {
        const Patch& p = *t.quilt[b];
        if (!p.leaf) continue;
        g.leaves.push_back(b);
        for (int dim = 0; dim < 3; dim++)
            for (int s = 0; s < 2; s++) {
                int side = s ? 1 : -1;
                int n[3] = {p.ijk[0], p.ijk[1], p.ijk[2]};
                n[dim] += side;
                int q = find(p.lvl, n[0], n[1], n[2]);
                int kind = GH_COPY;
                if (q < 0 && p.lvl > 0) {
                    int c[3] = {p.ijk[0] / 2, p.ijk[1] / 2, p.ijk[2] / 2};
                    c[dim] += side;
                    q = find(p.lvl - 1, c[0], c[1], c[2]);
                    kind = GH_PROLONG;
                }
                if (q < 0) kind = GH_BNDY;
                g.tab[(size_t)b * 12 + (2 * dim + s) * 2] = kind;
                g.tab[(size_t)b * 12 + (2 * dim + s) * 2 + 1] = q;
                if (kind == GH_PROLONG && t.quilt[q]->leaf) {
                    if (q < b) pred[b].push_back(q);
                    else pred[q].push_back(b);
                }
            }
    }
    std::vector<int> wave(nb, 0);
    int nw = 0;
    for (int b : g.leaves) {
        int w = 0;
        for (int q : pred[b]) w = std::max(w, wave[q] + 1);
        wave[b] = w;
        nw = std::max(nw, w + 1);
    }
    g.waves.assign(nw, {});
    for (int b : g.leaves) g.waves[wave[b]].push_back(b);
    g.same.assign((size_t)nb * 6, 0);
}

namespace {
int cell_on_axis(int dim, int a, int t1, int t2) {
    return dim == 0 ? lindx(a, t1, t2) : dim == 1 ? lindx(t1, a, t2) : lindx(t1, t2, a);
}
}  // namespace

// A copied ghost cell (grid.cpp:390-425) holds its source's U; its con2prim
// sees the same U and, when the two cells' metric bits agree, returns the
// source's result.  Cell centres on either side of a same-level face are
// normally computed to the same bits; this checks it for every such pair.
void mark_exact_copies(const GeomHost& geo, GhostPlan& g) {
    const size_t nb = g.tab.size() / 12;
    g.same.assign(nb * 6, 0);
    for (int b : g.leaves)
        for (int side = 0; side < 6; side++) {
            const int kind = g.tab[(size_t)b * 12 + side * 2], q = g.tab[(size_t)b * 12 + side * 2 + 1];
            if (kind != GH_COPY) continue;
            const int dim = side / 2, s = side % 2;
            bool eq = true;
            for (int gl = 0; gl < GH && eq; gl++)
                for (int t1 = 0; t1 < BLK && eq; t1++)
                    for (int t2 = 0; t2 < BLK && eq; t2++) {
                        const int c = cell_on_axis(dim, s ? BLK + gl : gl - GH, t1, t2);
                        const int sc = cell_on_axis(dim, (s ? 0 : BLK - GH) + gl, t1, t2);
                        for (int v = 0; v < NGCT; v++) {
                            const double x = geo.cell_tree[((size_t)b * NGCT + v) * NCELL + c];
                            const double y = geo.cell_tree[((size_t)q * NGCT + v) * NCELL + sc];
                            if (std::memcmp(&x, &y, sizeof x) != 0) { eq = false; break; }
                        }
                    }
            g.same[(size_t)b * 6 + side] = eq ? 1 : 0;
        }
}

}  // namespace bh
