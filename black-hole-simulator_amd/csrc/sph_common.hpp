// Layout, tables and run records of the spherical Kerr-Schild engine: the
// GRMHD step of src/fluid (integrator.cpp:191-275) on blocks of a
// logarithmic-r x theta x phi grid, with the physics made self-consistent
// ("consistent" preset; see sph_physics.hpp).
//
// A block holds 16 x 16 x BZ active zones (r, theta, phi) with two ghost
// layers per side; BZ = 16 for three-dimensional runs, 1 for axisymmetric
// ones (the phi ghosts of an axisymmetric block are copies of its single
// layer).  Cell index = ((k+2)*SY + (j+2))*SX + (i+2): r fastest, phi
// slowest, so that both shapes share the same inner layout.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SPH_HD __host__ __device__ __forceinline__
#else
#define SPH_HD inline
#endif

namespace sph {

using std::isfinite;
using std::sqrt;

constexpr int BX = 16, BY = 16;  // active zones per block in r and theta
constexpr int G = 2;             // ghost layers
constexpr int SX = BX + 2 * G;   // 20
constexpr int SY = BY + 2 * G;   // 20

// State components.  U (densitised by sqrt(gamma)): D, S_r, S_theta, S_phi,
// tau, B^r, B^theta, B^phi.  W: rho, eps, u~^r, u~^theta, u~^phi (u~^i = W v^i,
// the spatial velocity times the Lorentz factor), B^r, B^theta, B^phi.
enum { UD = 0, US1 = 1, US2 = 2, US3 = 3, UTAU = 4, UB1 = 5, UB2 = 6, UB3 = 7 };
enum { WRHO = 0, WEPS = 1, WU1 = 2, WU2 = 3, WU3 = 4, WB1 = 5, WB2 = 6, WB3 = 7 };
// Flux words per face: D, S_1..3, tau, then the two transverse field fluxes
// in cyclic order (x-face: B^y, B^z; y-face: B^z, B^x; z-face: B^x, B^y).
constexpr int NFW = 7;

struct Lay {
    int bz, sz;     // active zones and stored cells in phi
    int ncell;      // stored cells per block
    int nint;       // active zones per block
    int nf;         // padded faces per direction (max of the three index spaces)
    int nff;        // flux words per direction and word (ffidx below)
};
SPH_HD Lay make_lay(int bz) {
    Lay L;
    L.bz = bz;
    L.sz = bz + 2 * G;
    L.ncell = SX * SY * L.sz;
    L.nint = BX * BY * bz;
    const int a = (SX + 1) * SY * L.sz, b = SX * SY * (L.sz + 1);
    L.nf = a > b ? a : b;
    // own faces plus four side slabs (ffidx): r and theta directions
    // 17 x 16 x bz + 2 (17 bz + 272), phi 256 (bz + 1) + 64 (bz + 1)
    const int f01 = (BX + 1) * BY * bz + 2 * (BX + 1) * bz + 2 * (BX + 1) * BY;
    const int f2 = BX * BY * (bz + 1) + 2 * BY * (bz + 1) + 2 * BX * (bz + 1);
    L.nff = f01 > f2 ? f01 : f2;
    return L;
}
SPH_HD int cidx(int i, int j, int k) { return ((k + G) * SY + (j + G)) * SX + (i + G); }
// Face index spaces: face d of cell (i, j, k) sits at its low side in d;
// x faces i in [-2, 18], y faces j in [-2, 18], z faces k in [-2, bz + 2].
SPH_HD int fidx(int d, int i, int j, int k) {
    return d == 0 ? ((k + G) * SY + (j + G)) * (SX + 1) + (i + G)
         : d == 1 ? ((k + G) * (SY + 1) + (j + G)) * SX + (i + G)
                  : ((k + G) * SY + (j + G)) * SX + (i + G);
}

// Flux words of a block, per direction d: the own faces (coordinate along d
// in [0, n_d], the two transverse ones inside the block) as a dense box with
// i fastest, then four slabs of side faces (one transverse coordinate at -1
// or n_e, the other inside), ordered by transverse direction e and low/high
// side, each slab dense in its two remaining coordinates (the less
// significant fastest, as the face lists order them).  Corner faces (both
// transverse coordinates outside) are never formed and have no word.  Every
// word of a slot is written by the flux kernel (own faces always, a side
// slab whenever the block does not leave that side to its neighbour), so
// full cache lines reach memory.
SPH_HD int ffown(int d, int i, int j, int k, int bz) {
    (void)bz;
    return d == 0 ? (k * BY + j) * (BX + 1) + i : d == 1 ? (k * (BY + 1) + j) * BX + i : (k * BY + j) * BX + i;
}
SPH_HD int ffidx(int d, int i, int j, int k, int bz) {
    const int nx = BX + (d == 0), ny = BY + (d == 1), nz = bz + (d == 2);
    const int own = nx * ny * nz;
    // the transverse coordinate outside, if any
    if (d != 0 && (i < 0 || i >= BX)) {
        // e = 0: j (ny) fastest, then k (nz)
        const int h = i < 0 ? 0 : 1;
        return own + h * ny * nz + k * ny + j;
    }
    const int s0 = d != 0 ? 2 * ny * nz : 0;  // the e = 0 slabs
    if (d != 1 && (j < 0 || j >= BY)) {
        // e = 1: i (nx) fastest, then k (nz)
        const int h = j < 0 ? 0 : 1;
        return own + s0 + h * nx * nz + k * nx + i;
    }
    const int s1 = s0 + (d != 1 ? 2 * nx * nz : 0);
    if (d != 2 && (k < 0 || k >= bz)) {
        // e = 2: i (nx) fastest, then j (ny)
        const int h = k < 0 ? 0 : 1;
        return own + s1 + h * nx * ny + j * nx + i;
    }
    return (k * ny + j) * nx + i;
}

// Geometry tables, per (run, level), over the points (ii, jj) of the level's
// r x theta grid with two ghost points on each side and one more for the
// last face: point = (jj + 2) * (nr + 5) + (ii + 2).
// Metric words (cell centres and faces):
enum {
    M_ALPHA = 0, M_B1 = 1, M_B2 = 2, M_B3 = 3,                       // lapse, shift beta^i
    M_G11 = 4, M_G12 = 5, M_G13 = 6, M_G22 = 7, M_G23 = 8, M_G33 = 9,   // gamma_ij
    M_I11 = 10, M_I12 = 11, M_I13 = 12, M_I22 = 13, M_I23 = 14, M_I33 = 15,  // gamma^ij
    M_SG = 16,                                                       // sqrt(gamma)
    NMET = 17
};
// Cell-centre words after the metric: d_r g_mn and d_theta g_mn (10 each, in
// the order 00 01 02 03 11 12 13 22 23 33), Gamma^0_mn (10), d_r alpha,
// d_theta alpha, the floors rho_fl(r), eps_fl(r), then r, theta, sin theta
// and cos theta of the centre.
enum {
    C_DG1 = NMET, C_DG2 = NMET + 10, C_GAM0 = NMET + 20, C_DA1 = NMET + 30, C_DA2 = NMET + 31,
    C_RFL = NMET + 32, C_EFL = NMET + 33, C_R = NMET + 34, C_TH = NMET + 35, C_SIN = NMET + 36,
    C_COS = NMET + 37, NGC = NMET + 38
};

// One level of one run.
struct RL {
    int run, level;
    int nr, nth, nph;        // zones of the level
    int nbr, nbth, nbph;     // blocks of the level
    int coarser, finer;      // RL index of the next coarser / finer level, -1
    int lut;                 // offset of the level's block lookup (nbr*nbth*nbph ints; -1 = no leaf)
    long long gc;            // offset of the cell-centre table (NGC words x npt)
    long long gf;            // offset of the r-face and theta-face tables (2 x NMET words x npt)
    int dr;                  // offset of the r-spacing table (nr + 5 words, ii = -2 ..)
    int npt;                 // (nr + 5) * (nth + 5)
    int zone0;               // index of the level's first zone in the run's perturbation stream
    double dth, dph;         // theta and phi spacing
};
SPH_HD int gpt(const RL& L, int ii, int jj) { return (jj + 2) * (L.nr + 5) + (ii + 2); }

// Physics of one run.
struct RunPar {
    double a, Q, gam, cfl;
    double rho_fl, eps_fl;   // floor coefficients: rho_fl * r^-3/2, eps_fl * r^-3/2
    double sigma_max;        // b^2 / rho ceiling
    double w_max;            // Lorentz-factor ceiling
    int nlev, rl0;           // levels and the RL index of level 0
    int nz;                  // 1: axisymmetric (no phi CFL), else BZ
    int riemann;             // 0: HLL, 1: HLLD
};

// Block record: its (run, level) and block coordinates at that level.
struct Blk {
    int rl, bi, bj, bk;
};

// Per-run counters and the step's dt, reset each step.
struct RunCtr {
    unsigned long long dtbits;   // min over zones of the CFL dt, as an ordered bit pattern
    unsigned long long c2p_fail; // con2prim failures (floor state applied)
    unsigned long long floors;   // zones floored (rho, eps, sigma or Lorentz ceiling)
    unsigned long long pad;
};

}  // namespace sph
