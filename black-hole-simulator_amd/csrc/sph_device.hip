// The spherical Kerr-Schild engine on gfx950: the GRMHD step of src/fluid
// (rk2integrator::step, src/fluid/solver/integrator.cpp:262-275) for any
// number of independent runs (tori of their own grid, spin and field) whose
// blocks of every refinement level are advanced together in shared launches.
//
// One step = begin (per-run dt reset) and two RK stages; a stage is
//   A  con2prim of every active zone with floors (integrator.cpp:201-222,
//      grid.cpp:331-358), CFL reduction per run in the first stage
//      (integrator.cpp:156-189)
//   -- the only point where blocks must wait for each other --
//   B  ghost fill of primitives and face fields from same-level, coarser or
//      finer leaves and the domain boundary (grid.cpp:390-571), PLM + HLL
//      fluxes on every face a block needs (grid.cpp:255-329), constrained
//      transport with the EMFs formed in place (transport.cpp:11-58), flux
//      divergence + geometric sources + RK2 and the cell field from the new
//      faces (integrator.cpp:18-153, 238-259, transport.cpp:60-75).
// Phase B reads other blocks only through primitives (written in A) and
// face fields of the previous stage (double-buffered), so every block of
// phase B is independent.  Two launch schedules give identical results:
// one kernel per sub-phase over all blocks, or one persistent kernel per
// stage whose workgroups pull blocks from a queue and cross one grid
// barrier between A and B.
//
// State is structure-of-arrays per block slot; fluxes are FP64 or (mixed
// precision) FP32 words.  Reconstruction, Riemann fluxes and EMFs run in the
// flux type; con2prim, sources, the update and all reductions in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/bhsph.h"
#include "bh_common.hpp"
#include "bh_crmath.hpp"
#include "bh_crtrig.hpp"
#include "sph_common.hpp"
#include "sph_geom.hpp"
#include "sph_physics.hpp"
#include "sph_torus.hpp"
#include "sph_ffloat.hpp"

namespace sph {

#define SCHK(x)                                                                                    \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct DevCR {
    __device__ static double log(double x) { return bh::cr_log(x); }
    __device__ static double exp(double x) { return bh::cr_exp(x); }
    __device__ static double pow(double x, double y) { return bh::cr_pow(x, y); }
};

// one level of one run, with its device tables
struct DRL {
    int run, level, nr, nth, nph, nbr, nbth, nbph;
    int coarser, finer;
    int npt;
    int ih;               // r index of the zone row holding the outer horizon (-1: none)
    long long zone0;
    double dth, dph;
    const int* lut;
    const double* gc;
    const double* gf;
    const double* dr;
};

// Table pointers reached through device memory (DRL records) are generic
// pointers to the compiler, which then emits flat loads (counted against
// both the vector-memory and the LDS/scalar wait counters).  The tables
// live in device global memory, so their loads are issued as global loads.
#define GLB __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ const GLB T* gl(const T* p) {
    return (const GLB T*)p;
}

struct KPS {
    int bz, sz, ncell, nint, nf, nff;
    const Blk* blk;       // [slot]
    const DRL* rl;        // [rl]
    const RunPar* run;    // [run]
    RunCtr* ctr;          // [run]
    double* U;            // [slot][8][ncell]
    double* W;            // [slot][8][ncell]
    double* Us;           // [slot][5][nint]
    double* Bf0;          // [slot][3][nf]
    double* Bf1;
    void* F;              // [slot][3][NFW][nf] flux words (double or float)
    const int* list;      // block slots to process
    int nlist;
    const int* ghc;       // ghost cells of a block: packed (k+2)<<16 | (j+2)<<8 | (i+2)
    int nghc;
    const int* ghf;       // ghost faces: d<<24 | packed cell
    int nghf;
    // the ghost cells, then the side faces (bit 30 set), that a block whose
    // shared sides form mask m fills: ghp[ghp_off[m] .. ghp_off[m + 1])
    const int* ghp;
    const int* ghp_off;
    const int* flf;       // flux faces
    int nflf;
    const int* owf;       // own faces
    int nowf;
    const int* flp;       // theta / phi face lines (bit 26 set), then r faces
    int nflp;
    // persistent schedule
    unsigned int* qctr;   // [0] stage-A queue, [1] barrier arrivals, [2] stage-B queue, [4..12] queue schedule,
                          // [15] barrier timeout flag
    unsigned int spin;    // polls of a grid-barrier wait before it gives up
    int xcd;              // 1: renumber workgroups so that each XCD takes a contiguous range
    // Side faces: a block computes fluxes on the faces one layer outside it
    // (for the EMFs of its boundary edges).  The flux of a face depends only
    // on its level and global position, so where the block across that
    // side is a leaf of the same level, which computes the same face as its
    // own, the block skips it and the EMFs read the neighbour's word.
    const int* side;      // [slot][6]: same-level leaf across r-, r+, theta-, theta+, phi-, phi+ (-1: none)
    const int* sgrp;      // [3 directions][6 sides][2]: offset and count of those side faces in flf
    int reuse;            // 1: skip and share side faces (the sub-phase schedule)
};

// Workgroups are dispatched round-robin over the eight XCDs (workgroup x to
// XCD x mod 8), each with its own L2.  With xcd = 1 the kernels that spread a
// block over several workgroups use the logical index this returns, so that
// the workgroups of one block run on one XCD and share its L2.
constexpr int NXCD = 8;
__device__ __forceinline__ int wg_index(const KPS& P) {
    const int x = blockIdx.x;
    if (!P.xcd) return x;
    const int n = gridDim.x;
    const int q = n / NXCD, r = n % NXCD;
    const int k = x % NXCD, loc = x / NXCD;
    return k * q + (k < r ? k : r) + loc;
}

// the side (0..5) a flux face of direction d lies outside the block on, or -1
__host__ __device__ __forceinline__ int face_side_bz(int d, int i, int j, int k, int bz) {
    if (d != 0) { if (i < 0) return 0; if (i >= BX) return 1; }
    if (d != 1) { if (j < 0) return 2; if (j >= BY) return 3; }
    if (d != 2) { if (k < 0) return 4; if (k >= bz) return 5; }
    return -1;
}
__device__ __forceinline__ int face_side_of(const KPS& P, int d, int i, int j, int k) {
    return face_side_bz(d, i, j, k, P.bz);
}
// whether any flux stencil reads ghost cell (i, j, k) of a block whose
// shared sides are the set bits of mask (bit e: side e): corner cells
// (three coordinates outside) never; a slab cell only where its side is not
// shared; an edge cell (two outside) only where one of its two sides is not
// shared (the side faces computed there read it)
__host__ __device__ __forceinline__ bool ghost_read_mask(int mask, int i, int j, int k, int bz) {
    const int si = i < 0 ? 0 : (i >= BX ? 1 : -1);
    const int sj = j < 0 ? 2 : (j >= BY ? 3 : -1);
    const int sk = k < 0 ? 4 : (k >= bz ? 5 : -1);
    const int nout = (si >= 0) + (sj >= 0) + (sk >= 0);
    if (nout == 3) return false;
    bool all = true;
    if (si >= 0) all = all && ((mask >> si) & 1);
    if (sj >= 0) all = all && ((mask >> sj) & 1);
    if (sk >= 0) all = all && ((mask >> sk) & 1);
    return !all;
}
// true where block b leaves face (d; i, j, k) to its same-level neighbour
__device__ __forceinline__ bool face_shared(const KPS& P, int b, int d, int i, int j, int k) {
    if (!P.reuse) return false;
    const int sd = face_side_of(P, d, i, j, k);
    return sd >= 0 && P.side[(size_t)b * 6 + sd] >= 0;
}

// The same-level leaves across the six sides of a block, where the block
// shares its side faces with them (else -1).
struct SideNb {
    int s[6];
};
__device__ __forceinline__ SideNb side_nb(const KPS& P, int b) {
    SideNb n;
#pragma unroll
    for (int e = 0; e < 6; e++) n.s[e] = P.reuse ? P.side[(size_t)b * 6 + e] : -1;
    return n;
}
// the shared sides of a block with sides n as a 6-bit mask
__device__ __forceinline__ int side_mask(const SideNb& n) {
    int m = 0;
#pragma unroll
    for (int e = 0; e < 6; e++) m |= (n.s[e] >= 0 ? 1 : 0) << e;
    return m;
}
__device__ __forceinline__ bool ghost_read(const KPS& P, const SideNb& n, int i, int j, int k) {
    return ghost_read_mask(side_mask(n), i, j, k, P.bz);
}

__device__ __forceinline__ double dtbits_to(unsigned long long b) { return __longlong_as_double((long long)b); }

__device__ __forceinline__ int leaf_at(const KPS& P, const DRL& v, int gi, int gj, int gk) {
    const int b1 = gi / BX, b2 = gj / BY, b3 = P.bz == 1 ? gk : gk / BX;  // bz is 16 or 1
    return gl(v.lut)[((size_t)b3 * v.nbth + b2) * v.nbr + b1];
}
__device__ __forceinline__ void bc_cell(const DRL& v, int& gi, int& gj, int& gk, int& flip) {
    if (gj < 0) { gj = -1 - gj; flip ^= 1; }
    if (gj >= v.nth) { gj = 2 * v.nth - 1 - gj; flip ^= 1; }
    if (gi < 0) gi = 0;
    if (gi >= v.nr) gi = v.nr - 1;
    if (gk < 0 || gk >= v.nph) gk = ((gk % v.nph) + v.nph) % v.nph;  // periodic in phi
}
// a source cell of a ghost fill: its slot, its cell index there, and
// whether u~^theta and B^theta change sign (reflection at a pole)
struct Src {
    int s, c, flip;
};
__device__ __forceinline__ void src_at(const KPS& P, int s, int gi, int gj, int gk, int flip, Src& o) {
    const Blk B = P.blk[s];
    o.s = s;
    o.c = cidx(gi - B.bi * BX, gj - B.bj * BY, gk - B.bk * P.bz);
    o.flip = flip;
}
__device__ __forceinline__ bool leaf_src(const KPS& P, const DRL& v, int gi, int gj, int gk, Src& o) {
    int flip = 0;
    bc_cell(v, gi, gj, gk, flip);
    const int s = leaf_at(P, v, gi, gj, gk);
    if (s < 0) return false;
    src_at(P, s, gi, gj, gk, flip, o);
    return true;
}
__device__ __forceinline__ double w_of(const KPS& P, const double* __restrict__ Wr, const Src& x, int q) {
    const double w = Wr[((size_t)x.s * 8 + q) * P.ncell + x.c];
    return (x.flip && (q == WU2 || q == WB2)) ? -w : w;
}

// ghost cell (i, j, k) of block b.  The source cells are located first,
// then the eight words are formed and stored one at a time.  Sources are
// active zones and the destination a ghost cell, so the reads (Wr) and the
// writes (Ww) of one call never meet, although both point into W: declared
// as distinct, the loads need not wait for earlier stores.
__device__ __forceinline__ void ghost_cell_rw(const KPS& P, const double* __restrict__ Wr, double* __restrict__ Ww,
                                              int b, int i, int j, int k) {
    const Blk B = P.blk[b];
    const DRL& v = P.rl[B.rl];
    int gi = B.bi * BX + i, gj = B.bj * BY + j, gk = B.bk * P.bz + k, flip = 0;
    bc_cell(v, gi, gj, gk, flip);
    double* Wc = Ww + (size_t)b * 8 * P.ncell + cidx(i, j, k);
    auto put = [&](int q, double x) { Wc[(size_t)q * P.ncell] = (flip && (q == WU2 || q == WB2)) ? -x : x; };
    const int s = leaf_at(P, v, gi, gj, gk);
    if (s >= 0) {
        // same level: copy
        Src x;
        src_at(P, s, gi, gj, gk, 0, x);
#pragma unroll
        for (int q = 0; q < 8; q++) put(q, w_of(P, Wr, x, q));
    } else {
        bool done = false;
        if (v.coarser >= 0) {
            // coarser leaf: its value plus minmod-limited slopes in each
            // direction (zero where a neighbour is not a leaf of that level)
            const DRL& vc = P.rl[v.coarser];
            const int c1 = gi >> 1, c2 = gj >> 1, c3 = gk >> 1;
            Src xc;
            if (leaf_src(P, vc, c1, c2, c3, xc)) {
                Src xm[3], xp[3];
                bool h[3];
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    const int o0 = d == 0, o1 = d == 1, o2 = d == 2;
                    const bool hm = leaf_src(P, vc, c1 - o0, c2 - o1, c3 - o2, xm[d]);
                    const bool hp = leaf_src(P, vc, c1 + o0, c2 + o1, c3 + o2, xp[d]);
                    h[d] = hm && hp;
                }
                const double off[3] = {(gi & 1) ? 0.25 : -0.25, (gj & 1) ? 0.25 : -0.25, (gk & 1) ? 0.25 : -0.25};
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const double wc = w_of(P, Wr, xc, q);
                    double acc = wc;
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        const double sl = h[d] ? minmod(wc - w_of(P, Wr, xm[d], q), w_of(P, Wr, xp[d], q) - wc) : 0.0;
                        acc = acc + sl * off[d];
                    }
                    put(q, acc);
                }
                done = true;
            }
        }
        if (!done) {
            // finer leaves: the mean of the eight cells covering this one
            const DRL& vf = P.rl[v.finer];
            Src xf[8];
#pragma unroll
            for (int n = 0; n < 8; n++) {
                const int fi = 2 * gi + (n & 1), fj = 2 * gj + ((n >> 1) & 1), fk = 2 * gk + (n >> 2);
                src_at(P, leaf_at(P, vf, fi, fj, fk), fi, fj, fk, 0, xf[n]);
            }
#pragma unroll
            for (int q = 0; q < 8; q++) {
                double acc = 0.0;
#pragma unroll
                for (int n = 0; n < 8; n++) acc += w_of(P, Wr, xf[n], q);
                put(q, acc * 0.125);
            }
        }
    }
}
__device__ void ghost_cell(const KPS& P, int b, int i, int j, int k) { ghost_cell_rw(P, P.W, P.W, b, i, j, k); }

__device__ __forceinline__ void bc_face(const DRL& v, int d, int& gi, int& gj, int& gk, double& sgn) {
    if (d == 1) {
        if (gj < 0) { gj = -gj; sgn = -sgn; }
        if (gj > v.nth) { gj = 2 * v.nth - gj; sgn = -sgn; }
    } else {
        if (gj < 0) gj = -1 - gj;
        if (gj >= v.nth) gj = 2 * v.nth - 1 - gj;
    }
    if (d != 0) {
        if (gi < 0) gi = 0;
        if (gi >= v.nr) gi = v.nr - 1;
    }
    if (gk < 0 || gk >= v.nph) gk = ((gk % v.nph) + v.nph) % v.nph;  // periodic in phi
}
// the stored field of face (d; gi, gj, gk) at level v from a leaf that has
// it as an own face: the leaf of the cell on the face's high side, else the
// one on its low side; false if neither is a leaf
__device__ __forceinline__ bool face_val(const KPS& P, const double* Bf, const DRL& v, int d, int gi, int gj, int gk,
                                         double& out) {
    const int N = d == 0 ? v.nr : d == 1 ? v.nth : v.nph;
    const int gd = d == 0 ? gi : d == 1 ? gj : gk;
    for (int side = 0; side < 2; side++) {
        const int cd = gd - side;
        if (cd < 0 || cd >= N) continue;
        const int s = leaf_at(P, v, d == 0 ? cd : gi, d == 1 ? cd : gj, d == 2 ? cd : gk);
        if (s < 0) continue;
        const Blk B = P.blk[s];
        out = Bf[((size_t)s * 3 + d) * P.nf + fidx(d, gi - B.bi * BX, gj - B.bj * BY, gk - B.bk * P.bz)];
        return true;
    }
    return false;
}

__device__ void ghost_face(const KPS& P, double* Bf, int b, int d, int i, int j, int k) {
    const Blk B = P.blk[b];
    const DRL& v = P.rl[B.rl];
    int gi = B.bi * BX + i, gj = B.bj * BY + j, gk = B.bk * P.bz + k;
    double sgn = 1.0;
    bc_face(v, d, gi, gj, gk, sgn);
    double x, res;
    if (face_val(P, Bf, v, d, gi, gj, gk, x)) {
        res = sgn * x;
    } else {
        const int g[3] = {gi, gj, gk};
        bool done = false;
        if (v.coarser >= 0) {
            const DRL& vc = P.rl[v.coarser];
            int c[3] = {g[0] >> 1, g[1] >> 1, g[2] >> 1};
            int cc[3] = {c[0], c[1], c[2]};
            const int nd = d == 0 ? vc.nr : d == 1 ? vc.nth : vc.nph;
            if (cc[d] == nd) cc[d] -= 1;
            if (leaf_at(P, vc, cc[0], cc[1], cc[2]) >= 0) {
                if ((g[d] & 1) == 0) {
                    c[d] = g[d] >> 1;
                    double y;
                    face_val(P, Bf, vc, d, c[0], c[1], c[2], y);
                    res = sgn * y;
                } else {
                    int lo[3] = {c[0], c[1], c[2]}, hi[3] = {c[0], c[1], c[2]};
                    lo[d] = (g[d] - 1) >> 1;
                    hi[d] = (g[d] + 1) >> 1;
                    double y0, y1;
                    face_val(P, Bf, vc, d, lo[0], lo[1], lo[2], y0);
                    face_val(P, Bf, vc, d, hi[0], hi[1], hi[2], y1);
                    res = sgn * (0.5 * (y0 + y1));
                }
                done = true;
            }
        }
        if (!done) {
            const DRL& vf = P.rl[v.finer];
            const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
            double acc = 0.0;
            for (int bb = 0; bb < 2; bb++)
                for (int a = 0; a < 2; a++) {
                    int f[3] = {2 * g[0], 2 * g[1], 2 * g[2]};
                    f[t1] += a;
                    f[t2] += bb;
                    double y;
                    face_val(P, Bf, vf, d, f[0], f[1], f[2], y);
                    acc += y;
                }
            res = sgn * (acc * 0.25);
        }
    }
    Bf[((size_t)b * 3 + d) * P.nf + fidx(d, i, j, k)] = res;
}

__device__ __forceinline__ void unpack(int x, int& i, int& j, int& k) {
    i = (x & 0xff) - 2;
    j = ((x >> 8) & 0xff) - 2;
    k = ((x >> 16) & 0xff) - 2;
}

// ---------------------------------------------------------------- phase A
// con2prim + floors of one active zone; returns the zone's dt candidate
// (before the CFL factor) when want_dt, else 1e300.
__device__ double c2p_zone(const KPS& P, int b, int a, bool want_dt, int& nfail, int& nfl) {
    const Blk B = P.blk[b];
    const DRL& v = P.rl[B.rl];
    const RunPar& R = P.run[v.run];
    const int i = a % BX, j = (a / BX) % BY, k = a / (BX * BY);
    const int gi = B.bi * BX + i, gj = B.bj * BY + j;
    const int c = cidx(i, j, k);
    const int p = (gj + 2) * (v.nr + 5) + (gi + 2);
    const Met<double> m = load_met<double>(gl(v.gc), v.npt, p);
    double* Ub = P.U + (size_t)b * 8 * P.ncell;
    double* Wb = P.W + (size_t)b * 8 * P.ncell;
    double u[8], w[8];
#pragma unroll
    for (int q = 0; q < 8; q++) u[q] = Ub[(size_t)q * P.ncell + c];
    const double rfl = gl(v.gc)[(size_t)C_RFL * v.npt + p], efl = gl(v.gc)[(size_t)C_EFL * v.npt + p];
    // con2prim; a failed inversion (cold inversion or, for an unrecoverable
    // state, the floor state) always rewrites U, floors where they change W
    const int rc = c2p(m, R.gam, u, w);
    if (rc == 2) floor_state(rfl, efl, w);
    nfail += rc != 0;
    const int fl = apply_floors(m, R.gam, rfl, efl, R.sigma_max, R.w_max, w);
    nfl += fl;
    const bool redo = rc != 0 || fl != 0;
    if (redo) {
        double un[8];
        ptoc(m, R.gam, w, un);
#pragma unroll
        for (int q = 0; q < 5; q++) Ub[(size_t)q * P.ncell + c] = un[q];
    }
#pragma unroll
    for (int q = 0; q < 8; q++) Wb[(size_t)q * P.ncell + c] = w[q];
    if (!want_dt) return 1e300;
    const double dx[3] = {gl(v.dr)[gi + 2], v.dth, v.dph};
    return cell_dt(m, R.gam, w, dx, P.bz == 1 ? 2 : 3);
}

// workgroup-wide min of a positive double and sums of two counters, then one
// atomic per workgroup into the run's counters
template <int TPB>
__device__ void wg_commit(const KPS& P, int run, double dtc, int nfail, int nfl, bool want_dt) {
    __shared__ unsigned long long s_dt[TPB / 64];
    __shared__ int s_f[TPB / 64], s_l[TPB / 64];
    unsigned long long bits = (unsigned long long)__double_as_longlong(dtc);
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long ob = __shfl_xor(bits, o);
        bits = ob < bits ? ob : bits;
        nfail += __shfl_xor(nfail, o);
        nfl += __shfl_xor(nfl, o);
    }
    const int wv = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) { s_dt[wv] = bits; s_f[wv] = nfail; s_l[wv] = nfl; }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long mb = s_dt[0];
        int f = s_f[0], l = s_l[0];
        for (int q = 1; q < TPB / 64; q++) {
            mb = s_dt[q] < mb ? s_dt[q] : mb;
            f += s_f[q];
            l += s_l[q];
        }
        if (want_dt) {
            // the CFL factor is applied to the minimum (exact: multiplication
            // by a positive constant preserves order)
            const double dt = P.run[run].cfl * __longlong_as_double((long long)mb);
            atomicMin(&P.ctr[run].dtbits, (unsigned long long)__double_as_longlong(dt));
        }
        if (f) atomicAdd(&P.ctr[run].c2p_fail, (unsigned long long)f);
        if (l) atomicAdd(&P.ctr[run].floors, (unsigned long long)l);
    }
    __syncthreads();
}

template <int TPB>
__device__ void phase_a_block(const KPS& P, int b, bool want_dt, int h = 0, int parts = 1) {
    double mn = 1e300;
    int nfail = 0, nfl = 0;
    for (int a = h * TPB + threadIdx.x; a < P.nint; a += parts * TPB) mn = fmin(mn, c2p_zone(P, b, a, want_dt, nfail, nfl));
    wg_commit<TPB>(P, P.rl[P.blk[b].rl].run, mn, nfail, nfl, want_dt);
}

// ---------------------------------------------------------------- phase B
// the face of direction D (a template parameter: every state index is then
// static and the states stay in registers).  RS = 1: every run of the launch
// uses the HLLD solver, and only its code is compiled in (the HLL branch of
// RS = 0, which follows the run's setting, raises the register count by the
// values the two branches would share)
template <class Rt, int D, int RS = 0>
__device__ __forceinline__ void flux_face_val(const KPS& P, const double* Bf, int b, int i, int j, int k,
                                              Rt out[NFW]) {
    constexpr int d = D;
    const Blk B = P.blk[b];
    const DRL& v = P.rl[B.rl];
    const RunPar& R = P.run[v.run];
    const bool use_hlld = RS == 1 || R.riemann;
    const int gi = B.bi * BX + i, gj = B.bj * BY + j;
    const int p = (gj + 2) * (v.nr + 5) + (gi + 2);
    const GLB double* tab = gl(d == 0 ? v.gf : d == 1 ? v.gf + (size_t)NMET * v.npt : v.gc);
    // the face metric: of gamma^ij only gamma^nn enters the flux
    auto face_met = [&]() {
        Met<Rt> m;
        const long long np = v.npt;
        m.alpha = (Rt)tab[M_ALPHA * np + p];
        m.beta[0] = (Rt)tab[M_B1 * np + p];
        m.beta[1] = m.beta[2] = Rt(0);
#pragma unroll
        for (int q = 0; q < 6; q++) {
            m.g[q] = (q == 1 || q == 4) ? Rt(0) : (Rt)tab[(M_G11 + q) * np + p];
            m.gi[q] = Rt(0);
        }
        constexpr int nn = d == 0 ? 0 : d == 1 ? 3 : 5;
        m.gi[nn] = (Rt)tab[(M_I11 + nn) * np + p];
        m.sg = (Rt)tab[M_SG * np + p];
        if (use_hlld) {
            // the face frame of the HLLD solver also needs gamma^r phi and,
            // on theta faces, gamma^rr
            m.gi[2] = (Rt)tab[M_I13 * np + p];
            if (d == 1) m.gi[0] = (Rt)tab[M_I11 * np + p];
        }
        return m;
    };
    Met<Rt> m = face_met();
    constexpr int o0 = d == 0, o1 = d == 1, o2 = d == 2;
    constexpr int t1 = d == 2 ? 0 : d + 1, t2 = d == 0 ? 2 : d - 1;
    // each stencil cell's word 0, in this block or, for an own face whose
    // stencil crosses a shared side, in the same-level neighbour (a side
    // face computed here lies on an unshared side: all its cells are local)
    const double* Wc[4];
    {
        const SideNb n = side_nb(P, b);
        const int N = d == 0 ? BX : d == 1 ? BY : P.bz;
        const bool own = (d == 0 || (i >= 0 && i < BX)) && (d == 1 || (j >= 0 && j < BY)) &&
                         (d == 2 || (k >= 0 && k < P.bz));
        const int lo = n.s[2 * d], hi = n.s[2 * d + 1];
        const int a0 = d == 0 ? i : d == 1 ? j : k;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            int s = b, a = a0 + c - 2;
            if (own) {
                if (a < 0 && lo >= 0) { s = lo; a += N; }
                else if (a >= N && hi >= 0) { s = hi; a -= N; }
            }
            Wc[c] = P.W + (size_t)s * 8 * P.ncell + (d == 0 ? cidx(a, j, k) : d == 1 ? cidx(i, a, k) : cidx(i, j, a));
        }
    }
    // one variable at a time: four stencil loads, then its two face values
    auto ld = [&](int q, Rt x[4]) {
#pragma unroll
        for (int c = 0; c < 4; c++) x[c] = (Rt)Wc[c][(size_t)q * P.ncell];
    };
    Pt<Rt> PL, PR;
    {
        Rt r4[4], e4[4], pc[4];
        ld(WRHO, r4);
        ld(WEPS, e4);
#pragma unroll
        for (int c = 0; c < 4; c++) pc[c] = (Rt)(R.gam - 1.0) * r4[c] * e4[c];
        plm(r4[0], r4[1], r4[2], r4[3], PL.rho, PR.rho);
        plm(pc[0], pc[1], pc[2], pc[3], PL.p, PR.p);
    }
#pragma unroll
    for (int e = 0; e < 3; e++) {
        Rt x[4];
        ld(WU1 + e, x);
        plm(x[0], x[1], x[2], x[3], PL.u[e], PR.u[e]);
    }
    {
        Rt x[4];
        ld(WB1 + t1, x);
        plm(x[0], x[1], x[2], x[3], PL.B[t1], PR.B[t1]);
        ld(WB1 + t2, x);
        plm(x[0], x[1], x[2], x[3], PL.B[t2], PR.B[t2]);
    }
    const Rt bn = m.sg > Rt(0) ? (Rt)Bf[((size_t)b * 3 + d) * P.nf + fidx(d, i, j, k)] / m.sg : Rt(0);
    PL.B[d] = bn;
    PR.B[d] = bn;
    const Rt gam = (Rt)R.gam;
    if (use_hlld) {
        if constexpr (sizeof(Rt) == 8) {
            // FP64: the metric is read again for the map back to
            // coordinates (the asm statement keeps the compiler from reusing
            // the first loads), so that neither it nor the frame is held in
            // registers through the solve: 166 VGPRs, three wavefronts per
            // SIMD (184 and two without)
            hlld_m(m, m.gi[2], gam, PL, PR, d, out, [&](Rt& g13) {
                __asm__ volatile("" ::: "memory");
                const Met<Rt> m2 = face_met();
                g13 = m2.gi[2];
                return m2;
            });
        } else {
            hlld(m, m.gi[2], gam, PL, PR, d, out);
        }
        return;
    }
    Side<Rt> SL, SR;
    face_side(m, gam, PL, d, SL);
    face_side(m, gam, PR, d, SR);
    hll(m, SL, SR, out);
}
template <class Rt, int D, int RS = 0>
__device__ __forceinline__ void flux_face_d(const KPS& P, const double* Bf, int b, int i, int j, int k) {
    Rt out[NFW];
    flux_face_val<Rt, D, RS>(P, Bf, b, i, j, k, out);
    Rt* F = (Rt*)P.F + (size_t)b * 3 * NFW * P.nff + (size_t)D * NFW * P.nff + ffidx(D, i, j, k, P.bz);
#pragma unroll
    for (int q = 0; q < NFW; q++) F[(size_t)q * P.nff] = out[q];
}
template <class Rt, int RS = 0>
__device__ void flux_face(const KPS& P, const double* Bf, int b, int d, int i, int j, int k) {
    if (d == 0) flux_face_d<Rt, 0, RS>(P, Bf, b, i, j, k);
    else if (d == 1) flux_face_d<Rt, 1, RS>(P, Bf, b, i, j, k);
    else flux_face_d<Rt, 2, RS>(P, Bf, b, i, j, k);
}

// All faces of direction D (1: theta, 2: phi) along one line of the block at
// fixed transverse position (i, jk): faces 0 .. n_D of the line.  The
// reconstruction is carried along the line -- per variable the two previous
// cell values and the previous limited slope -- so that each cell is loaded
// once; the arithmetic is that of flux_face_d face by face.
template <class Rt, int D, int RS = 0>
__device__ void flux_pencil(const KPS& P, const double* Bf, int b, int i, int jk) {
    constexpr int d = D;
    constexpr int t1 = d == 2 ? 0 : d + 1, t2 = d == 0 ? 2 : d - 1;
    constexpr int NV = 7;  // rho, p, u~^1..3, B^t1, B^t2
    const Blk B = P.blk[b];
    const DRL& v = P.rl[B.rl];
    const RunPar& R = P.run[v.run];
    const int n = d == 1 ? BY : P.bz;
    const int j0 = d == 1 ? 0 : jk, k0 = d == 1 ? jk : 0;
    // an own line (transverse position inside the block) reads the cells
    // beyond a shared side in the same-level neighbour; a side line is local
    const SideNb nbs = side_nb(P, b);
    const bool own = i >= 0 && i < BX && (d == 1 ? (k0 >= 0 && k0 < P.bz) : (j0 >= 0 && j0 < BY));
    const int slo = own ? nbs.s[2 * d] : -1, shi = own ? nbs.s[2 * d + 1] : -1;
    const Rt gm1 = (Rt)(R.gam - 1.0);
    const Rt gam = (Rt)R.gam;
    auto cell = [&](int c, Rt w[NV]) {
        int s = b;
        if (c < 0 && slo >= 0) { s = slo; c += n; }
        else if (c >= n && shi >= 0) { s = shi; c -= n; }
        const double* Wb = P.W + (size_t)s * 8 * P.ncell + (d == 1 ? cidx(i, c, k0) : cidx(i, j0, c));
        const Rt rho = (Rt)Wb[(size_t)WRHO * P.ncell];
        const Rt eps = (Rt)Wb[(size_t)WEPS * P.ncell];
        w[0] = rho;
        w[1] = gm1 * rho * eps;
        w[2] = (Rt)Wb[(size_t)WU1 * P.ncell];
        w[3] = (Rt)Wb[(size_t)WU2 * P.ncell];
        w[4] = (Rt)Wb[(size_t)WU3 * P.ncell];
        w[5] = (Rt)Wb[(size_t)(WB1 + t1) * P.ncell];
        w[6] = (Rt)Wb[(size_t)(WB1 + t2) * P.ncell];
    };
    const int gi = B.bi * BX + i;
    const long long np = v.npt;
    auto load_m = [&](const GLB double* tab, int p, Met<Rt>& m) {
        m.alpha = (Rt)tab[M_ALPHA * np + p];
        m.beta[0] = (Rt)tab[M_B1 * np + p];
        m.beta[1] = m.beta[2] = Rt(0);
#pragma unroll
        for (int q = 0; q < 6; q++) {
            m.g[q] = (q == 1 || q == 4) ? Rt(0) : (Rt)tab[(M_G11 + q) * np + p];
            m.gi[q] = Rt(0);
        }
        constexpr int nn = d == 0 ? 0 : d == 1 ? 3 : 5;
        m.gi[nn] = (Rt)tab[(M_I11 + nn) * np + p];
        m.sg = (Rt)tab[M_SG * np + p];
        if (RS == 1 || R.riemann) {
            // the face frame of the HLLD solver also needs gamma^r phi and,
            // on theta faces, gamma^rr
            m.gi[2] = (Rt)tab[M_I13 * np + p];
            if (d == 1) m.gi[0] = (Rt)tab[M_I11 * np + p];
        }
    };
    // the face metric's table and point of face f (phi faces: the cell
    // centres, the same point for the whole line)
    const GLB double* mtab = gl(d == 1 ? v.gf + (size_t)NMET * v.npt : v.gc);
    auto mpt = [&](int f) { return (B.bj * BY + (d == 1 ? f : j0) + 2) * (v.nr + 5) + (gi + 2); };
    // entering face f: prev = cell f-1, cur = cell f, sp = limited slope of cell f-1
    Rt prev[NV], cur[NV], sp[NV];
    {
        Rt a[NV];
        cell(-2, a);
        cell(-1, prev);
        cell(0, cur);
#pragma unroll
        for (int q = 0; q < NV; q++) sp[q] = minmod(prev[q] - a[q], cur[q] - prev[q]);
    }
    Rt* Fd = (Rt*)P.F + (size_t)b * 3 * NFW * P.nff + (size_t)d * NFW * P.nff;
    const double* Bfd = Bf + ((size_t)b * 3 + d) * P.nf;
    for (int f = 0; f <= n; f++) {
        Rt nx[NV], sc[NV];
        cell(f + 1, nx);
        Pt<Rt> PL, PR;
        Rt L[NV], Rr[NV];
#pragma unroll
        for (int q = 0; q < NV; q++) {
            sc[q] = minmod(cur[q] - prev[q], nx[q] - cur[q]);
            L[q] = prev[q] + Rt(0.5) * sp[q];
            Rr[q] = cur[q] - Rt(0.5) * sc[q];
        }
        PL.rho = L[0]; PR.rho = Rr[0];
        PL.p = L[1]; PR.p = Rr[1];
#pragma unroll
        for (int e = 0; e < 3; e++) { PL.u[e] = L[2 + e]; PR.u[e] = Rr[2 + e]; }
        PL.B[t1] = L[5]; PR.B[t1] = Rr[5];
        PL.B[t2] = L[6]; PR.B[t2] = Rr[6];
        const int fj = d == 1 ? f : j0, fk = d == 1 ? k0 : f;
        Met<Rt> m;
        load_m(mtab, mpt(f), m);
        const int fx = fidx(d, i, fj, fk);
        const Rt bn = m.sg > Rt(0) ? (Rt)Bfd[fx] / m.sg : Rt(0);
        PL.B[d] = bn;
        PR.B[d] = bn;
        Rt out[NFW];
        if (RS == 1 || R.riemann) {
            if constexpr (sizeof(Rt) == 8) {
                // FP64: the metric read again after the solve, as in
                // flux_face_val
                hlld_m(m, m.gi[2], gam, PL, PR, d, out, [&](Rt& g13) {
                    __asm__ volatile("" ::: "memory");
                    Met<Rt> m2;
                    load_m(mtab, mpt(f), m2);
                    g13 = m2.gi[2];
                    return m2;
                });
            } else {
                hlld(m, m.gi[2], gam, PL, PR, d, out);
            }
        } else {
            Side<Rt> SL, SR;
            face_side(m, gam, PL, d, SL);
            face_side(m, gam, PR, d, SR);
            hll(m, SL, SR, out);
        }
#pragma unroll
        for (int q = 0; q < NFW; q++) Fd[(size_t)q * P.nff + ffidx(d, i, fj, fk, P.bz)] = out[q];
#pragma unroll
        for (int q = 0; q < NV; q++) {
            prev[q] = cur[q];
            cur[q] = nx[q];
            sp[q] = sc[q];
        }
    }
}

template <class Rt>
__device__ __forceinline__ Rt Fat(const KPS& P, const Rt* Fb, int d, int q, int i, int j, int k) {
    return Fb[((size_t)d * NFW + q) * P.nff + ffown(d, i, j, k, P.bz)];
}
// flux word of face (d; i, j, k) of block b, from the same-level neighbour
// where b left that side face to it (at most one transverse coordinate is
// outside the block: corner faces are never read)
template <class Rt>
__device__ __forceinline__ Rt Fsh(const KPS& P, int b, const SideNb& n, int d, int q, int i, int j, int k) {
    int s = b;
    if (d != 0) {
        if (i < 0 && n.s[0] >= 0) { s = n.s[0]; i += BX; }
        else if (i >= BX && n.s[1] >= 0) { s = n.s[1]; i -= BX; }
    }
    if (d != 1) {
        if (j < 0 && n.s[2] >= 0) { s = n.s[2]; j += BY; }
        else if (j >= BY && n.s[3] >= 0) { s = n.s[3]; j -= BY; }
    }
    if (d != 2) {
        if (k < 0 && n.s[4] >= 0) { s = n.s[4]; k += P.bz; }
        else if (k >= P.bz && n.s[5] >= 0) { s = n.s[5]; k -= P.bz; }
    }
    return ((const Rt*)P.F)[(((size_t)s * 3 + d) * NFW + q) * P.nff + ffidx(d, i, j, k, P.bz)];
}

template <class Rt>
__device__ __forceinline__ Rt emf_z(const KPS& P, int b, const SideNb& n, int gj0, int nth, int i, int j, int k) {
    const int gj = gj0 + j;
    if (gj == 0 || gj == nth) return Rt(0);
    return Rt(0.25) * (Fsh<Rt>(P, b, n, 1, 6, i, j, k) + Fsh<Rt>(P, b, n, 1, 6, i - 1, j, k) - Fsh<Rt>(P, b, n, 0, 5, i, j, k) -
                       Fsh<Rt>(P, b, n, 0, 5, i, j - 1, k));
}
template <class Rt>
__device__ __forceinline__ Rt emf_x(const KPS& P, int b, const SideNb& n, int gj0, int nth, int i, int j, int k) {
    const int gj = gj0 + j;
    if (gj == 0 || gj == nth) return Rt(0);
    return Rt(0.25) * (Fsh<Rt>(P, b, n, 2, 6, i, j, k) + Fsh<Rt>(P, b, n, 2, 6, i, j - 1, k) - Fsh<Rt>(P, b, n, 1, 5, i, j, k) -
                       Fsh<Rt>(P, b, n, 1, 5, i, j, k - 1));
}
template <class Rt>
__device__ __forceinline__ Rt emf_y(const KPS& P, int b, const SideNb& n, int i, int j, int k) {
    return Rt(0.25) * (Fsh<Rt>(P, b, n, 0, 6, i, j, k) + Fsh<Rt>(P, b, n, 0, 6, i, j, k - 1) - Fsh<Rt>(P, b, n, 2, 5, i, j, k) -
                       Fsh<Rt>(P, b, n, 2, 5, i - 1, j, k));
}

// own face (d; i, j, k) of block b: face field of the stage
template <class Rt>
__device__ void face_update(const KPS& P, int b, int d, int i, int j, int k, double dt, int stage) {
    const Blk B = P.blk[b];
    const DRL& v = P.rl[B.rl];
    const SideNb n = side_nb(P, b);
    const int gj0 = B.bj * BY;
    const int gi = B.bi * BX + i;
    double rhs;
    if (d == 0) {
        const double a1 = ((double)emf_z<Rt>(P, b, n, gj0, v.nth, i, j + 1, k) - (double)emf_z<Rt>(P, b, n, gj0, v.nth, i, j, k));
        const double a2 = ((double)emf_y<Rt>(P, b, n, i, j, k + 1) - (double)emf_y<Rt>(P, b, n, i, j, k));
        rhs = -a1 / v.dth + a2 / v.dph;
    } else if (d == 1) {
        const double dri = gl(v.dr)[gi + 2];
        const double a1 = ((double)emf_x<Rt>(P, b, n, gj0, v.nth, i, j, k + 1) - (double)emf_x<Rt>(P, b, n, gj0, v.nth, i, j, k));
        const double a2 = ((double)emf_z<Rt>(P, b, n, gj0, v.nth, i + 1, j, k) - (double)emf_z<Rt>(P, b, n, gj0, v.nth, i, j, k));
        rhs = -a1 / v.dph + a2 / dri;
    } else {
        const double dri = gl(v.dr)[gi + 2];
        const double a1 = ((double)emf_y<Rt>(P, b, n, i + 1, j, k) - (double)emf_y<Rt>(P, b, n, i, j, k));
        const double a2 = ((double)emf_x<Rt>(P, b, n, gj0, v.nth, i, j + 1, k) - (double)emf_x<Rt>(P, b, n, gj0, v.nth, i, j, k));
        rhs = -a1 / dri + a2 / v.dth;
    }
    const size_t x = ((size_t)b * 3 + d) * P.nf + fidx(d, i, j, k);
    if (stage == 0) P.Bf1[x] = P.Bf0[x] + dt * rhs;
    else P.Bf0[x] = 0.5 * (P.Bf0[x] + (P.Bf1[x] + dt * rhs));
}

// flux divergence, sources and the RK2 combination of D, S_j, tau in zone a
// of block b; fx(d, q, i, j, k) gives flux word q of face (d; i, j, k)
template <class Rt, class FX>
__device__ __forceinline__ void cell_hydro(const KPS& P, int b, int a, double dt, int stage, FX fx) {
    const Blk B = P.blk[b];
    const DRL& v = P.rl[B.rl];
    const RunPar& R = P.run[v.run];
    const int i = a % BX, j = (a / BX) % BY, k = a / (BX * BY);
    const int gi = B.bi * BX + i, gj = B.bj * BY + j;
    const int c = cidx(i, j, k);
    const int p = (gj + 2) * (v.nr + 5) + (gi + 2);
    const Met<double> m = load_met<double>(gl(v.gc), v.npt, p);
    double* Ub = P.U + (size_t)b * 8 * P.ncell;
    const double* Wb = P.W + (size_t)b * 8 * P.ncell;
    double* Usb = P.Us + (size_t)b * 5 * P.nint;
    double w[8];
#pragma unroll
    for (int q = 0; q < 8; q++) w[q] = Wb[(size_t)q * P.ncell + c];
    double src[5];
    source(gl(v.gc), v.npt, p, m, R.gam, w, src);
    const double dri = gl(v.dr)[gi + 2];
#pragma unroll
    for (int q = 0; q < 5; q++) {
        const double dx = ((double)fx(0, q, i + 1, j, k) - (double)fx(0, q, i, j, k)) / dri;
        const double dy = ((double)fx(1, q, i, j + 1, k) - (double)fx(1, q, i, j, k)) / v.dth;
        const double dz = ((double)fx(2, q, i, j, k + 1) - (double)fx(2, q, i, j, k)) / v.dph;
        const double dU = src[q] - (dx + dy + dz);
        double* u = Ub + (size_t)q * P.ncell + c;
        if (stage == 0) {
            Usb[(size_t)q * P.nint + a] = *u;
            *u = *u + dt * dU;
        } else {
            *u = 0.5 * (Usb[(size_t)q * P.nint + a] + (*u + dt * dU));
        }
    }
}
// the cell field of zone (i, j, k) from the stage's new face fields
__device__ __forceinline__ void cell_bfield(const KPS& P, int b, int i, int j, int k, int stage) {
    double* Ub = P.U + (size_t)b * 8 * P.ncell;
    const int c = cidx(i, j, k);
    // the stage's new face fields (selected as integers: a select of the two
    // pointers was placed in scratch memory)
    const uintptr_t a0 = (uintptr_t)P.Bf0, a1 = (uintptr_t)P.Bf1;
    const double* Bn = (const double*)(stage == 0 ? a1 : a0) + (size_t)b * 3 * P.nf;
    Ub[(size_t)UB1 * P.ncell + c] = 0.5 * (Bn[fidx(0, i, j, k)] + Bn[fidx(0, i + 1, j, k)]);
    Ub[(size_t)UB2 * P.ncell + c] = 0.5 * (Bn[P.nf + fidx(1, i, j, k)] + Bn[P.nf + fidx(1, i, j + 1, k)]);
    Ub[(size_t)UB3 * P.ncell + c] = 0.5 * (Bn[2 * P.nf + fidx(2, i, j, k)] + Bn[2 * P.nf + fidx(2, i, j, k + 1)]);
}
template <class Rt>
__device__ void cell_update(const KPS& P, int b, int a, double dt, int stage) {
    const Rt* Fb = (const Rt*)P.F + (size_t)b * 3 * NFW * P.nff;
    cell_hydro<Rt>(P, b, a, dt, stage,
                   [&](int d, int q, int i, int j, int k) { return Fat(P, Fb, d, q, i, j, k); });
    cell_bfield(P, b, a % BX, (a / BX) % BY, a / (BX * BY), stage);
}

__device__ __forceinline__ double run_dt(const KPS& P, int b) {
    return dtbits_to(P.ctr[P.rl[P.blk[b].rl].run].dtbits);
}

// ---------------------------------------------------- multi-kernel schedule
constexpr int TPB = 256;

__global__ void k_begin(RunCtr* ctr, const int* runs, int nruns, unsigned int* qctr) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nruns) ctr[runs[t]].dtbits = 0x7ff0000000000000ull;  // +inf
    if (t < 15) qctr[t] = 0;  // not qctr[15], the barrier timeout flag
}

// Launch variants the autotuner chooses from: workgroup size TB and a
// minimum of MW wavefronts per SIMD (MW > 1 caps the registers per thread).
template <int TB, int MW>
// parts workgroups per block, each a strided quarter (or all) of its zones:
// the run's dt is a minimum and the counters are sums, so how the zones
// are dealt out does not change them
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(MW, 8))) void k_c2p(KPS P, int want_dt, int parts) {
    const int w = wg_index(P);  // a block's parts on one XCD: they share its metric rows in L2
    phase_a_block<TB>(P, P.list[w / parts], want_dt != 0, w % parts, parts);
}

// The ghost cells a stencil reads (ghost_read) and the side faces of the
// sides the block does not share, filled by ghost_cell and ghost_face in one
// launch from the list of the block's sharing pattern, so that no thread is
// spent on a cell or face the block skips (on the 3-level torus 13 % of the
// ghost cells are read): parts workgroups per block.  The ghost cells are
// written from interiors and the side faces from own faces, so the two
// kinds of work are independent.
__device__ void ghost_part(const KPS& P, int b, int h, int parts, int buf) {
    const int m = side_mask(side_nb(P, b));
    const int o1 = P.ghp_off[m + 1];
    double* Bf = buf ? P.Bf1 : P.Bf0;
    for (int t = P.ghp_off[m] + h * TPB + threadIdx.x; t < o1; t += parts * TPB) {
        const int x = P.ghp[t];
        int i, j, k;
        unpack(x & ((1 << 24) - 1), i, j, k);
        if (x & (1 << 30)) ghost_face(P, Bf, b, (x >> 24) & 3, i, j, k);
        else ghost_cell(P, b, i, j, k);
    }
}
__global__ __launch_bounds__(TPB) void k_ghost(KPS P, int buf, int parts) {
    const int w = wg_index(P);
    ghost_part(P, P.list[w / parts], w % parts, parts, buf);
}

// fluxes with the theta and phi faces by lines (flux_pencil) and the r faces
// one per thread; P.flp lists the lines first (their threads run longest)
template <class Rt, int TB, int MW, int RS>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(MW, 8))) void k_flux_lines(KPS P, int buf) {
    const int per = (P.nflp + TB - 1) / TB;
    const int w = wg_index(P);
    const int b = P.list[w / per];
    const int t = (w % per) * TB + threadIdx.x;
    if (t >= P.nflp) return;
    int i, j, k;
    const int x = P.flp[t];
    unpack(x, i, j, k);
    const int d = (x >> 24) & 3;
    const double* Bf = buf ? P.Bf1 : P.Bf0;
    if (face_shared(P, b, d, i, j, k)) return;  // a line lies on one side
    if (x & (1 << 26)) {
        if (d == 1) flux_pencil<Rt, 1, RS>(P, Bf, b, i, k);
        else flux_pencil<Rt, 2, RS>(P, Bf, b, i, j);
    } else {
        flux_face_d<Rt, 0, RS>(P, Bf, b, i, j, k);
    }
}

template <class Rt, int TB, int MW, int RS>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(MW, 8))) void k_flux(KPS P, int buf) {
    const int per = (P.nflf + TB - 1) / TB;
    const int w = wg_index(P);
    const int b = P.list[w / per];
    const int t = (w % per) * TB + threadIdx.x;
    if (t >= P.nflf) return;
    int i, j, k;
    const int x = P.flf[t];
    unpack(x, i, j, k);
    if (face_shared(P, b, x >> 24, i, j, k)) return;
    flux_face<Rt, RS>(P, buf ? P.Bf1 : P.Bf0, b, x >> 24, i, j, k);
}

template <class Rt>
__global__ __launch_bounds__(TPB) void k_face(KPS P, int stage) {
    const int per = (P.nowf + TPB - 1) / TPB;
    const int w = wg_index(P);
    const int b = P.list[w / per];
    const int t = (w % per) * TPB + threadIdx.x;
    if (t >= P.nowf) return;
    int i, j, k;
    const int x = P.owf[t];
    unpack(x, i, j, k);
    face_update<Rt>(P, b, x >> 24, i, j, k, run_dt(P, b), stage);
}

// The face-field update with each edge's EMF formed once.  k_face forms the
// four EMFs of a face from its four flux words each, so every EMF is formed
// by each of the (up to four) faces around its edge, and most of the work is
// flux addressing.  Here nsplit workgroups share a block's phi layers: part
// h takes layers [k0, k1), first forms the EMFs of every edge one of its own
// faces reads into LDS (the same emf_x/y/z calls, stored in Rt as k_face
// uses them), then updates those faces from LDS with face_update's
// expressions.  The phi faces at k1 belong to the next part, except the
// block's last one (k1 = bz), which the last part updates.
//   Ez(i, j, kk): edges along phi at (r_{i-1/2}, theta_{j-1/2}, phi_{k0+kk}),
//                 i, j in [0, 16], kk in [0, KL)
//   Ey(i, j, kk): edges along theta at (r_{i-1/2}, theta_j, phi_{k0+kk-1/2}),
//                 i in [0, 16], j in [0, 16), kk in [0, KL]
//   Ex(i, j, kk): edges along r at (r_i, theta_{j-1/2}, phi_{k0+kk-1/2}),
//                 i in [0, 16), j in [0, 16], kk in [0, KL]
constexpr int FE_Z = (BX + 1) * (BY + 1), FE_Y = (BX + 1) * BY, FE_X = BX * (BY + 1);
__host__ __device__ constexpr int face_edges(int kl) { return FE_Z * kl + (FE_Y + FE_X) * (kl + 1); }

template <class Rt>
__device__ void face_edges_part(const KPS& P, int b, int h, int nsplit, int stage, Rt* sE) {
    const int k0 = h * P.bz / nsplit, k1 = (h + 1) * P.bz / nsplit;
    const int kl = k1 - k0;
    Rt* Ez = sE;
    Rt* Ey = Ez + FE_Z * kl;
    Rt* Ex = Ey + FE_Y * (kl + 1);
    const Blk B = P.blk[b];
    const DRL& v = P.rl[B.rl];
    const int gj0 = B.bj * BY;
    {
        const SideNb n = side_nb(P, b);
        const int ne = face_edges(kl);
        for (int e = threadIdx.x; e < ne; e += TPB) {
            int r = e;
            if (r < FE_Z * kl) {
                const int kk = r / FE_Z, q = r % FE_Z;
                Ez[r] = emf_z<Rt>(P, b, n, gj0, v.nth, q % (BX + 1), q / (BX + 1), k0 + kk);
                continue;
            }
            r -= FE_Z * kl;
            if (r < FE_Y * (kl + 1)) {
                const int kk = r / FE_Y, q = r % FE_Y;
                Ey[r] = emf_y<Rt>(P, b, n, q % (BX + 1), q / (BX + 1), k0 + kk);
                continue;
            }
            r -= FE_Y * (kl + 1);
            const int kk = r / FE_X, q = r % FE_X;
            Ex[r] = emf_x<Rt>(P, b, n, gj0, v.nth, q % BX, q / BX, k0 + kk);
        }
    }
    __syncthreads();
    const double dt = run_dt(P, b);
    const bool last = k1 == P.bz;
    const int nr = (BX + 1) * BY * kl, nt = BX * (BY + 1) * kl, np = BX * BY * (kl + (last ? 1 : 0));
    auto ez = [&](int i, int j, int kk) { return (double)Ez[(kk * (BY + 1) + j) * (BX + 1) + i]; };
    auto ey = [&](int i, int j, int kk) { return (double)Ey[(kk * BY + j) * (BX + 1) + i]; };
    auto ex = [&](int i, int j, int kk) { return (double)Ex[(kk * (BY + 1) + j) * BX + i]; };
    for (int f = threadIdx.x; f < nr + nt + np; f += TPB) {
        int d, i, j, kk;
        double rhs;
        if (f < nr) {
            d = 0;
            i = f % (BX + 1);
            j = (f / (BX + 1)) % BY;
            kk = f / ((BX + 1) * BY);
            const double a1 = (ez(i, j + 1, kk) - ez(i, j, kk));
            const double a2 = (ey(i, j, kk + 1) - ey(i, j, kk));
            rhs = -a1 / v.dth + a2 / v.dph;
        } else if (f < nr + nt) {
            const int g = f - nr;
            d = 1;
            i = g % BX;
            j = (g / BX) % (BY + 1);
            kk = g / (BX * (BY + 1));
            const double dri = gl(v.dr)[B.bi * BX + i + 2];
            const double a1 = (ex(i, j, kk + 1) - ex(i, j, kk));
            const double a2 = (ez(i + 1, j, kk) - ez(i, j, kk));
            rhs = -a1 / v.dph + a2 / dri;
        } else {
            const int g = f - nr - nt;
            d = 2;
            i = g % BX;
            j = (g / BX) % BY;
            kk = g / (BX * BY);
            const double dri = gl(v.dr)[B.bi * BX + i + 2];
            const double a1 = (ey(i + 1, j, kk) - ey(i, j, kk));
            const double a2 = (ex(i, j + 1, kk) - ex(i, j, kk));
            rhs = -a1 / dri + a2 / v.dth;
        }
        const size_t x = ((size_t)b * 3 + d) * P.nf + fidx(d, i, j, k0 + kk);
        if (stage == 0) P.Bf1[x] = P.Bf0[x] + dt * rhs;
        else P.Bf0[x] = 0.5 * (P.Bf0[x] + (P.Bf1[x] + dt * rhs));
    }
}
template <class Rt>
__global__ __launch_bounds__(TPB) void k_face_edges(KPS P, int stage, int nsplit) {
    extern __shared__ __align__(16) unsigned char face_lds[];
    const int w = wg_index(P);
    face_edges_part<Rt>(P, P.list[w / nsplit], w % nsplit, nsplit, stage, (Rt*)face_lds);
}

template <class Rt, int TB, int MW>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(MW, 8))) void k_update(KPS P, int stage) {
    const int per = (P.nint + TB - 1) / TB;
    const int w = wg_index(P);
    const int b = P.list[w / per];
    const int a = (w % per) * TB + threadIdx.x;
    if (a >= P.nint) return;
    cell_update<Rt>(P, b, a, run_dt(P, b), stage);
}

// Stage 0's update with stage 1's con2prim of the same zone fused behind it
// (schedule 3): con2prim is a function of the zone's new U alone, so doing it
// here instead of in a separate pass gives the same W, U and counters, and
// saves a pass over U and the launch.
template <class Rt, int TB, int MW>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(MW, 8))) void k_update_c2p(KPS P) {
    const int per = (P.nint + TB - 1) / TB;
    const int w = wg_index(P);
    const int b = P.list[w / per];
    const int a = (w % per) * TB + threadIdx.x;
    int nfail = 0, nfl = 0;
    if (a < P.nint) {
        cell_update<Rt>(P, b, a, run_dt(P, b), 0);
        c2p_zone(P, b, a, false, nfail, nfl);
    }
    // every thread of the workgroup reaches the commit's barriers
    wg_commit<TB>(P, P.rl[P.blk[b].rl].run, 1e300, nfail, nfl, false);
}

// ---------------------------------------------------- slab schedule
// Fluxes and the hydro update fused per block: one workgroup walks the
// block's phi layers k = 0 .. bz-1; for layer k it forms the fluxes of the
// layer's r and theta faces and of the phi faces at k + 1 (those at k are
// kept from the layer before), keeps words 0-4 of each in LDS and stores the
// EMF words 5, 6 for k_face, then updates the layer's zones from LDS.  Words
// 0-4 never reach HBM.  The side faces of the sides the block does not share
// are formed first (EMF words only); the cell field follows k_face (k_cellb).  Every
// flux and update is computed by the functions of the sub-phase kernels, so
// the results are the same bit for bit.
constexpr int SL_X = (BX + 1) * BY, SL_Y = BX * (BY + 1), SL_Z = BX * BY;
// nsplit workgroups share a block: part h takes layers [h bz/nsplit,
// (h+1) bz/nsplit) and forms the phi faces at its first layer itself (those
// at a shared boundary layer are formed by both parts, with the same bits);
// the side faces are dealt out over the parts.
template <class Rt, int RS>
__global__ __launch_bounds__(TPB) void k_slab(KPS P, int buf, int stage, int nsplit) {
    __shared__ Rt sx[5][SL_X], sy[5][SL_Y], sz[2][5][SL_Z];
    const int wg = wg_index(P);
    const int b = P.list[wg / nsplit], part = wg % nsplit;
    const int k0 = part * P.bz / nsplit, k1 = (part + 1) * P.bz / nsplit;
    const double* Bf = buf ? P.Bf1 : P.Bf0;
    const double dt = run_dt(P, b);
    Rt* Fb = (Rt*)P.F + (size_t)b * 3 * NFW * P.nff;
    auto emf_words = [&](int D, int i, int j, int k, const Rt out[NFW]) {
        const size_t x = ffidx(D, i, j, k, P.bz);
        Fb[((size_t)D * NFW + 5) * P.nff + x] = out[5];
        Fb[((size_t)D * NFW + 6) * P.nff + x] = out[6];
    };
    // the side faces of the sides this block does not share (only their EMF
    // words are read; a shared side's faces are the neighbour's own)
    {
        const SideNb n = side_nb(P, b);
        for (int g = part; g < 18; g += nsplit) {
            const int e = g % 6, D = g / 6;
            if ((e == 0 ? n.s[0] : e == 1 ? n.s[1] : e == 2 ? n.s[2] : e == 3 ? n.s[3] : e == 4 ? n.s[4] : n.s[5]) >= 0)
                continue;
            const int off = P.sgrp[2 * g], cnt = P.sgrp[2 * g + 1];
            for (int t = threadIdx.x; t < cnt; t += TPB) {
                int i, j, k;
                unpack(P.flf[off + t], i, j, k);
                Rt out[NFW];
                if (D == 0) flux_face_val<Rt, 0, RS>(P, Bf, b, i, j, k, out);
                else if (D == 1) flux_face_val<Rt, 1, RS>(P, Bf, b, i, j, k, out);
                else flux_face_val<Rt, 2, RS>(P, Bf, b, i, j, k, out);
                emf_words(D, i, j, k, out);
            }
        }
    }
    for (int t = threadIdx.x; t < SL_Z; t += TPB) {
        const int i = t % BX, j = t / BX;
        Rt out[NFW];
        flux_face_val<Rt, 2, RS>(P, Bf, b, i, j, k0, out);
#pragma unroll
        for (int q = 0; q < 5; q++) sz[k0 & 1][q][t] = out[q];
        emf_words(2, i, j, k0, out);
    }
    __syncthreads();
    for (int k = k0; k < k1; k++) {
        const int cur = k & 1;
        for (int t = threadIdx.x; t < SL_X + SL_Y + SL_Z; t += TPB) {
            Rt out[NFW];
            if (t < SL_X) {
                const int i = t % (BX + 1), j = t / (BX + 1);
                flux_face_val<Rt, 0, RS>(P, Bf, b, i, j, k, out);
#pragma unroll
                for (int q = 0; q < 5; q++) sx[q][t] = out[q];
                emf_words(0, i, j, k, out);
            } else if (t < SL_X + SL_Y) {
                const int u = t - SL_X, i = u % BX, j = u / BX;
                flux_face_val<Rt, 1, RS>(P, Bf, b, i, j, k, out);
#pragma unroll
                for (int q = 0; q < 5; q++) sy[q][u] = out[q];
                emf_words(1, i, j, k, out);
            } else {
                const int u = t - SL_X - SL_Y, i = u % BX, j = u / BX;
                flux_face_val<Rt, 2, RS>(P, Bf, b, i, j, k + 1, out);
#pragma unroll
                for (int q = 0; q < 5; q++) sz[cur ^ 1][q][u] = out[q];
                emf_words(2, i, j, k + 1, out);
            }
        }
        __syncthreads();
        for (int t = threadIdx.x; t < BX * BY; t += TPB)
            cell_hydro<Rt>(P, b, k * BX * BY + t, dt, stage, [&](int d, int q, int ii, int jj, int kk) -> Rt {
                if (d == 0) return sx[q][jj * (BX + 1) + ii];
                if (d == 1) return sy[q][jj * BX + ii];
                return sz[kk == k ? cur : cur ^ 1][q][jj * BX + ii];
            });
        __syncthreads();
    }
}
// the cell field of every active zone from the stage's new faces
__global__ __launch_bounds__(TPB) void k_cellb(KPS P, int stage) {
    const int per = (P.nint + TPB - 1) / TPB;
    const int w = wg_index(P);
    const int b = P.list[w / per];
    const int a = (w % per) * TPB + threadIdx.x;
    if (a >= P.nint) return;
    cell_bfield(P, b, a % BX, (a / BX) % BY, a / (BX * BY), stage);
}

// ---------------------------------------------------- persistent schedule
// One launch per stage.  Workgroups pull blocks for phase A from a queue,
// meet at one grid barrier (every workgroup of the launch is resident: the
// grid is sized from the occupancy of this kernel), then pull blocks again
// for phase B and run its four sub-phases with workgroup barriers between
// them.  The barrier wait is bounded (P.spin polls); a launch that is not
// co-resident sets the timeout flag qctr[QFLAG] and every workgroup still
// exits.  The flag is not cleared by k_begin (so a timeout in stage 0 is
// still seen after stage 1), only by the host once it has read it.
constexpr unsigned SPIN_LIMIT = 1u << 26;
constexpr int QFLAG = 15;

__device__ __forceinline__ void grid_barrier(unsigned int* arrive, unsigned int* flag, unsigned int n,
                                             unsigned int limit) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        atomicAdd(arrive, 1u);
        unsigned spins = 0;
        while (__hip_atomic_load(arrive, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < n) {
            __builtin_amdgcn_s_sleep(2);
            if (++spins > limit) {
                atomicExch(flag, 1u);
                break;
            }
        }
        __threadfence();
    }
    __syncthreads();
}

template <class Rt>
__global__ __launch_bounds__(TPB) void k_stage(KPS P, int stage) {
    __shared__ int s_b;
    const bool want_dt = stage == 0;
    for (;;) {
        if (threadIdx.x == 0) s_b = (int)atomicAdd(&P.qctr[0], 1u);
        __syncthreads();
        const int q = s_b;
        __syncthreads();
        if (q >= P.nlist) break;
        phase_a_block<TPB>(P, P.list[q], want_dt);
    }
    grid_barrier(&P.qctr[1], &P.qctr[QFLAG], gridDim.x, P.spin);
    if (P.qctr[QFLAG]) return;
    double* Bc = stage == 0 ? P.Bf0 : P.Bf1;
    for (;;) {
        if (threadIdx.x == 0) s_b = (int)atomicAdd(&P.qctr[2], 1u);
        __syncthreads();
        const int q = s_b;
        __syncthreads();
        if (q >= P.nlist) break;
        const int b = P.list[q];
        const double dt = run_dt(P, b);
        for (int t = threadIdx.x; t < P.nghc; t += TPB) {
            int i, j, k;
            unpack(P.ghc[t], i, j, k);
            if (ghost_read(P, side_nb(P, b), i, j, k)) ghost_cell(P, b, i, j, k);
        }
        for (int t = threadIdx.x; t < P.nghf; t += TPB) {
            int i, j, k;
            const int x = P.ghf[t];
            unpack(x, i, j, k);
            ghost_face(P, Bc, b, x >> 24, i, j, k);
        }
        __syncthreads();
        for (int t = threadIdx.x; t < P.nflf; t += TPB) {
            int i, j, k;
            const int x = P.flf[t];
            unpack(x, i, j, k);
            flux_face<Rt>(P, Bc, b, x >> 24, i, j, k);
        }
        __syncthreads();
        for (int t = threadIdx.x; t < P.nowf; t += TPB) {
            int i, j, k;
            const int x = P.owf[t];
            unpack(x, i, j, k);
            face_update<Rt>(P, b, x >> 24, i, j, k, dt, stage);
        }
        __syncthreads();
        for (int a = threadIdx.x; a < P.nint; a += TPB) cell_update<Rt>(P, b, a, dt, stage);
        __syncthreads();
    }
}


// ---------------------------------------------------- queue schedule
// One launch per RK stage (schedule 4).  Every phase of the stage is a queue
// of work items over all active blocks of all levels and runs -- (block,
// part) items for con2prim, the ghost fill and the face update, (block,
// chunk of faces or zones) items for the fluxes and the zone update --
// that resident workgroups pull with one atomic per item; one grid barrier
// separates consecutive phases (each phase reads what the one before wrote
// in other blocks).  Stage 0: con2prim (with the run's CFL dt) | ghost |
// fluxes | face update | zone update with stage 1's con2prim; stage 1:
// ghost | fluxes | face update | zone update.  The work of every item is
// that of the sub-phase kernels, by the same functions, so the results are
// those of schedule 3 bit for bit.  Queue counters qctr[4..12] and the
// barrier counters are cleared by k_begin before each stage; the grid is
// sized from the kernel's occupancy so that every workgroup is resident,
// and a barrier wait that runs out sets qctr[QFLAG] and every workgroup exits.
constexpr int QPARTS = 4;  // (block, part) items per block of the con2prim and ghost phases
template <class Rt, int RS, int MW>
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(MW, 8))) void k_queue(KPS P, int stage, int pe) {
    extern __shared__ __align__(16) unsigned char q_lds[];
    __shared__ int s_item;
    auto next = [&](int c) {
        __syncthreads();
        if (threadIdx.x == 0) s_item = (int)atomicAdd(&P.qctr[c], 1u);
        __syncthreads();
        return s_item;
    };
    const int nl = P.nlist;
    const unsigned n = gridDim.x;
    const double* Bf = stage ? P.Bf1 : P.Bf0;
    if (stage == 0) {
        for (int q = next(4); q < nl * QPARTS; q = next(4)) {
            const int b = P.list[q / QPARTS], h = q % QPARTS;
            double mn = 1e300;
            int nfail = 0, nfl = 0;
            for (int a = h * TPB + threadIdx.x; a < P.nint; a += QPARTS * TPB)
                mn = fmin(mn, c2p_zone(P, b, a, true, nfail, nfl));
            wg_commit<TPB>(P, P.rl[P.blk[b].rl].run, mn, nfail, nfl, true);
        }
        grid_barrier(&P.qctr[5], &P.qctr[QFLAG], n, P.spin);
        if (P.qctr[QFLAG]) return;
    }
    for (int q = next(6); q < nl * QPARTS; q = next(6)) ghost_part(P, P.list[q / QPARTS], q % QPARTS, QPARTS, stage);
    grid_barrier(&P.qctr[7], &P.qctr[QFLAG], n, P.spin);
    if (P.qctr[QFLAG]) return;
    const int nfc = (P.nflf + TPB - 1) / TPB;
    for (int q = next(8); q < nl * nfc; q = next(8)) {
        const int b = P.list[q / nfc];
        const int t = (q % nfc) * TPB + threadIdx.x;
        if (t < P.nflf) {
            int i, j, k;
            const int x = P.flf[t];
            unpack(x, i, j, k);
            if (!face_shared(P, b, x >> 24, i, j, k)) flux_face<Rt, RS>(P, Bf, b, x >> 24, i, j, k);
        }
    }
    grid_barrier(&P.qctr[9], &P.qctr[QFLAG], n, P.spin);
    if (P.qctr[QFLAG]) return;
    for (int q = next(10); q < nl * pe; q = next(10)) face_edges_part<Rt>(P, P.list[q / pe], q % pe, pe, stage, (Rt*)q_lds);
    grid_barrier(&P.qctr[11], &P.qctr[QFLAG], n, P.spin);
    if (P.qctr[QFLAG]) return;
    const int nuc = (P.nint + TPB - 1) / TPB;
    for (int q = next(12); q < nl * nuc; q = next(12)) {
        const int b = P.list[q / nuc];
        const int a = (q % nuc) * TPB + threadIdx.x;
        int nfail = 0, nfl = 0;
        if (a < P.nint) {
            cell_update<Rt>(P, b, a, run_dt(P, b), stage);
            if (stage == 0) c2p_zone(P, b, a, false, nfail, nfl);
        }
        if (stage == 0) wg_commit<TPB>(P, P.rl[P.blk[b].rl].run, 1e300, nfail, nfl, false);
    }
}

// ---------------------------------------------------- initial data
struct TorusDev {
    TorusPar T;
    double beta, acut, pamp;
    U128 st, inc;
};

// torus density of every stored cell (for the corner averages) into rt
// (0 outside the torus), the primitive state of the active zones with the
// pressure in the eps slot, and the run's rho max (bit pattern) in mx[0]
__global__ __launch_bounds__(TPB) void k_ic_torus(KPS P, TorusDev D, double* rt, unsigned long long* mx) {
    const int per = (P.ncell + TPB - 1) / TPB;
    const int b = P.list[blockIdx.x / per];
    const int c = (blockIdx.x % per) * TPB + threadIdx.x;
    if (c >= P.ncell) return;
    const int i = c % SX - G, j = (c / SX) % SY - G, k = c / (SX * SY) - G;
    const Blk B = P.blk[b];
    const DRL& v = P.rl[B.rl];
    const int gi = B.bi * BX + i, gj = B.bj * BY + j;
    const int p = (gj + 2) * (v.nr + 5) + (gi + 2);
    const long long np = v.npt;
    const double rr = v.gc[C_R * np + p], s = v.gc[C_SIN * np + p], co = v.gc[C_COS * np + p];
    const double al = v.gc[M_ALPHA * np + p], b1 = v.gc[M_B1 * np + p];
    const double g11 = v.gc[M_G11 * np + p], g13 = v.gc[M_G13 * np + p], g33 = v.gc[M_G33 * np + p];
    const double beta_l1 = g11 * b1, beta_l3 = g13 * b1;
    const double gtt = -al * al + beta_l1 * b1, gtp = beta_l3, gpp = g33;
    double rho, pp, ur, uph;
    const bool in = fm_cell<DevCR>(D.T, rr, s, co, gtt, gtp, gpp, b1, rho, pp, ur, uph) != 0;
    rt[(size_t)b * P.ncell + c] = in ? rho : 0.0;
    const bool act = i >= 0 && i < BX && j >= 0 && j < BY && k >= 0 && k < P.bz;
    if (in && act) {
        double* Wb = P.W + (size_t)b * 8 * P.ncell;
        Wb[(size_t)WRHO * P.ncell + c] = rho;
        Wb[(size_t)WEPS * P.ncell + c] = pp;
        Wb[(size_t)WU1 * P.ncell + c] = ur;
        Wb[(size_t)WU2 * P.ncell + c] = 0.0;
        Wb[(size_t)WU3 * P.ncell + c] = uph;
        atomicMax(mx, (unsigned long long)__double_as_longlong(rho));
    }
}

__device__ __forceinline__ double ic_corner(const KPS& P, const double* rt, int b, double irm, double acut, int i,
                                            int j, int k) {
    const double* r = rt + (size_t)b * P.ncell;
    const double q = 0.25 * (r[cidx(i - 1, j - 1, k)] * irm + r[cidx(i, j - 1, k)] * irm +
                             r[cidx(i - 1, j, k)] * irm + r[cidx(i, j, k)] * irm);
    return fmax(q - acut, 0.0);
}

// face fields of the own r and theta faces from the corner potential
__global__ __launch_bounds__(TPB) void k_ic_faces(KPS P, TorusDev D, const double* rt,
                                                  const unsigned long long* mx) {
    const int per = (P.nowf + TPB - 1) / TPB;
    const int b = P.list[blockIdx.x / per];
    const int t = (blockIdx.x % per) * TPB + threadIdx.x;
    if (t >= P.nowf) return;
    int i, j, k;
    const int x = P.owf[t];
    unpack(x, i, j, k);
    const int d = x >> 24;
    const Blk B = P.blk[b];
    const DRL& v = P.rl[B.rl];
    const double irm = 1.0 / __longlong_as_double((long long)mx[0]);
    double val = 0.0;
    if (d == 0) {
        val = (ic_corner(P, rt, b, irm, D.acut, i, j + 1, k) - ic_corner(P, rt, b, irm, D.acut, i, j, k)) / v.dth;
    } else if (d == 1) {
        const int gi = B.bi * BX + i;
        val = -(ic_corner(P, rt, b, irm, D.acut, i + 1, j, k) - ic_corner(P, rt, b, irm, D.acut, i, j, k)) /
              v.dr[gi + 2];
    }
    P.Bf0[((size_t)b * 3 + d) * P.nf + fidx(d, i, j, k)] = val;
}

__device__ __forceinline__ void ic_cell_b(const KPS& P, int b, int i, int j, int k, double sg, double* B) {
    const double* F0 = P.Bf0 + (size_t)b * 3 * P.nf;
    B[0] = 0.5 * (F0[fidx(0, i, j, k)] + F0[fidx(0, i + 1, j, k)]) / sg;
    B[1] = 0.5 * (F0[P.nf + fidx(1, i, j, k)] + F0[P.nf + fidx(1, i, j + 1, k)]) / sg;
    B[2] = 0.5 * (F0[2 * P.nf + fidx(2, i, j, k)] + F0[2 * P.nf + fidx(2, i, j, k + 1)]) / sg;
}

// normalised torus / atmosphere state of the active zones, its cell field,
// and the maxima of p_gas (mx[1]) and p_mag (mx[2])
__global__ __launch_bounds__(TPB) void k_ic_cells(KPS P, TorusDev D, const double* rt, unsigned long long* mx) {
    const int per = (P.nint + TPB - 1) / TPB;
    const int b = P.list[blockIdx.x / per];
    const int a = (blockIdx.x % per) * TPB + threadIdx.x;
    if (a >= P.nint) return;
    const int i = a % BX, j = (a / BX) % BY, k = a / (BX * BY);
    const Blk B = P.blk[b];
    const DRL& v = P.rl[B.rl];
    const RunPar& R = P.run[v.run];
    const int gi = B.bi * BX + i, gj = B.bj * BY + j;
    const int p = (gj + 2) * (v.nr + 5) + (gi + 2);
    const int cc = cidx(i, j, k);
    const long long np = v.npt;
    const Met<double> m = load_met<double>(v.gc, np, p);
    const double irm = 1.0 / __longlong_as_double((long long)mx[0]);
    double* Wb = P.W + (size_t)b * 8 * P.ncell;
    double w[8];
    if (rt[(size_t)b * P.ncell + cc] > 0.0) {
        w[WRHO] = Wb[(size_t)WRHO * P.ncell + cc] * irm;
        w[WEPS] = Wb[(size_t)WEPS * P.ncell + cc] * irm;
        w[WU1] = Wb[(size_t)WU1 * P.ncell + cc];
        w[WU2] = 0.0;
        w[WU3] = Wb[(size_t)WU3 * P.ncell + cc];
    } else {
        w[WRHO] = v.gc[C_RFL * np + p];
        w[WEPS] = (R.gam - 1.0) * w[WRHO] * v.gc[C_EFL * np + p];
        w[WU1] = w[WU2] = w[WU3] = 0.0;
    }
    ic_cell_b(P, b, i, j, k, m.sg, w + WB1);
#pragma unroll
    for (int q = 0; q < 8; q++) Wb[(size_t)q * P.ncell + cc] = w[q];
    double Bl[3], ul[3];
    lower3(m.g, w + WB1, Bl);
    lower3(m.g, w + WU1, ul);
    const double W2 = 1.0 + dot3(ul, w + WU1);
    const double Bu = dot3(Bl, w + WU1);
    const double b2 = (dot3(Bl, w + WB1) + Bu * Bu) / W2;
    atomicMax(mx + 1, (unsigned long long)__double_as_longlong(w[WEPS]));
    atomicMax(mx + 2, (unsigned long long)__double_as_longlong(0.5 * b2));
}

__global__ __launch_bounds__(TPB) void k_ic_scale(KPS P, const unsigned long long* mx, double beta, double* scale_out) {
    const int per = (3 * P.nf + TPB - 1) / TPB;
    const int b = P.list[blockIdx.x / per];
    const int t = (blockIdx.x % per) * TPB + threadIdx.x;
    const double pg = __longlong_as_double((long long)mx[1]), pm = __longlong_as_double((long long)mx[2]);
    const double scale = pm > 0.0 ? sqrt(pg / (beta * pm)) : 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) *scale_out = scale;
    if (t >= 3 * P.nf) return;
    P.Bf0[(size_t)b * 3 * P.nf + t] *= scale;
}

__global__ __launch_bounds__(TPB) void k_ic_final(KPS P, TorusDev D, const double* rt) {
    const int per = (P.nint + TPB - 1) / TPB;
    const int b = P.list[blockIdx.x / per];
    const int a = (blockIdx.x % per) * TPB + threadIdx.x;
    if (a >= P.nint) return;
    const int i = a % BX, j = (a / BX) % BY, k = a / (BX * BY);
    const Blk B = P.blk[b];
    const DRL& v = P.rl[B.rl];
    const RunPar& R = P.run[v.run];
    const int gi = B.bi * BX + i, gj = B.bj * BY + j, gk = B.bk * P.bz + k;
    const int p = (gj + 2) * (v.nr + 5) + (gi + 2);
    const int cc = cidx(i, j, k);
    const long long np = v.npt;
    const Met<double> m = load_met<double>(v.gc, np, p);
    double* Wb = P.W + (size_t)b * 8 * P.ncell;
    double* Ub = P.U + (size_t)b * 8 * P.ncell;
    double w[8];
#pragma unroll
    for (int q = 0; q < 8; q++) w[q] = Wb[(size_t)q * P.ncell + cc];
    ic_cell_b(P, b, i, j, k, m.sg, w + WB1);
    double pp = w[WEPS];
    if (rt[(size_t)b * P.ncell + cc] > 0.0) {
        const unsigned long long n = (unsigned long long)(v.zone0 + ((long long)gk * v.nth + gj) * v.nr + gi);
        const double u = pcg64_uniform(D.st, D.inc, n);
        pp = pp * (1.0 + D.pamp * (2.0 * u - 1.0));
    }
    w[WEPS] = pp / ((R.gam - 1.0) * w[WRHO]);
    const double rfl = v.gc[C_RFL * np + p], efl = v.gc[C_EFL * np + p];
    if (w[WRHO] < rfl) w[WRHO] = rfl;
    if (w[WEPS] < efl) w[WEPS] = efl;
#pragma unroll
    for (int q = 0; q < 8; q++) Wb[(size_t)q * P.ncell + cc] = w[q];
    double u8[8];
    ptoc(m, R.gam, w, u8);
#pragma unroll
    for (int q = 0; q < 5; q++) Ub[(size_t)q * P.ncell + cc] = u8[q];
    const double* F0 = P.Bf0 + (size_t)b * 3 * P.nf;
    Ub[(size_t)UB1 * P.ncell + cc] = 0.5 * (F0[fidx(0, i, j, k)] + F0[fidx(0, i + 1, j, k)]);
    Ub[(size_t)UB2 * P.ncell + cc] = 0.5 * (F0[P.nf + fidx(1, i, j, k)] + F0[P.nf + fidx(1, i, j + 1, k)]);
    Ub[(size_t)UB3 * P.ncell + cc] = 0.5 * (F0[2 * P.nf + fidx(2, i, j, k)] + F0[2 * P.nf + fidx(2, i, j, k + 1)]);
}

// ---------------------------------------------------- diagnostics
// per block: mass, Killing energy (fixed-order tree sums over the block's
// zones), max |div B| and the divergence stencil's scale
__global__ __launch_bounds__(TPB) void k_diag(KPS P, double* out) {
    __shared__ double sm[6][TPB];
    const int b = P.list[blockIdx.x];
    const Blk B = P.blk[b];
    const DRL& v = P.rl[B.rl];
    const double* Ub = P.U + (size_t)b * 8 * P.ncell;
    const double* F0 = P.Bf0 + (size_t)b * 3 * P.nf;
    double ms = 0.0, en = 0.0, dm = 0.0, sc = 0.0, md = 0.0, ph = 0.0;
    for (int a = threadIdx.x; a < P.nint; a += TPB) {
        const int i = a % BX, j = (a / BX) % BY, k = a / (BX * BY);
        const int gi = B.bi * BX + i, gj = B.bj * BY + j;
        const int c = cidx(i, j, k);
        const int p = (gj + 2) * (v.nr + 5) + (gi + 2);
        const long long np = v.npt;
        const double dri = v.dr[gi + 2];
        const double vol = dri * v.dth * v.dph;
        ms += Ub[(size_t)UD * P.ncell + c] * vol;
        const double al = v.gc[M_ALPHA * np + p];
        const double bS = v.gc[M_B1 * np + p] * Ub[(size_t)US1 * P.ncell + c] +
                          v.gc[M_B2 * np + p] * Ub[(size_t)US2 * P.ncell + c] +
                          v.gc[M_B3 * np + p] * Ub[(size_t)US3 * P.ncell + c];
        en += (al * (Ub[(size_t)UTAU * P.ncell + c] + Ub[(size_t)UD * P.ncell + c]) - bS) * vol;
        const double x0 = F0[fidx(0, i + 1, j, k)], x1 = F0[fidx(0, i, j, k)];
        const double y0 = F0[P.nf + fidx(1, i, j + 1, k)], y1 = F0[P.nf + fidx(1, i, j, k)];
        const double z0 = F0[2 * P.nf + fidx(2, i, j, k + 1)], z1 = F0[2 * P.nf + fidx(2, i, j, k)];
        const double t0 = (x0 - x1) / dri, t1 = (y0 - y1) / v.dth, t2 = (z0 - z1) / v.dph;
        dm = fmax(dm, fabs(t0 + t1 + t2));
        sc = fmax(sc, (fabs(x0) + fabs(x1)) / dri + (fabs(y0) + fabs(y1)) / v.dth + (fabs(z0) + fabs(z1)) / v.dph);
        if (gi == v.ih) {
            // accretion rate through the horizon row, -sqrt(gamma) D (alpha v^r - beta^r)
            // dtheta dphi, and half the unsigned radial flux sqrt(gamma) |B^r| (faces)
            const double* Wb = P.W + (size_t)b * 8 * P.ncell;
            double ut[3], ul[3], g6[6];
            for (int q = 0; q < 3; q++) ut[q] = Wb[(size_t)(WU1 + q) * P.ncell + c];
            for (int q = 0; q < 6; q++) g6[q] = v.gc[(M_G11 + q) * np + p];
            lower3(g6, ut, ul);
            const double Wl = sqrt(1.0 + dot3(ul, ut));
            const double vr = ut[0] / Wl;
            md -= Ub[(size_t)UD * P.ncell + c] * (al * vr - v.gc[M_B1 * np + p]) * v.dth * v.dph;
            ph += 0.5 * fabs(x1) * v.dth * v.dph;
        }
    }
    sm[0][threadIdx.x] = ms; sm[1][threadIdx.x] = en; sm[2][threadIdx.x] = dm; sm[3][threadIdx.x] = sc;
    sm[4][threadIdx.x] = md; sm[5][threadIdx.x] = ph;
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            sm[0][threadIdx.x] += sm[0][threadIdx.x + o];
            sm[1][threadIdx.x] += sm[1][threadIdx.x + o];
            sm[2][threadIdx.x] = fmax(sm[2][threadIdx.x], sm[2][threadIdx.x + o]);
            sm[3][threadIdx.x] = fmax(sm[3][threadIdx.x], sm[3][threadIdx.x + o]);
            sm[4][threadIdx.x] += sm[4][threadIdx.x + o];
            sm[5][threadIdx.x] += sm[5][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) out[(size_t)blockIdx.x * 6 + threadIdx.x] = sm[threadIdx.x][0];
}

// ---------------------------------------------------- allocation-free diagnostics
// Scratch of bhs_run_diagnostics_static: the per-block partials of k_diag
// (6 words each) for up to DIAG_MAX_BLOCKS blocks, then the 6 reduced
// words.  One copy per GPU, held until the process ends; calls that overlap
// on one GPU share it and race.
constexpr int DIAG_MAX_BLOCKS = 65536;
__device__ __attribute__((used)) __align__(256) double scratch[6 * DIAG_MAX_BLOCKS + 8];

// the host loop of Engine::diagnostics (sums in block order, maxima), on
// the device: one thread, the same additions in the same order
__global__ void k_diag_seq(const double* __restrict__ part, int nl, double* __restrict__ out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double ms = 0.0, en = 0.0, dm = 0.0, sc = 0.0, md = 0.0, ph = 0.0;
    for (int q = 0; q < nl; q++) {
        ms += part[6 * q];
        en += part[6 * q + 1];
        dm = dm < part[6 * q + 2] ? part[6 * q + 2] : dm;
        sc = sc < part[6 * q + 3] ? part[6 * q + 3] : sc;
        md += part[6 * q + 4];
        ph += part[6 * q + 5];
    }
    out[0] = ms; out[1] = en; out[2] = dm; out[3] = sc; out[4] = md; out[5] = ph;
}

// Analyser quantities: per block the sums of analysis_terms times the
// volume element sqrt(-g) dr dtheta dphi (words 0-5) or, on the horizon row,
// the surface element sqrt(-g) dtheta dphi (words 6-8), and half the
// unsigned radial face flux there (word 9, as Phi_B in k_diag); fixed-order
// workgroup sums.
constexpr int NANA_W = NANA + 1;
__global__ __launch_bounds__(TPB) void k_analysis(KPS P, double rho_min, double* out) {
    __shared__ double sm[NANA_W][TPB];
    const int b = P.list[blockIdx.x];
    const Blk B = P.blk[b];
    const DRL& v = P.rl[B.rl];
    const RunPar& R = P.run[v.run];
    const double* F0 = P.Bf0 + (size_t)b * 3 * P.nf;
    double acc[NANA_W];
#pragma unroll
    for (int q = 0; q < NANA_W; q++) acc[q] = 0.0;
    for (int a = threadIdx.x; a < P.nint; a += TPB) {
        const int i = a % BX, j = (a / BX) % BY, k = a / (BX * BY);
        const int gi = B.bi * BX + i, gj = B.bj * BY + j;
        const int c = cidx(i, j, k);
        const int p = (gj + 2) * (v.nr + 5) + (gi + 2);
        const Met<double> m = load_met<double>(gl(v.gc), v.npt, p);
        double w[8];
#pragma unroll
        for (int q = 0; q < 8; q++) w[q] = P.W[((size_t)b * 8 + q) * P.ncell + c];
        const bool hz = gi == v.ih;
        double t[NANA];
        analysis_terms(m, R.gam, w, rho_min, hz, t);
        const double sa = m.alpha * m.sg * v.dth * v.dph, sv = sa * gl(v.dr)[gi + 2];
#pragma unroll
        for (int q = 0; q < 6; q++) acc[q] += t[q] * sv;
#pragma unroll
        for (int q = 6; q < NANA; q++) acc[q] += t[q] * sa;
        if (hz) acc[NANA] += 0.5 * fabs(F0[fidx(0, i, j, k)]) * v.dth * v.dph;
    }
#pragma unroll
    for (int q = 0; q < NANA_W; q++) sm[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o)
#pragma unroll
            for (int q = 0; q < NANA_W; q++) sm[q][threadIdx.x] += sm[q][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < NANA_W) out[(size_t)blockIdx.x * NANA_W + threadIdx.x] = sm[threadIdx.x][0];
}

// ---------------------------------------------------- launch variants
constexpr int NVAR = 6;
constexpr int VAR_TB[NVAR] = {256, 256, 256, 128, 128, 128};
constexpr int VAR_MW[NVAR] = {1, 3, 4, 1, 3, 4};

#define SPH_VARIANTS(M) M(0, 256, 1) M(1, 256, 3) M(2, 256, 4) M(3, 128, 1) M(4, 128, 3) M(5, 128, 4)

// flux variants 0 .. 5: one face per thread; 6 .. 11: theta and phi faces
// by lines (same workgroup shapes).  hlld_all: every active run uses the
// HLLD solver (the per-face kernels then carry its code only)
constexpr int NVAR_FLUX = 2 * NVAR;
template <class Rt>
void launch_flux(int v, int nl, int nitems, int nlines, hipStream_t st, KPS P, int buf, bool hlld_all) {
    if (hlld_all) {
        switch (v) {
#define L_(i, tb, mw) \
    case i: hipLaunchKernelGGL((k_flux<Rt, tb, mw, 1>), dim3(nl * ((nitems + tb - 1) / tb)), dim3(tb), 0, st, P, buf); break;
            SPH_VARIANTS(L_)
#undef L_
#define L_(i, tb, mw) \
    case NVAR + i: hipLaunchKernelGGL((k_flux_lines<Rt, tb, mw, 1>), dim3(nl * ((nlines + tb - 1) / tb)), dim3(tb), 0, st, P, buf); break;
            SPH_VARIANTS(L_)
#undef L_
        }
        return;
    }
    switch (v) {
#define L_(i, tb, mw) \
    case i: hipLaunchKernelGGL((k_flux<Rt, tb, mw, 0>), dim3(nl * ((nitems + tb - 1) / tb)), dim3(tb), 0, st, P, buf); break;
        SPH_VARIANTS(L_)
#undef L_
#define L_(i, tb, mw) \
    case NVAR + i: hipLaunchKernelGGL((k_flux_lines<Rt, tb, mw, 0>), dim3(nl * ((nlines + tb - 1) / tb)), dim3(tb), 0, st, P, buf); break;
        SPH_VARIANTS(L_)
#undef L_
    }
}
template <class Rt>
void launch_update(int v, int nl, int nitems, hipStream_t st, KPS P, int stage) {
    switch (v) {
#define L_(i, tb, mw) \
    case i: hipLaunchKernelGGL((k_update<Rt, tb, mw>), dim3(nl * ((nitems + tb - 1) / tb)), dim3(tb), 0, st, P, stage); break;
        SPH_VARIANTS(L_)
#undef L_
    }
}
template <class Rt>
void launch_update_c2p(int v, int nl, int nitems, hipStream_t st, KPS P) {
    switch (v) {
#define L_(i, tb, mw) \
    case i: hipLaunchKernelGGL((k_update_c2p<Rt, tb, mw>), dim3(nl * ((nitems + tb - 1) / tb)), dim3(tb), 0, st, P); break;
        SPH_VARIANTS(L_)
#undef L_
    }
}
// face variants: 0 one face per thread (k_face); 1 .. 4 the edge-EMF kernel
// with 2, 4, 8, 16 workgroups per block (at most bz), LDS for the edges of
// ceil(bz / parts) layers
constexpr int NVAR_FACE = 5;
inline int face_parts(int v, int bz) { return std::min(1 << v, bz); }
template <class Rt>
void launch_face(int v, int nl, int nown, int bz, hipStream_t st, KPS P, int stage) {
    if (v <= 0) {
        hipLaunchKernelGGL(k_face<Rt>, dim3(nl * ((nown + TPB - 1) / TPB)), dim3(TPB), 0, st, P, stage);
        return;
    }
    const int ns = face_parts(v, bz);
    const size_t lds = sizeof(Rt) * (size_t)face_edges((bz + ns - 1) / ns);
    hipLaunchKernelGGL(k_face_edges<Rt>, dim3(nl * ns), dim3(TPB), lds, st, P, stage, ns);
}
inline void launch_c2p(int v, int nl, hipStream_t st, KPS P, int want_dt) {
    // four workgroups per 16^3 block: 7 904 wavefronts of 16 zones each on
    // the 3-level torus left the last of 2.6 rounds of resident waves half
    // empty
    const int parts = P.bz > 1 ? 4 : 1;
    switch (v) {
#define L_(i, tb, mw) \
    case i: hipLaunchKernelGGL((k_c2p<tb, mw>), dim3(nl * parts), dim3(tb), 0, st, P, want_dt, parts); break;
        SPH_VARIANTS(L_)
#undef L_
    }
}

// ---------------------------------------------------- host side
template <class T>
struct DB {
    T* p = nullptr;
    size_t n = 0;
    DB() = default;
    DB(const DB&) = delete;
    DB& operator=(const DB&) = delete;
    ~DB() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t m) {
        if (m <= n && p) return;
        release();
        SCHK(hipMalloc(&p, std::max<size_t>(1, m) * sizeof(T)));
        n = m;
    }
    void upload(const T* h, size_t m, hipStream_t s) {
        alloc(m);
        if (m) SCHK(hipMemcpyAsync(p, h, m * sizeof(T), hipMemcpyHostToDevice, s));
    }
};

struct HostRun {
    bool live = false;
    RunPar par{};
    bhs_grid grid{};
    std::vector<int> rls;      // RL indices of its levels
    std::vector<int> slots;    // its block slots, in the caller's block order
    std::vector<std::unique_ptr<DB<double>>> tabs;  // per level: gc, gf, dr
    std::vector<std::unique_ptr<DB<int>>> luts;
    std::unique_ptr<DB<int>> slotlist;                // its slots on the device
};

inline int pack_cell(int i, int j, int k) { return ((k + 2) << 16) | ((j + 2) << 8) | (i + 2); }

class Engine {
public:
    Lay L;
    long long cap = 0;
    hipStream_t stream = nullptr;
    DB<double> U, W, Us, Bf0, Bf1, F, icwork;  // icwork: initial-data work array
    DB<Blk> blk;
    DB<int> side;                 // [slot][6] same-level leaf across each side, -1
    std::vector<int> hside;
    DB<DRL> rl;
    DB<RunPar> runp;
    DB<RunCtr> ctr;
    DB<int> ghc, ghf, flf, owf;
    DB<int> ghp, ghp_off;         // ghost cells and side faces to fill, per sharing pattern
    int nghc = 0, nghf = 0, nflf = 0, nowf = 0, nflp = 0;
    DB<int> flp;
    DB<int> sgrp;                 // side-face groups of flf (offset, count) per direction and side
    std::vector<char> own_mask;   // [3][nf]: own faces (the evolved face field)
    DB<int> list;                 // active block slots (sorted by run)
    DB<int> runlist;              // active runs
    DB<unsigned int> qctr;
    DB<double> diagbuf;
    DB<unsigned long long> icmax;
    DB<double> icscale;
    std::vector<Blk> hblk;
    std::vector<DRL> hrl;
    std::vector<HostRun> runs;
    std::vector<int> free_slots;
    std::vector<int> act_runs;    // runs of the current active list
    bool act_hlld = false;        // every active run uses the HLLD solver
    std::vector<int> act_slots;
    bool mixed = false;           // FP32 fluxes and EMFs
    bool fastf = false;           // with mixed: FP32 with hardware reciprocal and square root (ffloat)
    int mode = 0;                 // 0 multi-kernel, 1 persistent, 2 slab (fused flux + update), 3 multi-kernel with
                                  // stage 1's con2prim fused into stage 0's update
    int wg_per_cu = 0;            // persistent grid: 0 = occupancy
    unsigned spin_limit = SPIN_LIMIT;  // polls of a grid-barrier wait (schedules 1 and 4)
    int var_flux = 0, var_update = 0, var_c2p = 1;  // launch variants (VAR_TB, VAR_MW)
    int var_face = 2;             // face-update variant (launch_face)
    bool xcd_remap = true;        // XCD-contiguous workgroup numbering
    bool tables_dirty = true;
    std::string err;
    // graph of the step for the current active set
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool use_graph = true;
    long long graph_key = -1;

    double* scratch_dev = nullptr;  // this device's copy of the static scratch array

    explicit Engine(int bz, long long capacity) {
        L = make_lay(bz);
        cap = capacity;
        SCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        // the device address of the static scratch array on the current GPU
        SCHK(hipGetSymbolAddress((void**)&scratch_dev, HIP_SYMBOL(scratch)));
        U.alloc((size_t)cap * 8 * L.ncell);
        W.alloc((size_t)cap * 8 * L.ncell);
        Us.alloc((size_t)cap * 5 * L.nint);
        Bf0.alloc((size_t)cap * 3 * L.nf);
        Bf1.alloc((size_t)cap * 3 * L.nf);
        F.alloc((size_t)cap * 3 * NFW * L.nff);  // FP64 words; the FP32 path uses the first half of each block's
        SCHK(hipMemsetAsync(Bf0.p, 0, Bf0.n * sizeof(double), stream));
        SCHK(hipMemsetAsync(Bf1.p, 0, Bf1.n * sizeof(double), stream));
        SCHK(hipMemsetAsync(F.p, 0, F.n * sizeof(double), stream));
        SCHK(hipMemsetAsync(W.p, 0, W.n * sizeof(double), stream));
        SCHK(hipMemsetAsync(U.p, 0, U.n * sizeof(double), stream));
        blk.alloc(cap);
        hblk.assign(cap, Blk{-1, 0, 0, 0});
        side.alloc((size_t)cap * 6);
        hside.assign((size_t)cap * 6, -1);
        for (long long s = cap - 1; s >= 0; s--) free_slots.push_back((int)s);
        qctr.alloc(16);
        SCHK(hipMemsetAsync(qctr.p, 0, 16 * sizeof(unsigned), stream));
        icmax.alloc(4);
        icscale.alloc(1);
        build_item_lists();
    }
    ~Engine() {
        graph_release();
        if (stream) (void)hipStreamDestroy(stream);
    }
    void graph_release() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
        graph_key = -1;
    }

    void build_item_lists() {
        std::vector<int> gc_, gf_, ff_, ow_;
        const int n[3] = {BX, BY, L.bz};
        for (int k = -G; k < L.bz + G; k++)
            for (int j = -G; j < BY + G; j++)
                for (int i = -G; i < BX + G; i++)
                    if (!(i >= 0 && i < BX && j >= 0 && j < BY && k >= 0 && k < L.bz)) gc_.push_back(pack_cell(i, j, k));
        // flux faces: the own faces of the three directions, then the side
        // faces grouped by direction and side (a block that shares a side
        // skips whole wavefronts of it)
        std::vector<int> sides[3][6];
        for (int d = 0; d < 3; d++) {
            int lo[3], hi[3];
            for (int e = 0; e < 3; e++) { lo[e] = e == d ? 0 : -1; hi[e] = n[e]; }
            for (int k = lo[2]; k <= hi[2]; k++)
                for (int j = lo[1]; j <= hi[1]; j++)
                    for (int i = lo[0]; i <= hi[0]; i++) {
                        const int c[3] = {i, j, k};
                        // a face outside the block in both transverse
                        // directions enters no EMF of an own face
                        int out = 0, sd = -1;
                        for (int e = 0; e < 3; e++)
                            if (e != d && (c[e] < 0 || c[e] >= n[e])) {
                                out++;
                                sd = 2 * e + (c[e] < 0 ? 0 : 1);
                            }
                        if (out == 2) continue;
                        const int x = (d << 24) | pack_cell(i, j, k);
                        if (out == 0) {
                            ow_.push_back(x);
                            ff_.push_back(x);
                        } else {
                            sides[d][sd].push_back(x);
                        }
                    }
        }
        std::vector<int> sg(36);
        for (int d = 0; d < 3; d++)
            for (int e = 0; e < 6; e++) {
                sg[2 * (6 * d + e)] = (int)ff_.size();
                sg[2 * (6 * d + e) + 1] = (int)sides[d][e].size();
                ff_.insert(ff_.end(), sides[d][e].begin(), sides[d][e].end());
                gf_.insert(gf_.end(), sides[d][e].begin(), sides[d][e].end());
            }
        sgrp.upload(sg.data(), sg.size(), stream);
        // per sharing pattern m (bit e: side e shared with a same-level
        // leaf): the ghost cells some stencil reads, then the side faces of
        // the unshared sides (bit 30)
        {
            std::vector<int> gp, go(65, 0);
            for (int m = 0; m < 64; m++) {
                go[m] = (int)gp.size();
                for (int x : gc_) {
                    const int i = (x & 0xff) - 2, j = ((x >> 8) & 0xff) - 2, k = ((x >> 16) & 0xff) - 2;
                    if (ghost_read_mask(m, i, j, k, L.bz)) gp.push_back(x);
                }
                for (int d = 0; d < 3; d++)
                    for (int e = 0; e < 6; e++)
                        if (!((m >> e) & 1))
                            for (int x : sides[d][e]) gp.push_back(x | (1 << 30));
            }
            go[64] = (int)gp.size();
            ghp.upload(gp.data(), gp.size(), stream);
            ghp_off.upload(go.data(), go.size(), stream);
        }
        own_mask.assign((size_t)3 * L.nf, 0);
        for (int x : ow_) {
            int i, j, k;
            const int d = x >> 24;
            i = (x & 0xff) - 2; j = ((x >> 8) & 0xff) - 2; k = ((x >> 16) & 0xff) - 2;
            own_mask[(size_t)d * L.nf + fidx(d, i, j, k)] = 1;
        }
        // theta and phi face lines (transverse ranges as the face ranges;
        // those inside the block first), then the r faces in the order above
        std::vector<int> fl_, fls_;
        for (int k = -1; k <= L.bz; k++)
            for (int i = -1; i <= BX; i++) {
                const int oi = i < 0 || i >= BX, ok = k < 0 || k >= L.bz;
                if (oi && ok) continue;
                ((oi || ok) ? fls_ : fl_).push_back((1 << 26) | (1 << 24) | pack_cell(i, 0, k));
            }
        for (int j = -1; j <= BY; j++)
            for (int i = -1; i <= BX; i++) {
                const int oi = i < 0 || i >= BX, oj = j < 0 || j >= BY;
                if (oi && oj) continue;
                ((oi || oj) ? fls_ : fl_).push_back((1 << 26) | (2 << 24) | pack_cell(i, j, 0));
            }
        fl_.insert(fl_.end(), fls_.begin(), fls_.end());
        for (int x : ff_)
            if ((x >> 24) == 0) fl_.push_back(x);
        nflp = (int)fl_.size();
        flp.upload(fl_.data(), fl_.size(), stream);
        nghc = (int)gc_.size(); nghf = (int)gf_.size(); nflf = (int)ff_.size(); nowf = (int)ow_.size();
        ghc.upload(gc_.data(), gc_.size(), stream);
        ghf.upload(gf_.data(), gf_.size(), stream);
        flf.upload(ff_.data(), ff_.size(), stream);
        owf.upload(ow_.data(), ow_.size(), stream);
        SCHK(hipStreamSynchronize(stream));
    }

    KPS kp(const int* lst, int nlst) {
        KPS P;
        P.bz = L.bz; P.sz = L.sz; P.ncell = L.ncell; P.nint = L.nint; P.nf = L.nf; P.nff = L.nff;
        P.spin = spin_limit;
        P.blk = blk.p; P.rl = rl.p; P.run = runp.p; P.ctr = ctr.p;
        P.U = U.p; P.W = W.p; P.Us = Us.p; P.Bf0 = Bf0.p; P.Bf1 = Bf1.p; P.F = F.p;
        P.list = lst; P.nlist = nlst;
        P.ghc = ghc.p; P.nghc = nghc; P.ghf = ghf.p; P.nghf = nghf;
        P.ghp = ghp.p; P.ghp_off = ghp_off.p;
        P.flf = flf.p; P.nflf = nflf; P.owf = owf.p; P.nowf = nowf;
        P.flp = flp.p; P.nflp = nflp;
        P.qctr = qctr.p;
        P.xcd = xcd_remap ? 1 : 0;
        P.side = side.p;
        P.sgrp = sgrp.p;
        P.reuse = mode != 1 ? 1 : 0;  // the queue schedule (4) separates its phases by grid barriers
        return P;
    }

    // ---- runs
    int add_run(const bhs_physics& ph, const bhs_grid& g, int nb, const int* blocks) {
        if (nb <= 0) throw std::invalid_argument("a run needs at least one block");
        if (g.nlevels < 1 || g.nlevels > 8) throw std::invalid_argument("nlevels must be 1..8");
        if (L.bz == 1 && g.nlevels != 1) throw std::invalid_argument("axisymmetric runs have one level");
        if (g.nr % BX || g.nth % BY || (L.bz > 1 && g.nph % L.bz)) throw std::invalid_argument("grid not a whole number of blocks");
        if (!(g.r_in > 0.0) || !(g.r_out > g.r_in)) throw std::invalid_argument("need 0 < r_in < r_out");
        if (ph.riemann != BHS_HLL && ph.riemann != BHS_HLLD) throw std::invalid_argument("riemann must be BHS_HLL or BHS_HLLD");
        if ((long long)free_slots.size() < nb) throw std::length_error("not enough free block slots");
        int id = -1;
        for (size_t r = 0; r < runs.size(); r++)
            if (!runs[r].live) { id = (int)r; break; }
        if (id < 0) { id = (int)runs.size(); runs.emplace_back(); }
        HostRun& R = runs[id];
        R = HostRun();
        R.live = true;
        R.grid = g;
        R.par.a = ph.a; R.par.Q = ph.Q; R.par.gam = ph.gamma; R.par.cfl = ph.cfl;
        R.par.rho_fl = ph.rho_floor; R.par.eps_fl = ph.eps_floor; R.par.sigma_max = ph.sigma_max;
        R.par.w_max = ph.gamma_max; R.par.nlev = g.nlevels; R.par.nz = L.bz;
        R.par.riemann = ph.riemann;
        // levels and their tables
        std::vector<std::vector<int>> lut(g.nlevels);
        long long z0 = 0;
        std::vector<DRL> lv(g.nlevels);
        for (int l = 0; l < g.nlevels; l++) {
            DRL v{};
            const int f = 1 << l;
            v.run = id; v.level = l;
            v.nr = g.nr * f; v.nth = g.nth * f; v.nph = L.bz == 1 ? 1 : g.nph * f;
            v.nbr = v.nr / BX; v.nbth = v.nth / BY; v.nbph = v.nph / L.bz;
            v.npt = (v.nr + 5) * (v.nth + 5);
            v.zone0 = z0;
            z0 += (long long)v.nr * v.nth * v.nph;
            v.dth = M_PI / v.nth;
            v.dph = 2.0 * M_PI / v.nph;
            {
                // the zone row whose r faces bracket the outer horizon
                const double disc = 1.0 - ph.a * ph.a - ph.Q * ph.Q;
                const double rh = 1.0 + std::sqrt(disc > 0.0 ? disc : 0.0);
                const double dl = (std::log(g.r_out) - std::log(g.r_in)) / v.nr;
                const int ih = (int)std::floor((std::log(rh) - std::log(g.r_in)) / dl);
                v.ih = (ih >= 0 && ih < v.nr) ? ih : -1;
            }
            lut[l].assign((size_t)v.nbr * v.nbth * v.nbph, -1);
            lv[l] = v;
        }
        for (int q = 0; q < nb; q++) {
            const int l = blocks[4 * q], bi = blocks[4 * q + 1], bj = blocks[4 * q + 2], bk = blocks[4 * q + 3];
            if (l < 0 || l >= g.nlevels) throw std::invalid_argument("block level out of range");
            const DRL& v = lv[l];
            if (bi < 0 || bi >= v.nbr || bj < 0 || bj >= v.nbth || bk < 0 || bk >= v.nbph)
                throw std::invalid_argument("block coordinates out of range");
            int& e = lut[l][((size_t)bk * v.nbth + bj) * v.nbr + bi];
            if (e >= 0) throw std::invalid_argument("block listed twice");
            e = q;
        }
        check_cover(lv, lut, blocks, nb);
        // slots
        R.slots.resize(nb);
        for (int q = 0; q < nb; q++) {
            R.slots[q] = free_slots.back();
            free_slots.pop_back();
        }
        for (int l = 0; l < g.nlevels; l++)
            for (int& e : lut[l])
                if (e >= 0) e = R.slots[e];
        R.slotlist = std::make_unique<DB<int>>();
        R.slotlist->upload(R.slots.data(), R.slots.size(), stream);
        const double lnr0 = std::log(g.r_in);
        for (int l = 0; l < g.nlevels; l++) {
            DRL& v = lv[l];
            const double dlnr = (std::log(g.r_out) - std::log(g.r_in)) / v.nr;
            std::vector<double> gc, gf, dr;
            build_level_tables(ph.a, ph.Q, ph.rho_floor, ph.eps_floor, v.nr, v.nth, lnr0, dlnr, gc, gf, dr);
            std::vector<double> all;
            all.reserve(gc.size() + gf.size() + dr.size());
            all.insert(all.end(), gc.begin(), gc.end());
            all.insert(all.end(), gf.begin(), gf.end());
            all.insert(all.end(), dr.begin(), dr.end());
            auto t = std::make_unique<DB<double>>();
            t->upload(all.data(), all.size(), stream);
            v.gc = t->p;
            v.gf = t->p + gc.size();
            v.dr = t->p + gc.size() + gf.size();
            auto lu = std::make_unique<DB<int>>();
            lu->upload(lut[l].data(), lut[l].size(), stream);
            v.lut = lu->p;
            SCHK(hipStreamSynchronize(stream));
            R.tabs.push_back(std::move(t));
            R.luts.push_back(std::move(lu));
        }
        // RL records
        for (int l = 0; l < g.nlevels; l++) {
            int slot = -1;
            for (size_t x = 0; x < hrl.size(); x++)
                if (hrl[x].run < 0) { slot = (int)x; break; }
            if (slot < 0) { slot = (int)hrl.size(); hrl.push_back(DRL{}); }
            hrl[slot] = lv[l];
            R.rls.push_back(slot);
        }
        for (int l = 0; l < g.nlevels; l++) {
            hrl[R.rls[l]].coarser = l > 0 ? R.rls[l - 1] : -1;
            hrl[R.rls[l]].finer = l + 1 < g.nlevels ? R.rls[l + 1] : -1;
        }
        R.par.rl0 = R.rls[0];
        for (int q = 0; q < nb; q++)
            hblk[R.slots[q]] = Blk{R.rls[blocks[4 * q]], blocks[4 * q + 1], blocks[4 * q + 2], blocks[4 * q + 3]};
        // the same-level leaf across each side (r and theta: none beyond the
        // domain or the axis; phi: periodic)
        for (int q = 0; q < nb; q++) {
            const int l = blocks[4 * q], bi = blocks[4 * q + 1], bj = blocks[4 * q + 2], bk = blocks[4 * q + 3];
            const DRL& v = lv[l];
            auto at = [&](int i, int j, int k) { return lut[l][((size_t)k * v.nbth + j) * v.nbr + i]; };
            int* sd = &hside[(size_t)R.slots[q] * 6];
            sd[0] = bi > 0 ? at(bi - 1, bj, bk) : -1;
            sd[1] = bi + 1 < v.nbr ? at(bi + 1, bj, bk) : -1;
            sd[2] = bj > 0 ? at(bi, bj - 1, bk) : -1;
            sd[3] = bj + 1 < v.nbth ? at(bi, bj + 1, bk) : -1;
            // phi sides of an axisymmetric block (one cell deep) are its own
            // single layer; a stencil two cells deep cannot be mapped into it
            sd[4] = L.bz > 2 ? at(bi, bj, (bk + v.nbph - 1) % v.nbph) : -1;
            sd[5] = L.bz > 2 ? at(bi, bj, (bk + 1) % v.nbph) : -1;
        }
        tables_dirty = true;
        return id;
    }

    // Every ghost cell of every block must resolve: a leaf at its own level,
    // a coarser leaf (one level down), or finer leaves (one level up) whose
    // eight children are all leaves.
    static void check_cover(const std::vector<DRL>& lv, const std::vector<std::vector<int>>& lut, const int* blocks,
                            int nb) {
        auto leaf = [&](int l, int gi, int gj, int gk) -> bool {
            const DRL& v = lv[l];
            if (gj < 0) gj = -1 - gj;
            if (gj >= v.nth) gj = 2 * v.nth - 1 - gj;
            gi = std::min(std::max(gi, 0), v.nr - 1);
            gk = ((gk % v.nph) + v.nph) % v.nph;
            const int bz = v.nph / v.nbph;
            return lut[l][((size_t)(gk / bz) * v.nbth + gj / BY) * v.nbr + gi / BX] >= 0;
        };
        for (int q = 0; q < nb; q++) {
            const int l = blocks[4 * q];
            const DRL& v = lv[l];
            const int bz = v.nph / v.nbph;
            // sample the block's neighbourhood one and two cells outside its
            // faces, edges and corners, at transverse offsets 0, 8 and 15
            // (every block one level finer that a face touches)
            const int offs[3] = {0, 8, 15};
            for (int dk = -1; dk <= 1; dk++)
                for (int dj = -1; dj <= 1; dj++)
                    for (int di = -1; di <= 1; di++)
                        for (int oc = 0; oc < 27; oc++) {
                        const int oi = offs[oc % 3], oj = offs[(oc / 3) % 3], ok = bz > 1 ? offs[oc / 9] : 0;
                        if ((di != 0 && oc % 3) || (dj != 0 && (oc / 3) % 3) || ((dk != 0 || bz == 1) && oc / 9)) continue;
                        const int gi = blocks[4 * q + 1] * BX + (di < 0 ? -1 : di > 0 ? BX : oi);
                        const int gj = blocks[4 * q + 2] * BY + (dj < 0 ? -1 : dj > 0 ? BY : oj);
                        const int gk = blocks[4 * q + 3] * bz + (dk < 0 ? -1 : dk > 0 ? bz : ok);
                        for (int s = 0; s < 2; s++) {  // both cells of the two ghost layers' coarse cell
                            const int ggi = gi + (di < 0 ? -s : di > 0 ? s : 0);
                            const int ggj = gj + (dj < 0 ? -s : dj > 0 ? s : 0);
                            const int ggk = gk + (dk < 0 ? -s : dk > 0 ? s : 0);
                            if (leaf(l, ggi, ggj, ggk)) continue;
                            if (l > 0 && leaf(l - 1, ggi >> 1, ggj >> 1, ggk >> 1)) continue;
                            bool fine = l + 1 < (int)lv.size();
                            if (fine) {
                                int Gi = ggi, Gj = ggj, Gk = ggk;
                                if (Gj < 0) Gj = -1 - Gj;
                                if (Gj >= v.nth) Gj = 2 * v.nth - 1 - Gj;
                                Gi = std::min(std::max(Gi, 0), v.nr - 1);
                                Gk = ((Gk % v.nph) + v.nph) % v.nph;
                                for (int c = 0; c < 8 && fine; c++)
                                    fine = leaf(l + 1, 2 * Gi + (c & 1), 2 * Gj + ((c >> 1) & 1), 2 * Gk + (c >> 2));
                            }
                            if (!fine)
                                throw std::invalid_argument("block table: a ghost region has no leaf at the same, the "
                                                            "next coarser or the next finer level (levels must be "
                                                            "nested and 2:1 balanced)");
                        }
                    }
        }
    }

    void remove_run(int id) {
        HostRun& R = runs.at(id);
        if (!R.live) throw std::invalid_argument("no such run");
        SCHK(hipStreamSynchronize(stream));
        for (int s : R.slots) {
            hblk[s] = Blk{-1, 0, 0, 0};
            for (int e = 0; e < 6; e++) hside[(size_t)s * 6 + e] = -1;
            free_slots.push_back(s);
        }
        for (int x : R.rls) hrl[x].run = -1;
        R = HostRun();
        tables_dirty = true;
        graph_release();
    }

    void sync_tables() {
        if (!tables_dirty) return;
        SCHK(hipStreamSynchronize(stream));
        blk.upload(hblk.data(), hblk.size(), stream);
        side.upload(hside.data(), hside.size(), stream);
        rl.upload(hrl.data(), std::max<size_t>(1, hrl.size()), stream);
        std::vector<RunPar> rp(std::max<size_t>(1, runs.size()));
        for (size_t r = 0; r < runs.size(); r++) rp[r] = runs[r].par;
        runp.upload(rp.data(), rp.size(), stream);
        if (ctr.n < rp.size()) {
            ctr.alloc(rp.size() + 16);
            SCHK(hipMemsetAsync(ctr.p, 0, ctr.n * sizeof(RunCtr), stream));
        }
        SCHK(hipStreamSynchronize(stream));
        tables_dirty = false;
        graph_release();
        act_runs.clear();
    }

    void set_active(const std::vector<int>& rr) {
        sync_tables();
        if (rr == act_runs && list.p) return;
        std::vector<int> sl;
        for (int r : rr) {
            if (r < 0 || r >= (int)runs.size() || !runs[r].live) throw std::invalid_argument("no such run");
            sl.insert(sl.end(), runs[r].slots.begin(), runs[r].slots.end());
        }
        list.upload(sl.data(), sl.size(), stream);
        runlist.upload(rr.data(), rr.size(), stream);
        SCHK(hipStreamSynchronize(stream));
        act_runs = rr;
        act_slots = sl;
        act_hlld = true;
        for (int r : rr) act_hlld = act_hlld && runs[r].par.riemann == BHS_HLLD;
        graph_release();
    }

    // ---- step
    // grid of the queue schedule: every workgroup resident (its occupancy
    // with the face phase's LDS), at most one per queue item of the
    // smallest phase
    template <class Rt>
    int queue_grid(size_t lds, bool w3) {
        int occ = 0;
        if (act_hlld && w3) SCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_queue<Rt, 1, 3>, TPB, lds));
        else if (act_hlld) SCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_queue<Rt, 1, 1>, TPB, lds));
        else if (w3) SCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_queue<Rt, 0, 3>, TPB, lds));
        else SCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_queue<Rt, 0, 1>, TPB, lds));
        int dev = 0, ncu = 0;
        SCHK(hipGetDevice(&dev));
        SCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        int per = occ;
        if (wg_per_cu > 0) per = std::min(per, wg_per_cu);
        per = std::max(per, 1);
        return std::max(1, std::min(ncu * per, (int)act_slots.size() * 2));
    }
    int persistent_grid() {
        int occ = 0;
        if (mixed && fastf) SCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_stage<ffloat>, TPB, 0));
        else if (mixed) SCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_stage<float>, TPB, 0));
        else SCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_stage<double>, TPB, 0));
        int dev = 0, ncu = 0;
        SCHK(hipGetDevice(&dev));
        SCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        int per = occ;
        if (wg_per_cu > 0) per = std::min(per, wg_per_cu);
        per = std::max(per, 1);
        return std::max(1, std::min(ncu * per, (int)act_slots.size()));
    }

    // Per-kernel timing with HIP events on the engine's stream (eager
    // launches only): every launch, or the launches of one kernel name.
    bool profile = false;
    std::string prof_only;
    struct PendEv {
        std::string name;
        hipEvent_t a, b;
    };
    std::vector<PendEv> pend;
    std::vector<hipEvent_t> evpool;
    std::map<std::string, std::pair<long long, double>> prof_acc;
    hipEvent_t ev_get() {
        if (!evpool.empty()) {
            hipEvent_t e = evpool.back();
            evpool.pop_back();
            return e;
        }
        hipEvent_t e;
        SCHK(hipEventCreate(&e));
        return e;
    }
    template <class Fn>
    void tm(const char* name, Fn&& fn) {
        if (!profile || (!prof_only.empty() && prof_only != name)) {
            fn();
            return;
        }
        PendEv p{name, ev_get(), ev_get()};
        SCHK(hipEventRecord(p.a, stream));
        fn();
        SCHK(hipEventRecord(p.b, stream));
        pend.push_back(p);
    }
    void prof_collect() {
        for (auto& p : pend) {
            float ms = 0.f;
            SCHK(hipEventElapsedTime(&ms, p.a, p.b));
            auto& acc = prof_acc[p.name];
            acc.first++;
            acc.second += ms;
            evpool.push_back(p.a);
            evpool.push_back(p.b);
        }
        pend.clear();
    }

    template <class Rt>
    void enqueue_step() {
        const int nl = (int)act_slots.size();
        const int nr = (int)act_runs.size();
        KPS P = kp(list.p, nl);
        tm("k_begin", [&] {
            hipLaunchKernelGGL(k_begin, dim3((std::max(nr, 4) + 63) / 64), dim3(64), 0, stream, ctr.p, runlist.p, nr,
                               qctr.p);
        });
        for (int stage = 0; stage < 2; stage++) {
            if (mode == 1) {
                if (stage == 1)
                    tm("k_begin", [&] {
                        hipLaunchKernelGGL(k_begin, dim3(1), dim3(64), 0, stream, ctr.p, runlist.p, 0, qctr.p);
                    });
                const int ng = persistent_grid();
                tm("k_stage", [&] { hipLaunchKernelGGL(k_stage<Rt>, dim3(ng), dim3(TPB), 0, stream, P, stage); });
                continue;
            }
            if (mode == 4) {
                if (stage == 1)
                    tm("k_begin", [&] {
                        hipLaunchKernelGGL(k_begin, dim3(1), dim3(64), 0, stream, ctr.p, runlist.p, 0, qctr.p);
                    });
                const int pe = face_parts(var_face > 0 ? var_face : 2, L.bz);
                const size_t lds = sizeof(Rt) * (size_t)face_edges((L.bz + pe - 1) / pe);
                // the flux variant's occupancy hint: 1 (registers as the
                // compiler needs) or 3 waves per SIMD
                const bool w3 = VAR_MW[var_flux % NVAR] != 1;
                const int ng = queue_grid<Rt>(lds, w3);
                tm("k_queue", [&] {
                    if (act_hlld && w3)
                        hipLaunchKernelGGL((k_queue<Rt, 1, 3>), dim3(ng), dim3(TPB), lds, stream, P, stage, pe);
                    else if (act_hlld)
                        hipLaunchKernelGGL((k_queue<Rt, 1, 1>), dim3(ng), dim3(TPB), lds, stream, P, stage, pe);
                    else if (w3)
                        hipLaunchKernelGGL((k_queue<Rt, 0, 3>), dim3(ng), dim3(TPB), lds, stream, P, stage, pe);
                    else
                        hipLaunchKernelGGL((k_queue<Rt, 0, 1>), dim3(ng), dim3(TPB), lds, stream, P, stage, pe);
                });
                continue;
            }
            // schedule 3: stage 1's con2prim ran inside stage 0's update
            if (!(mode == 3 && stage == 1))
                tm("k_c2p", [&] { launch_c2p(var_c2p, nl, stream, P, stage == 0 ? 1 : 0); });
            const int buf = stage;
            tm("k_ghost", [&] {
                const int parts = L.bz > 1 ? 4 : 1;
                hipLaunchKernelGGL(k_ghost, dim3(nl * parts), dim3(TPB), 0, stream, P, buf, parts);
            });
            if (mode == 2) {
                // the side faces (after the own faces in the face list), then
                // fluxes and hydro update fused per block
                // the flux variant v picks how many workgroups share a
                // block's layers: 2^(v mod 5), at most bz
                const int ns = std::min(1 << (var_flux % 5), L.bz);
                tm("k_slab", [&] {
                    if (act_hlld)
                        hipLaunchKernelGGL((k_slab<Rt, 1>), dim3(nl * ns), dim3(TPB), 0, stream, P, buf, stage, ns);
                    else
                        hipLaunchKernelGGL((k_slab<Rt, 0>), dim3(nl * ns), dim3(TPB), 0, stream, P, buf, stage, ns);
                });
                tm("k_face", [&] {
                    launch_face<Rt>(var_face, nl, nowf, L.bz, stream, P, stage);
                });
                tm("k_cellb", [&] {
                    hipLaunchKernelGGL(k_cellb, dim3(nl * ((L.nint + TPB - 1) / TPB)), dim3(TPB), 0, stream, P, stage);
                });
                continue;
            }
            tm("k_flux", [&] { launch_flux<Rt>(var_flux, nl, nflf, nflp, stream, P, buf, act_hlld); });
            tm("k_face", [&] {
                launch_face<Rt>(var_face, nl, nowf, L.bz, stream, P, stage);
            });
            if (mode == 3 && stage == 0)
                tm("k_update_c2p", [&] { launch_update_c2p<Rt>(var_update, nl, L.nint, stream, P); });
            else
                tm("k_update", [&] { launch_update<Rt>(var_update, nl, L.nint, stream, P, stage); });
        }
        SCHK(hipGetLastError());
    }
    void enqueue() {
        if (mixed && fastf) enqueue_step<ffloat>();
        else if (mixed) enqueue_step<float>();
        else enqueue_step<double>();
    }
    long long key() const {
        return (long long)(mixed ? 1 : 0) | ((long long)(fastf ? 1 : 0) << 1) | ((long long)mode << 2) |
               ((long long)wg_per_cu << 5) |
               ((long long)var_flux << 13) | ((long long)var_update << 19) | ((long long)var_c2p << 25) |
               ((long long)(xcd_remap ? 1 : 0) << 31) | ((long long)var_face << 32) |
               ((long long)(spin_limit != SPIN_LIMIT ? 1 : 0) << 36);
    }
    void step() {
        if (act_slots.empty()) throw std::invalid_argument("no active runs");
        if (!use_graph || profile) {
            enqueue();
            return;
        }
        if (!gexec || graph_key != key()) {
            graph_release();
            SCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            enqueue();
            SCHK(hipStreamEndCapture(stream, &graph));
            SCHK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
            graph_key = key();
        }
        SCHK(hipGraphLaunch(gexec, stream));
    }
    void sync() {
        SCHK(hipGetLastError());
        SCHK(hipStreamSynchronize(stream));
        if (!pend.empty()) prof_collect();
    }
    void check_barrier() {
        if (mode != 1 && mode != 4) return;
        unsigned f = 0;
        SCHK(hipMemcpyAsync(&f, qctr.p + QFLAG, sizeof f, hipMemcpyDeviceToHost, stream));
        SCHK(hipStreamSynchronize(stream));
        if (f) {
            // cleared here (never inside a captured step), so the next step
            // starts without it
            SCHK(hipMemsetAsync(qctr.p + QFLAG, 0, sizeof(unsigned), stream));
            SCHK(hipStreamSynchronize(stream));
            throw std::runtime_error("persistent stage: grid barrier timed out (workgroups not co-resident)");
        }
    }
    void read_ctr(std::vector<RunCtr>& out) {
        out.resize(runs.size());
        if (!runs.empty()) SCHK(hipMemcpyAsync(out.data(), ctr.p, runs.size() * sizeof(RunCtr), hipMemcpyDeviceToHost, stream));
        SCHK(hipStreamSynchronize(stream));
    }

    // con2prim of one run's active zones (W of the current U), no dt
    void c2p_run(int r) {
        set_active({r});
        KPS P = kp(list.p, (int)act_slots.size());
        launch_c2p(var_c2p, (int)act_slots.size(), stream, P, 0);
        sync();
    }

    void init_torus(int r, const bhs_torus& t, double* info) {
        set_active({r});
        const HostRun& R = runs[r];
        TorusDev D;
        D.T.a = R.par.a; D.T.gam = R.par.gam; D.T.r_in = t.r_in; D.T.r_max = t.r_max; D.T.kappa = t.kappa;
        D.T.l = fm_l(R.par.a, t.r_max);
        D.T.lnh_in = fm_lnh_raw<HostCR>(D.T, t.r_in, 1.0, 0.0);
        D.beta = t.beta; D.acut = t.a_cut; D.pamp = t.perturbation;
        D.st = U128{t.pcg_state_hi, t.pcg_state_lo};
        D.inc = U128{t.pcg_inc_hi, t.pcg_inc_lo};
        const int nl = (int)act_slots.size();
        icwork.alloc((size_t)cap * L.ncell);
        KPS P = kp(list.p, nl);
        // clear the run's state: face fields (incl. phi faces), W
        for (int s : R.slots) {
            SCHK(hipMemsetAsync(Bf0.p + (size_t)s * 3 * L.nf, 0, 3 * L.nf * sizeof(double), stream));
            SCHK(hipMemsetAsync(Bf1.p + (size_t)s * 3 * L.nf, 0, 3 * L.nf * sizeof(double), stream));
            SCHK(hipMemsetAsync(W.p + (size_t)s * 8 * L.ncell, 0, 8 * L.ncell * sizeof(double), stream));
            SCHK(hipMemsetAsync(U.p + (size_t)s * 8 * L.ncell, 0, 8 * L.ncell * sizeof(double), stream));
        }
        SCHK(hipMemsetAsync(icmax.p, 0, 4 * sizeof(unsigned long long), stream));
        hipLaunchKernelGGL(k_ic_torus, dim3(nl * ((L.ncell + TPB - 1) / TPB)), dim3(TPB), 0, stream, P, D, icwork.p,
                           icmax.p);
        hipLaunchKernelGGL(k_ic_faces, dim3(nl * ((nowf + TPB - 1) / TPB)), dim3(TPB), 0, stream, P, D, icwork.p,
                           icmax.p);
        hipLaunchKernelGGL(k_ic_cells, dim3(nl * ((L.nint + TPB - 1) / TPB)), dim3(TPB), 0, stream, P, D, icwork.p,
                           icmax.p);
        hipLaunchKernelGGL(k_ic_scale, dim3(nl * ((3 * L.nf + TPB - 1) / TPB)), dim3(TPB), 0, stream, P, icmax.p,
                           t.beta, icscale.p);
        hipLaunchKernelGGL(k_ic_final, dim3(nl * ((L.nint + TPB - 1) / TPB)), dim3(TPB), 0, stream, P, D, icwork.p);
        SCHK(hipGetLastError());
        unsigned long long mx[4];
        double sc = 0.0;
        SCHK(hipMemcpyAsync(mx, icmax.p, sizeof mx, hipMemcpyDeviceToHost, stream));
        SCHK(hipMemcpyAsync(&sc, icscale.p, sizeof sc, hipMemcpyDeviceToHost, stream));
        sync();
        if (info) {
            for (int q = 0; q < 3; q++) std::memcpy(&info[q], &mx[q], sizeof(double));
            info[3] = sc;
        }
        double rm;
        std::memcpy(&rm, &mx[0], sizeof rm);
        if (!(rm > 0.0)) throw std::domain_error("no torus zone on the grid");
    }
    struct HostCR {
        static double log(double x) { return bh::cr_log(x); }
        static double exp(double x) { return bh::cr_exp(x); }
        static double pow(double x, double y) { return bh::cr_pow(x, y); }
    };

    void get_state(int r, double* hU, double* hW, double* hB) {
        const HostRun& R = runs.at(r);
        sync();
        for (size_t q = 0; q < R.slots.size(); q++) {
            const int s = R.slots[q];
            if (hU) SCHK(hipMemcpyAsync(hU + q * 8 * L.ncell, U.p + (size_t)s * 8 * L.ncell, 8 * L.ncell * sizeof(double), hipMemcpyDeviceToHost, stream));
            if (hW) SCHK(hipMemcpyAsync(hW + q * 8 * L.ncell, W.p + (size_t)s * 8 * L.ncell, 8 * L.ncell * sizeof(double), hipMemcpyDeviceToHost, stream));
            if (hB) SCHK(hipMemcpyAsync(hB + q * 3 * L.nf, Bf0.p + (size_t)s * 3 * L.nf, 3 * L.nf * sizeof(double), hipMemcpyDeviceToHost, stream));
        }
        sync();
        // the state of a block is its active zones and its own faces; ghost
        // cells and ghost faces are refilled from their owners every stage
        // (as far as the step reads them) and are exported as 0
        for (size_t q = 0; q < R.slots.size(); q++) {
            for (int k = -G; k < L.bz + G; k++)
                for (int j = -G; j < BY + G; j++)
                    for (int i = -G; i < BX + G; i++) {
                        if (i >= 0 && i < BX && j >= 0 && j < BY && k >= 0 && k < L.bz) continue;
                        for (int w = 0; w < 8; w++) {
                            const size_t x = (q * 8 + w) * L.ncell + cidx(i, j, k);
                            if (hU) hU[x] = 0.0;
                            if (hW) hW[x] = 0.0;
                        }
                    }
            if (hB)
                for (size_t x = 0; x < (size_t)3 * L.nf; x++)
                    if (!own_mask[x]) hB[q * 3 * L.nf + x] = 0.0;
        }
    }
    void set_state(int r, const double* hU, const double* hW, const double* hB) {
        const HostRun& R = runs.at(r);
        sync();
        for (size_t q = 0; q < R.slots.size(); q++) {
            const int s = R.slots[q];
            if (hU) SCHK(hipMemcpyAsync(U.p + (size_t)s * 8 * L.ncell, hU + q * 8 * L.ncell, 8 * L.ncell * sizeof(double), hipMemcpyHostToDevice, stream));
            if (hW) SCHK(hipMemcpyAsync(W.p + (size_t)s * 8 * L.ncell, hW + q * 8 * L.ncell, 8 * L.ncell * sizeof(double), hipMemcpyHostToDevice, stream));
            if (hB) SCHK(hipMemcpyAsync(Bf0.p + (size_t)s * 3 * L.nf, hB + q * 3 * L.nf, 3 * L.nf * sizeof(double), hipMemcpyHostToDevice, stream));
        }
        sync();
    }

    // diagnostics without any allocation: the run's persistent slot list,
    // the static scratch array, one stream-ordered copy out
    void diagnostics_static(int r, double* out) {
        if (r < 0 || r >= (int)runs.size() || !runs[r].live) throw std::invalid_argument("no such run");
        if (tables_dirty) throw std::domain_error("the engine's block tables are not on the device yet (step or "
                                                  "initialise a run first)");
        const HostRun& R = runs[r];
        const int nl = (int)R.slots.size();
        if (nl > DIAG_MAX_BLOCKS) throw std::invalid_argument("run has more blocks than the static scratch holds");
        KPS P = kp(R.slotlist->p, nl);
        double* part = scratch_dev;
        double* fin = scratch_dev + (size_t)6 * DIAG_MAX_BLOCKS;
        hipLaunchKernelGGL(k_diag, dim3(nl), dim3(TPB), 0, stream, P, part);
        hipLaunchKernelGGL(k_diag_seq, dim3(1), dim3(64), 0, stream, (const double*)part, nl, fin);
        SCHK(hipGetLastError());
        double h[6];
        RunCtr c;
        SCHK(hipMemcpyAsync(h, fin, sizeof h, hipMemcpyDeviceToHost, stream));
        SCHK(hipMemcpyAsync(&c, ctr.p + r, sizeof c, hipMemcpyDeviceToHost, stream));
        SCHK(hipStreamSynchronize(stream));
        out[0] = h[0]; out[1] = h[1]; out[2] = h[2]; out[3] = h[3];
        out[4] = (double)c.c2p_fail;
        out[5] = (double)c.floors;
        out[6] = h[4];
        out[7] = h[5];
    }

    void diagnostics(int r, double* out) {
        set_active({r});
        const int nl = (int)act_slots.size();
        diagbuf.alloc((size_t)nl * 6);
        KPS P = kp(list.p, nl);
        hipLaunchKernelGGL(k_diag, dim3(nl), dim3(TPB), 0, stream, P, diagbuf.p);
        std::vector<double> h((size_t)nl * 6);
        SCHK(hipMemcpyAsync(h.data(), diagbuf.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
        std::vector<RunCtr> c;
        read_ctr(c);
        double ms = 0.0, en = 0.0, dm = 0.0, sc = 0.0, md = 0.0, ph = 0.0;
        for (int q = 0; q < nl; q++) {
            ms += h[6 * q];
            en += h[6 * q + 1];
            dm = std::max(dm, h[6 * q + 2]);
            sc = std::max(sc, h[6 * q + 3]);
            md += h[6 * q + 4];
            ph += h[6 * q + 5];
        }
        out[0] = ms; out[1] = en; out[2] = dm; out[3] = sc;
        out[4] = (double)c[r].c2p_fail;
        out[5] = (double)c[r].floors;
        out[6] = md;
        out[7] = ph;
    }

    // the analyser quantities of a run: out[10] = thermal energy, magnetic
    // energy, Ldot, L_BZ, eta_BZ = L_BZ/|Mdot|, phi_BH = Phi_B/sqrt(|Mdot|),
    // P_beta = p/p_mag, alpha_SS = (Reynolds + Maxwell)/p_tot and
    // Maxwell/Reynolds (both over rho >= rho_min), Mdot
    void analysis(int r, double rho_min, double* out) {
        set_active({r});
        const int nl = (int)act_slots.size();
        diagbuf.alloc((size_t)nl * NANA_W);
        KPS P = kp(list.p, nl);
        hipLaunchKernelGGL(k_analysis, dim3(nl), dim3(TPB), 0, stream, P, rho_min, diagbuf.p);
        SCHK(hipGetLastError());
        std::vector<double> h((size_t)nl * NANA_W);
        SCHK(hipMemcpyAsync(h.data(), diagbuf.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
        SCHK(hipStreamSynchronize(stream));
        double S[NANA_W] = {};
        for (int q = 0; q < nl; q++)
            for (int w = 0; w < NANA_W; w++) S[w] += h[(size_t)q * NANA_W + w];
        analysis_combine(S, out);
    }
};

}  // namespace sph

// ============================================================ C ABI
using namespace sph;

struct bhs_engine {
    std::unique_ptr<Engine> e;
    std::string err;
};

namespace {
thread_local std::string g_err;
template <class Fn>
int guard(bhs_engine* h, Fn&& fn) {
    if (!h) return BHS_EINVAL;
    try {
        return fn();
    } catch (const std::invalid_argument& x) {
        h->err = x.what();
        return BHS_EINVAL;
    } catch (const std::length_error& x) {
        h->err = x.what();
        return BHS_ENOMEM;
    } catch (const std::domain_error& x) {
        h->err = x.what();
        return BHS_ESTATE;
    } catch (const std::exception& x) {
        h->err = x.what();
        return BHS_EDEVICE;
    }
}
}  // namespace

extern "C" {

int bhs_create(int bz, long long max_blocks, bhs_engine** out) {
    if (!out) return BHS_EINVAL;
    *out = nullptr;
    if (bz != 1 && bz != BX) { g_err = "bz must be 1 (axisymmetric) or 16"; return BHS_EINVAL; }
    if (max_blocks <= 0) { g_err = "max_blocks must be positive"; return BHS_EINVAL; }
    try {
        auto h = std::make_unique<bhs_engine>();
        h->e = std::make_unique<Engine>(bz, max_blocks);
        *out = h.release();
        return BHS_OK;
    } catch (const std::exception& x) {
        g_err = x.what();
        return BHS_EDEVICE;
    }
}
void bhs_destroy(bhs_engine* h) { delete h; }
const char* bhs_last_error(const bhs_engine* h) { return h ? h->err.c_str() : g_err.c_str(); }

long long bhs_block_bytes(int bz) {
    const Lay L = make_lay(bz);
    return (long long)sizeof(double) * (2LL * 8 * L.ncell + 5LL * L.nint + 2LL * 3 * L.nf + 3LL * NFW * L.nff);
}

int bhs_run_add(bhs_engine* h, const bhs_physics* ph, const bhs_grid* g, int nblocks, const int* blocks, int* run) {
    return guard(h, [&] {
        if (!ph || !g || !blocks || !run) throw std::invalid_argument("null argument");
        *run = h->e->add_run(*ph, *g, nblocks, blocks);
        return BHS_OK;
    });
}
int bhs_run_remove(bhs_engine* h, int run) {
    return guard(h, [&] {
        h->e->remove_run(run);
        return BHS_OK;
    });
}
int bhs_run_nblocks(bhs_engine* h, int run, int* nb) {
    return guard(h, [&] {
        if (run < 0 || run >= (int)h->e->runs.size() || !h->e->runs[run].live) throw std::invalid_argument("no such run");
        *nb = (int)h->e->runs[run].slots.size();
        return BHS_OK;
    });
}
int bhs_run_init_torus(bhs_engine* h, int run, const bhs_torus* t, double* info) {
    return guard(h, [&] {
        if (!t) throw std::invalid_argument("null torus");
        h->e->init_torus(run, *t, info);
        return BHS_OK;
    });
}
int bhs_run_get_state(bhs_engine* h, int run, double* U, double* W, double* Bf) {
    return guard(h, [&] {
        h->e->get_state(run, U, W, Bf);
        return BHS_OK;
    });
}
int bhs_run_set_state(bhs_engine* h, int run, const double* U, const double* W, const double* Bf) {
    return guard(h, [&] {
        h->e->set_state(run, U, W, Bf);
        return BHS_OK;
    });
}
int bhs_set_precision(bhs_engine* h, int mixed) {
    return guard(h, [&] {
        if (mixed < 0 || mixed > 2) throw std::invalid_argument("precision must be 0 (FP64), 1 (mixed) or 2 (mixed, fast FP32)");
        h->e->mixed = mixed != 0;
        h->e->fastf = mixed == 2;
        return BHS_OK;
    });
}
int bhs_set_schedule(bhs_engine* h, int persistent, int wg_per_cu) {
    return guard(h, [&] {
        if (wg_per_cu < 0 || wg_per_cu > 32) throw std::invalid_argument("wg_per_cu out of range");
        if (persistent < 0 || persistent > 4) throw std::invalid_argument("schedule must be 0, 1, 2, 3 or 4");
        h->e->mode = persistent;
        h->e->wg_per_cu = wg_per_cu;
        return BHS_OK;
    });
}
int bhs_set_variants(bhs_engine* h, int flux, int update, int c2p) {
    return guard(h, [&] {
        if (flux < 0 || flux >= NVAR_FLUX) throw std::invalid_argument("flux variant out of range");
        for (int v : {update, c2p})
            if (v < 0 || v >= NVAR) throw std::invalid_argument("launch variant out of range");
        h->e->var_flux = flux;
        h->e->var_update = update;
        h->e->var_c2p = c2p;
        return BHS_OK;
    });
}
int bhs_set_face_variant(bhs_engine* h, int face) {
    return guard(h, [&] {
        if (face < 0 || face >= NVAR_FACE) throw std::invalid_argument("face variant out of range");
        h->e->var_face = face;
        return BHS_OK;
    });
}
int bhs_set_barrier_spin_limit(bhs_engine* h, unsigned int polls) {
    return guard(h, [&] {
        h->e->spin_limit = polls ? polls : SPIN_LIMIT;
        h->e->graph_release();
        return BHS_OK;
    });
}
int bhs_set_xcd_remap(bhs_engine* h, int enable) {
    return guard(h, [&] {
        h->e->xcd_remap = enable != 0;
        return BHS_OK;
    });
}
int bhs_variant_info(int v, int* tpb, int* min_waves) {
    if (v < 0 || v >= NVAR_FLUX || !tpb || !min_waves) return BHS_EINVAL;
    *tpb = VAR_TB[v % NVAR];
    *min_waves = VAR_MW[v % NVAR];
    return BHS_OK;
}
int bhs_use_graph(bhs_engine* h, int enable) {
    return guard(h, [&] {
        h->e->use_graph = enable != 0;
        h->e->graph_release();
        return BHS_OK;
    });
}
int bhs_steps(bhs_engine* h, int nruns, const int* runs, int nsteps, double* dt_out) {
    return guard(h, [&] {
        if (nruns <= 0 || !runs || nsteps < 0) throw std::invalid_argument("bad run list or step count");
        Engine& E = *h->e;
        E.set_active(std::vector<int>(runs, runs + nruns));
        std::vector<RunCtr> c;
        for (int s = 0; s < nsteps; s++) {
            E.step();
            if (dt_out) {
                // dt of each run: read after the step (it is also the value the
                // step used)
                E.read_ctr(c);
                for (int q = 0; q < nruns; q++)
                    dt_out[(size_t)s * nruns + q] = __builtin_bit_cast(double, c[runs[q]].dtbits);
            }
        }
        E.sync();
        E.check_barrier();
        return BHS_OK;
    });
}
int bhs_steps_timed(bhs_engine* h, int nruns, const int* runs, int nsteps, double* ms_per_step) {
    return guard(h, [&] {
        if (nruns <= 0 || !runs || nsteps < 0 || !ms_per_step) throw std::invalid_argument("bad arguments");
        Engine& E = *h->e;
        E.set_active(std::vector<int>(runs, runs + nruns));
        std::vector<hipEvent_t> ev(nsteps + 1);
        for (auto& e : ev) SCHK(hipEventCreate(&e));
        SCHK(hipEventRecord(ev[0], E.stream));
        for (int s = 0; s < nsteps; s++) {
            E.step();
            SCHK(hipEventRecord(ev[s + 1], E.stream));
        }
        E.sync();
        for (int s = 0; s < nsteps; s++) {
            float ms = 0.f;
            SCHK(hipEventElapsedTime(&ms, ev[s], ev[s + 1]));
            ms_per_step[s] = ms;
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
        E.check_barrier();
        return BHS_OK;
    });
}
int bhs_profile(bhs_engine* h, int enable, const char* only) {
    return guard(h, [&] {
        Engine& E = *h->e;
        E.sync();
        E.profile = enable != 0;
        E.prof_only = only ? only : "";
        E.prof_acc.clear();
        E.graph_release();
        return BHS_OK;
    });
}
int bhs_profile_report(bhs_engine* h, char* buf, int len) {
    return guard(h, [&] {
        Engine& E = *h->e;
        E.sync();
        std::string out;
        char line[256];
        for (auto& kv : E.prof_acc) {
            std::snprintf(line, sizeof line, "%s %lld %.6f\n", kv.first.c_str(), kv.second.first, kv.second.second);
            out += line;
        }
        if (buf && len > 0) {
            std::strncpy(buf, out.c_str(), (size_t)len - 1);
            buf[len - 1] = 0;
        }
        return (int)out.size();
    });
}
int bhs_c2p(bhs_engine* h, int run) {
    return guard(h, [&] {
        h->e->c2p_run(run);
        return BHS_OK;
    });
}
int bhs_run_diagnostics(bhs_engine* h, int run, double* out) {
    return guard(h, [&] {
        h->e->diagnostics(run, out);
        return BHS_OK;
    });
}
int bhs_run_analysis(bhs_engine* h, int run, double rho_min, double* out) {
    return guard(h, [&] {
        if (!out) throw std::invalid_argument("null output");
        h->e->analysis(run, rho_min, out);
        return BHS_OK;
    });
}
int bhs_run_diagnostics_static(bhs_engine* h, int run, double* out) {
    return guard(h, [&] {
        if (!out) throw std::invalid_argument("null output");
        h->e->diagnostics_static(run, out);
        return BHS_OK;
    });
}
int bhs_sync(bhs_engine* h) {
    return guard(h, [&] {
        h->e->sync();
        return BHS_OK;
    });
}

}  // extern "C"
