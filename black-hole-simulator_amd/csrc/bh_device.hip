// gfx950 kernels for one RK2 step of the GRMHD fluid solver, the device-side
// tree, and the C ABI declared in include/bhfluid.h.
//
// One launch per phase of the reference's rk2integrator::rkstg
// (src/fluid/solver/integrator.cpp:191-260), each over every block of the
// tree at once: ghost fill (in ordered waves), con2prim, PLM+HLLD fluxes,
// EMF average, face-field update, face-to-cell field average, conserved
// update; plus the CFL reduction and the floors pass.  State is SoA per
// block in HBM; all static geometry is precomputed.  dt stays on the device
// between launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "bh_common.hpp"
#include "bh_crmath.hpp"
#include "bh_physics.hpp"
#include "bh_init.hpp"
#include "bh_tree.hpp"
#include "../../include/bhfluid.h"

namespace bh {

#define HIPCHK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) throw DevError(std::string(#x) + ": " + hipGetErrorString(e_));      \
    } while (0)

struct DevError {
    std::string msg;
    explicit DevError(std::string m) : msg(std::move(m)) {}
};

constexpr int TPB = 256;
constexpr int WG_GHOST = (NGHOST + TPB - 1) / TPB;   // 16
constexpr int WG_CELL = (NCELL + TPB - 1) / TPB;     // 32
// active zones within two cells of a block face (the rest, 12^3, are deep)
constexpr int NBORDER = NINT - (BLK - 2 * GH) * (BLK - 2 * GH) * (BLK - 2 * GH);  // 2368
constexpr int NC2P = NGHOST + NBORDER;                                             // 6272
constexpr int WG_C2P = (NC2P + TPB - 1) / TPB;                                     // 25
constexpr int WG_FACE = (3 * NFC + TPB - 1) / TPB;   // 51
constexpr int WG_EDGE = (19 * 19 * 19 + TPB - 1) / TPB;  // 27
constexpr int WG_BF = (18 * 18 * 18 + TPB - 1) / TPB;    // 23
constexpr int WG_INT = NINT / TPB;                   // 16
constexpr int BRENT_WG = 2048;                       // grid of the Brent pass
// The Brent queue is split into NSUBQ sub-queues, one counter each on its
// own 128-byte line (see subq_of).  Appends to one counter serialise at its
// L2 channel (one wavefront atomic at a time), which made a single counter
// the bottleneck of the con2prim passes; NSUBQ counters spread them over
// the channels.
constexpr int NSUBQ = 256;
constexpr int QN_STRIDE = 32;                        // uints per counter line
static_assert(NSUBQ <= TPB, "k_qreset clears one counter per thread");
static_assert(BRENT_WG % NSUBQ == 0, "every sub-queue gets the same number of Brent workgroups");

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup p runs on
// XCD p % 8) and each XCD has its own L2.  Kernels whose neighbouring work
// items share inputs (shared faces) renumber their workgroups so that each
// XCD takes a contiguous range of the work: the workgroups of one block then
// run on one XCD and share its L2 instead of fetching the block's data into
// several.
constexpr int NXCD = 8;
__device__ __forceinline__ int xcd_local(int p, int G) {
    const int per = G / NXCD;
    if (p >= per * NXCD) return p;  // the last G % 8 workgroups keep their number
    return (p % NXCD) * per + p / NXCD;
}

// Kernel view of the device tree.
struct KP {
    double gamma;
    int limiter;
    double* U;           // [nb][8][8000]
    double* W;           // [nb][8][8000]
    double* Bf;          // [nb][3][8400]
    double* Us;          // [nb][5][4096] RK store of D, S, tau
    double* F;           // [nb][3][8][4352] fluxes on the computed faces
    const double* gct;   // [nb][8][8000]
    const double* gcc;   // [nb][11][8000]
    const double* gfl;   // [nb][2][8000]
    const double* gfc;   // [nb][3][8][4352]
    const double* gchr;  // [nb][40][4096]
    const double* dxyz;  // [nb][3]
    const double* org;   // [nb][3] first x, y, z cell edge of each block
    const double* gth;   // [nb][8000] polar angle of each cell centre
    const int* ghtab;    // [nb][12]
    const unsigned char* ghsame;  // [nb][6] copied ghosts share their source's metric bits
    // [nb] bit 2d+s: the two layers next to side (d, s) of the block are read
    // by a same-level copy into a side whose ghosts do not reuse the source's
    // con2prim outcome (ghsame 0); such a ghost keeps the source's
    // pre-con2prim v and B when its own con2prim fails (grid.cpp:559-566)
    const int* pmask;
    // [nb][2] the leaf whose high (0) / low (1) z-side ghosts copy this
    // leaf's low (0) / high (1) z layers and are filled by push_zghost, or -1
    const int* push_to;
    // [8000] visiting order of the ghost fill (its first 3904 entries) and
    // of the con2prim phase (its first 6272): the ghost cells in storage
    // order, then the border active zones, then the deep ones, each in
    // storage order (neighbouring lanes touch neighbouring addresses; the
    // z-side ghost rows, two cells per side and row, meet across rows)
    const short* cell_order;
    unsigned long long* dtbits;  // dt as an ordered bit pattern
    unsigned long long* faults;
    // cst[b][c] = 1 when W holds the successful con2prim of the cell's current
    // U; a repeated con2prim of the same U is then skipped (it would return
    // the same W).  Every kernel that writes U keeps it current.  Between the
    // early con2prim pass and the con2prim phase, an active zone may also hold
    // CST_FAIL (+1 when the failing call counted equation-of-state faults,
    // always 3 then): con2prim of the current U fails; W and U are untouched
    // (CST_FLFAIL, +1 likewise, where they already hold the floor state);
    // or CST_DONE: the con2prim phase's work on the zone is already done (a
    // failure applied early where no ghost fill reads the zone).  A ghost
    // cell or an active zone after the floors may hold CST_FLOOR: W is the
    // floor state of the cell (floored rho and eps, v = B = +0) and U its
    // ptoc, as the failure paths leave them; flc then gives the outcome of
    // the next con2prim of that U.
    unsigned char* cst;
    // [nb][8000] con2prim of the cell's floor state (ptoc of rho, eps
    // floored, v = B = +0, the cell's tree metric): 0 = solved, else 1 + the
    // equation-of-state faults of the failed con2prim + 8 x those of the
    // floor state's ptoc (flc_faults).  Static per cell.
    const unsigned char* flc;
    double* lg;  // [nb][2][8000] log(rho), log(eps) of the current W
    // Cells whose con2prim needs the Brent fallback are queued by the first
    // pass (Newton for every cell) and finished by k_brent over a compact
    // list, so that a few slow cells do not hold whole wavefronts.
    unsigned int* qn;     // sub-queue lengths, qn[k * QN_STRIDE]
    unsigned int qcap;    // capacity of each sub-queue
    int* qb;              // block
    int* qc;              // cell | kind << 16
    double* qx;           // Newton bracket upper end
    // con2prim input of each queued cell, [QS_WORDS][NSUBQ * qcap]: U, then
    // sqrt(gamma), gamma^13, gamma^33 of the tree metric, stored by the pusher
    // so that the Brent pass reads its inputs contiguously
    double* qs;
    unsigned long long* qstat;  // queued cells per pass (diagnostic)
};
enum { Q_C2P = 0, Q_GHOST = 1, Q_PROLONG = 2, Q_FLOOR = 3, Q_PRE = 4 };
constexpr unsigned char CST_FAIL = 2;
constexpr unsigned char CST_DONE = 4;
constexpr unsigned char CST_FLOOR = 5;
// CST_FAIL for a zone that already holds its floor state (CST_FLOOR before
// the early pass), so that the failure path's floor would change nothing
constexpr unsigned char CST_FLFAIL = 6;
// faults of a failed floor-state con2prim (flc code > 0): the con2prim alone,
// or with the floor path's ptoc that follows it
__device__ __forceinline__ int flc_c2p_faults(unsigned char code) { return (code - 1) & 7; }
__device__ __forceinline__ int flc_faults(unsigned char code) { return ((code - 1) & 7) + ((code - 1) >> 3); }
// status of a failed con2prim and the faults it counted (0 or 3: the three
// equation-of-state calls of the primitive recovery all see the same rho)
__device__ __forceinline__ unsigned char fail_status(int nf) { return CST_FAIL + (nf > 0 ? 1 : 0); }
__device__ __forceinline__ int fail_faults(unsigned char st) { return (st & 1) ? 3 : 0; }
// CST_FAIL, CST_FLFAIL, each + 1 with faults
__device__ __forceinline__ bool is_fail(unsigned char st) {
    return (st & ~1) == CST_FAIL || (st & ~1) == CST_FLFAIL;
}

constexpr int chr_slot(int m, int n, int l) {
    int lo = n < l ? n : l, hi = n < l ? l : n;
    for (int z = 0; z < NCHR; z++)
        if (kChrIdx[z][0] == m && kChrIdx[z][1] == lo && kChrIdx[z][2] == hi) return z;
    return -1;
}

// Fluxes are stored compactly on the computed faces only, [b][dir][comp][4352]
// (x-faces i in [-1,15], j,k in [0,15]; y and z likewise).  The reference
// keeps them in padded arrays whose other entries stay zero and are read as
// zero by the EMF average; fB() returns that zero.
__device__ __forceinline__ int face_lf(int dir, int i, int j, int k) {
    return dir == 0 ? ((i + 1) * 16 + j) * 16 + k : (dir == 1 ? (i * 17 + (j + 1)) * 16 + k : (i * 16 + j) * 17 + (k + 1));
}
__device__ __forceinline__ bool face_in(int dir, int i, int j, int k) {
    int lo0 = dir == 0 ? -1 : 0, lo1 = dir == 1 ? -1 : 0, lo2 = dir == 2 ? -1 : 0;
    return i >= lo0 && i < BLK && j >= lo1 && j < BLK && k >= lo2 && k < BLK;
}

__device__ __forceinline__ double& Ua(const KP& P, int b, int v, int c) { return P.U[((size_t)b * 8 + v) * NCELL + c]; }
__device__ __forceinline__ double& Wa(const KP& P, int b, int v, int c) { return P.W[((size_t)b * 8 + v) * NCELL + c]; }

__device__ __forceinline__ Cons loadU(const KP& P, int b, int c) {
    Cons q;
    q.D = Ua(P, b, UD, c);
    q.S[0] = Ua(P, b, US0, c); q.S[1] = Ua(P, b, US1, c); q.S[2] = Ua(P, b, US2, c);
    q.tau = Ua(P, b, UTAU, c);
    q.B[0] = Ua(P, b, UB0, c); q.B[1] = Ua(P, b, UB1, c); q.B[2] = Ua(P, b, UB2, c);
    return q;
}
__device__ __forceinline__ void storeU(const KP& P, int b, int c, const Cons& q) {
    Ua(P, b, UD, c) = q.D;
    Ua(P, b, US0, c) = q.S[0]; Ua(P, b, US1, c) = q.S[1]; Ua(P, b, US2, c) = q.S[2];
    Ua(P, b, UTAU, c) = q.tau;
    Ua(P, b, UB0, c) = q.B[0]; Ua(P, b, UB1, c) = q.B[1]; Ua(P, b, UB2, c) = q.B[2];
}
__device__ __forceinline__ Prim loadW(const KP& P, int b, int c) {
    Prim w;
    w.rho = Wa(P, b, WRHO, c); w.eps = Wa(P, b, WEPS, c);
    for (int d = 0; d < 3; d++) { w.v[d] = Wa(P, b, WV0 + d, c); w.B[d] = Wa(P, b, WB0 + d, c); }
    return w;
}
__device__ __forceinline__ void storeW(const KP& P, int b, int c, const Prim& w) {
    Wa(P, b, WRHO, c) = w.rho; Wa(P, b, WEPS, c) = w.eps;
    for (int d = 0; d < 3; d++) { Wa(P, b, WV0 + d, c) = w.v[d]; Wa(P, b, WB0 + d, c) = w.B[d]; }
}
// tree metric at a cell centre
__device__ __forceinline__ GPt cell_tree(const KP& P, int b, int c) {
    const double* g = P.gct + (size_t)b * NGCT * NCELL + c;
    GPt m;
    m.sqrtdetg = 0.0; m.beta0 = 0.0;
    m.sqrtdetgam = g[0]; m.alpha = g[NCELL]; m.g11 = g[2 * NCELL]; m.g13 = g[3 * NCELL];
    m.g22 = g[4 * NCELL]; m.g33 = g[5 * NCELL]; m.gi13 = g[6 * NCELL]; m.gi33 = g[7 * NCELL];
    return m;
}
__device__ __forceinline__ double cell_sqrtdetg(const KP& P, int b, int c) {
    return P.gcc[(size_t)b * NGCC * NCELL + c];
}
__device__ __forceinline__ double gcc_at(const KP& P, int b, int v, int c) {
    return P.gcc[((size_t)b * NGCC + v) * NCELL + c];
}
// Store x unless the location already holds the same bits (where the
// field does not change, e.g. B in the field-free atmosphere, the store is
// skipped and the write traffic with it).
__device__ __forceinline__ void store_changed(double* p, double old, double x) {
    if (__double_as_longlong(x) != __double_as_longlong(old)) *p = x;
}
// storeU of the ptoc of a floored state: its U.B (sqrt(gamma) B) keeps the
// bits U.B holds wherever B = 0, so U.B is stored where it differs
__device__ __forceinline__ void storeU_floor(const KP& P, int b, int c, const Cons& q) {
    double* p = P.U + (size_t)b * 8 * NCELL + c;
    const double o0 = p[UB0 * NCELL], o1 = p[UB1 * NCELL], o2 = p[UB2 * NCELL];
    p[UD * NCELL] = q.D;
    p[US0 * NCELL] = q.S[0]; p[US1 * NCELL] = q.S[1]; p[US2 * NCELL] = q.S[2];
    p[UTAU * NCELL] = q.tau;
    store_changed(p + UB0 * NCELL, o0, q.B[0]);
    store_changed(p + UB1 * NCELL, o1, q.B[1]);
    store_changed(p + UB2 * NCELL, o2, q.B[2]);
}
// storeW of a con2prim result: its B (U.B / sqrt(gamma)) mostly has the bits
// W.B already holds (always where B = 0), so B is stored where it differs
__device__ __forceinline__ void storeW_c2p(const KP& P, int b, int c, const Prim& w) {
    double* p = P.W + (size_t)b * 8 * NCELL + c;
    const double o0 = p[WB0 * NCELL], o1 = p[WB1 * NCELL], o2 = p[WB2 * NCELL];
    p[WRHO * NCELL] = w.rho; p[WEPS * NCELL] = w.eps;
    p[WV0 * NCELL] = w.v[0]; p[WV1 * NCELL] = w.v[1]; p[WV2 * NCELL] = w.v[2];
    store_changed(p + WB0 * NCELL, o0, w.B[0]);
    store_changed(p + WB1 * NCELL, o1, w.B[1]);
    store_changed(p + WB2 * NCELL, o2, w.B[2]);
}
__device__ __forceinline__ void add_faults(const KP& P, int nf) {
    if (nf) atomicAdd(P.faults, (unsigned long long)nf);
}

// con2prim of one cell with the floor fallback of the c2p phase
// (integrator.cpp:205-220): floor state keeps the cell's B, U recomputed.
__device__ __forceinline__ bool c2p_cell(const KP& P, int b, int c, int& nf) {
    GPt m = cell_tree(P, b, c);
    Prim pv;
    if (ctop(m, P.gamma, loadU(P, b, c), pv, nf)) {
        storeW(P, b, c, pv);
        return true;
    } else {
        const double* fl = P.gfl + (size_t)b * 2 * NCELL;
        Prim f = floor_prim(fl[c], fl[NCELL + c]);
        f.B[0] = Wa(P, b, WB0, c); f.B[1] = Wa(P, b, WB1, c); f.B[2] = Wa(P, b, WB2, c);
        storeW(P, b, c, f);
        storeU(P, b, c, ptoc(m, P.gamma, f, nf));
        return false;
    }
}
__device__ __forceinline__ unsigned char& cst_at(const KP& P, int b, int c) { return P.cst[(size_t)b * NCELL + c]; }

// Sub-queue of a pushing workgroup: each run of NSUBQ consecutive workgroups
// covers every sub-queue once, rotated by 37 per run, so that a sub-queue
// collects workgroups from different positions within their blocks (cells
// that need the fallback cluster by position) and at most
// ceil(grid / NSUBQ) workgroups.
__device__ __forceinline__ unsigned int subq_of(unsigned int wg) {
    return (wg % NSUBQ + (wg / NSUBQ) * 37u) % NSUBQ;
}
constexpr int QS_WORDS = 11;

// Queue slots are claimed once per wavefront: the lowest pushing lane adds
// the number of pushing lanes to its workgroup's sub-queue length and each
// lane takes the slot given by its rank among them.  A sub-queue receives
// the cells of at most ceil(grid / NSUBQ) workgroups, which its capacity
// covers.
__device__ __forceinline__ void push_brent(const KP& P, int b, int c, int kind, double xi_h, const Cons& u,
                                           const GPt& m) {
    const unsigned long long mask = __ballot(1);
    const int lane = (int)__lane_id();
    const int leader = __ffsll((unsigned long long)mask) - 1;
    const unsigned int q = subq_of(blockIdx.x);
    unsigned int base = 0;
    if (lane == leader) base = atomicAdd(P.qn + q * QN_STRIDE, (unsigned int)__popcll(mask));
    base = __shfl(base, leader);
    const size_t i = (size_t)q * P.qcap + base + (unsigned int)__popcll(mask & ((1ull << lane) - 1ull));
    P.qb[i] = b;
    P.qc[i] = c | (kind << 16);
    P.qx[i] = xi_h;
    const size_t st = (size_t)NSUBQ * P.qcap;
    double* e = P.qs + i;
    e[0] = u.D; e[st] = u.S[0]; e[2 * st] = u.S[1]; e[3 * st] = u.S[2]; e[4 * st] = u.tau;
    e[5 * st] = u.B[0]; e[6 * st] = u.B[1]; e[7 * st] = u.B[2];
    e[8 * st] = m.sqrtdetgam; e[9 * st] = m.gi13; e[10 * st] = m.gi33;
}
// con2prim set-up of queue entry i from its stored input
__device__ __forceinline__ void queued_setup(const KP& P, size_t i, C2P& s) {
    const size_t st = (size_t)NSUBQ * P.qcap;
    const double* q = P.qs + i;
    Cons u;
    u.D = q[0]; u.S[0] = q[st]; u.S[1] = q[2 * st]; u.S[2] = q[3 * st]; u.tau = q[4 * st];
    u.B[0] = q[5 * st]; u.B[1] = q[6 * st]; u.B[2] = q[7 * st];
    GPt m;
    m.sqrtdetg = 0.0; m.alpha = 0.0; m.beta0 = 0.0; m.g11 = 0.0; m.g13 = 0.0; m.g22 = 0.0; m.g33 = 0.0;
    m.sqrtdetgam = q[8 * st]; m.gi13 = q[9 * st]; m.gi33 = q[10 * st];
    ctop_setup(m, u, s);
}
// storeW of a floor state whose B was read from this W: B is left in place,
// and rho, eps and v are stored where they differ from what W holds (a zone
// floored again keeps the same words)
__device__ __forceinline__ void store_floorW(const KP& P, int b, int c, const Prim& f) {
    double* w = P.W + (size_t)b * 8 * NCELL + c;
    const double r = w[WRHO * NCELL], e = w[WEPS * NCELL];
    const double v0 = w[WV0 * NCELL], v1 = w[WV1 * NCELL], v2 = w[WV2 * NCELL];
    store_changed(w + WRHO * NCELL, r, f.rho);
    store_changed(w + WEPS * NCELL, e, f.eps);
    store_changed(w + WV0 * NCELL, v0, f.v[0]);
    store_changed(w + WV1 * NCELL, v1, f.v[1]);
    store_changed(w + WV2 * NCELL, v2, f.v[2]);
}
// c2p failure path (integrator.cpp:211-218): floor state with the cell's B
__device__ __forceinline__ void c2p_floor(const KP& P, int b, int c, const GPt& m, int& nf) {
    const double* fl = P.gfl + (size_t)b * 2 * NCELL;
    Prim f = floor_prim(fl[c], fl[NCELL + c]);
    f.B[0] = Wa(P, b, WB0, c); f.B[1] = Wa(P, b, WB1, c); f.B[2] = Wa(P, b, WB2, c);
    store_floorW(P, b, c, f);
    storeU_floor(P, b, c, ptoc(m, P.gamma, f, nf));
}
// true when v and B are all +0.0 (bit patterns zero): W with floored rho
// and eps is then exactly the cell's floor state
__device__ __forceinline__ bool vb_zero(const Prim& w) {
    unsigned long long o = 0;
#pragma unroll
    for (int d = 0; d < 3; d++)
        o |= (unsigned long long)__double_as_longlong(w.v[d]) | (unsigned long long)__double_as_longlong(w.B[d]);
    return o == 0;
}
// ghost con2prim failure path (grid.cpp:559-566): rho, eps floored, v and B
// kept.  Returns the memo state of the result (CST_FLOOR or 0).
__device__ __forceinline__ unsigned char ghost_floor(const KP& P, int b, int c, const GPt& m, int& nf) {
    const double* fl = P.gfl + (size_t)b * 2 * NCELL;
    Prim w = loadW(P, b, c);
    w.rho = fl[c];
    w.eps = fl[NCELL + c];
    Wa(P, b, WRHO, c) = w.rho;
    Wa(P, b, WEPS, c) = w.eps;
    storeU_floor(P, b, c, ptoc(m, P.gamma, w, nf));
    return vb_zero(w) ? CST_FLOOR : 0;
}
// First con2prim pass: 1 = solved (pv set), 0 = failed, -1 = queued.  A
// Brent fallback whose outcome is known without its loop (the atmosphere
// form) is finished here; the others go to the queue.
__device__ __forceinline__ int c2p_first(const KP& P, int b, int c, const GPt& m, int kind, Prim& pv, int& nf) {
    C2P s;
    const Cons u = loadU(P, b, c);
    ctop_setup(m, u, s);
    double xi, xi_h;
    C2PEnds ends;
    if (!ctop_newton(P.gamma, s, xi, xi_h, &ends)) {
        const int r = ctop_brent_short(P.gamma, s, xi_h, xi, &ends);
        if (r < 0) {
            push_brent(P, b, c, kind, xi_h, u, m);
            return -1;
        }
        if (r == 0) return 0;
    }
    return ctop_finish(P.gamma, s, xi, pv, nf) ? 1 : 0;
}

__device__ __forceinline__ int cell_of(int dim, int a, int t1, int t2) {
    return dim == 0 ? lindx(a, t1, t2) : dim == 1 ? lindx(t1, a, t2) : lindx(t1, t2, a);
}

// Active zones whose pre-con2prim W a ghost fill reads: those a
// domain-boundary side mirrors, and those a same-level copy without outcome
// reuse reads (pmask).  The early pass leaves them to the con2prim phase.
__device__ __forceinline__ bool mirror_source(const KP& P, int b, int i, int j, int k) {
    const int* gt = P.ghtab + b * 12;
    const int pm = P.pmask[b];
    const int x[3] = {i, j, k};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if ((gt[(2 * d) * 2] == GH_BNDY || (pm >> (2 * d) & 1)) && x[d] < GH) return true;
        if ((gt[(2 * d + 1) * 2] == GH_BNDY || (pm >> (2 * d + 1) & 1)) && x[d] >= BLK - GH) return true;
    }
    return false;
}

// ---------------------------------------------------------------- ghosts
// amrtree::ghosts for the leaves of one wave, grid.cpp:390-571: each thread
// fills one ghost cell from its face neighbour (copy, coarse prolongation or
// domain-boundary mirror; edge and corner ghosts are not filled, quirk Q27),
// then runs the ghost con2prim (failure keeps v and B, quirk Q29).
// Four wavefronts per SIMD (128 VGPRs) measured 3 % faster than the three
// that its register count would otherwise allow.
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_ghost(KP P, const int* __restrict__ blocks) {
    const int b = blocks[blockIdx.x / WG_GHOST];
    const int t = (blockIdx.x % WG_GHOST) * TPB + threadIdx.x;
    if (t >= NGHOST) return;
    const int c = P.cell_order[t];
    const int i = c / 400 - GH, j = (c / 20) % 20 - GH, k = c % 20 - GH;
    int nf = 0;
    const int oi = (i < 0 || i >= BLK), oj = (j < 0 || j >= BLK), ok = (k < 0 || k >= BLK);
    if (oi + oj + ok == 1) {
        const int dim = oi ? 0 : (oj ? 1 : 2);
        const int a = dim == 0 ? i : (dim == 1 ? j : k);
        const int s = a < 0 ? 0 : 1;
        const int g = s ? a - BLK : a + GH;
        const int t1 = dim == 0 ? j : i;
        const int t2 = dim == 2 ? j : k;
        const int kind = P.ghtab[b * 12 + (2 * dim + s) * 2];
        const int nb = P.ghtab[b * 12 + (2 * dim + s) * 2 + 1];
        if (kind == GH_COPY || kind == GH_BNDY) {
            int src_b, src_c;
            if (kind == GH_COPY) {
                src_b = nb;
                src_c = cell_of(dim, (s ? 0 : BLK - GH) + g, t1, t2);
                // already written by the source's early con2prim (push_zghost)
                if (dim == 2 && P.push_to[nb * 2 + (s ? 0 : 1)] == b && P.ghsame[b * 6 + 4 + s] &&
                    !mirror_source(P, nb, t1, t2, (s ? 0 : BLK - GH) + g))
                    return;
            }
            else { src_b = b; src_c = cell_of(dim, s ? BLK - 1 - g : g, t1, t2); }
            // all sixteen loads (and the source's con2prim status) ahead of
            // the stores into the same arrays
            double u[8], w[8];
            for (int v = 0; v < 8; v++) {
                u[v] = Ua(P, src_b, v, src_c);
                w[v] = Wa(P, src_b, v, src_c);
            }
            const unsigned char st = cst_at(P, src_b, src_c);
            const bool reuse = kind == GH_COPY && P.ghsame[b * 6 + 2 * dim + s];
            // A source with a known failure and v = B = +0, copied into a
            // ghost that holds its own floor state: the copy and the ghost
            // floor path below would leave the ghost as it is.
            if (reuse && is_fail(st) && cst_at(P, b, c) == CST_FLOOR &&
                ((st & ~1) == CST_FLFAIL || vb_zero(Prim{w[0], w[1], {w[2], w[3], w[4]}, {w[5], w[6], w[7]}}))) {
                nf = fail_faults(st) + (P.gfl[(size_t)b * 2 * NCELL + c] <= 0.0 ? 2 : 0);
                add_faults(P, nf);
                return;
            }
            {
                // U.B and W.B of the ghost keep their words where the bits agree
                double* up = P.U + (size_t)b * 8 * NCELL + c;
                double* wp = P.W + (size_t)b * 8 * NCELL + c;
                double ob[6];
                for (int v = 0; v < 3; v++) { ob[v] = up[(UB0 + v) * NCELL]; ob[3 + v] = wp[(WB0 + v) * NCELL]; }
                for (int v = 0; v < 5; v++) { up[v * NCELL] = u[v]; wp[v * NCELL] = w[v]; }
                for (int v = 0; v < 3; v++) {
                    store_changed(up + (UB0 + v) * NCELL, ob[v], u[UB0 + v]);
                    store_changed(wp + (WB0 + v) * NCELL, ob[3 + v], w[WB0 + v]);
                }
            }
            // The source holds this same U at the same metric bits, so its
            // known con2prim outcome is the ghost's: success returns the
            // source's W (copied above), failure takes the ghost floor path.
            if (reuse) {
                if (st == 1) {
                    cst_at(P, b, c) = 1;
                    return;
                }
                if (is_fail(st)) {
                    nf = fail_faults(st);
                    cst_at(P, b, c) = ghost_floor(P, b, c, cell_tree(P, b, c), nf);
                    add_faults(P, nf);
                    return;
                }
            }
        } else {
            // gh_prolong, grid.cpp:427-492 (quirk Q25: coarse index ignores octant and side)
            const int ci = dim == 0 ? g / 2 : t1 / 2;
            const int cj = dim == 1 ? g / 2 : (dim == 0 ? t1 / 2 : t2 / 2);
            const int ck = dim == 2 ? g / 2 : t2 / 2;
            const double offg = (g % 2 == 0) ? -0.25 : 0.25;
            const double offt1 = (t1 % 2 == 0) ? -0.25 : 0.25;
            const double offt2 = (t2 % 2 == 0) ? -0.25 : 0.25;
            double on, ot1, ot2;
            if (dim == 0) { on = offg; ot1 = offt1; ot2 = offt2; }
            else if (dim == 1) { on = offt1; ot1 = offg; ot2 = offt2; }
            else { on = offt1; ot1 = offt2; ot2 = offg; }
            const int cc = lindx(ci, cj, ck);
            double nu[8];
            for (int v = 0; v < 8; v++) {
                const double vc = Ua(P, nb, v, cc);
                double sn = minmod(vc - Ua(P, nb, v, lindx(ci - 1, cj, ck)), Ua(P, nb, v, lindx(ci + 1, cj, ck)) - vc);
                double st1 = minmod(vc - Ua(P, nb, v, lindx(ci, cj - 1, ck)), Ua(P, nb, v, lindx(ci, cj + 1, ck)) - vc);
                double st2 = minmod(vc - Ua(P, nb, v, lindx(ci, cj, ck - 1)), Ua(P, nb, v, lindx(ci, cj, ck + 1)) - vc);
                nu[v] = vc + sn * on + st1 * ot1 + st2 * ot2;
            }
            for (int v = 0; v < 8; v++) Ua(P, b, v, c) = nu[v];
            GPt mp = cell_tree(P, b, c);
            Prim pv;
            int r = c2p_first(P, b, c, mp, Q_PROLONG, pv, nf);
            if (r < 0) { add_faults(P, nf); return; }
            if (r > 0) {
                // the ghost con2prim below would repeat this one on the same U
                storeW(P, b, c, pv);
                cst_at(P, b, c) = 1;
                add_faults(P, nf);
                return;
            }
            c2p_floor(P, b, c, mp, nf);
        }
    } else if (cst_at(P, b, c) == 1) {
        // an edge or corner ghost (not filled) whose W is already the
        // con2prim of its U
        return;
    }
    // ghost con2prim, grid.cpp:538-570
    GPt m = cell_tree(P, b, c);
    Prim pv;
    int r = c2p_first(P, b, c, m, Q_GHOST, pv, nf);
    if (r > 0) {
        storeW_c2p(P, b, c, pv);
        cst_at(P, b, c) = 1;
    } else if (r == 0) {
        cst_at(P, b, c) = ghost_floor(P, b, c, m, nf);
    }
    add_faults(P, nf);
}

// con2prim on every stored cell of every leaf, integrator.cpp:201-222.
// Ghost cells first, in storage order, then the border active zones (mostly
// settled by the early pass).  The deep zones are not visited: the early
// pass leaves each of them solved (memo 1) or floored (CST_DONE).
__global__ __launch_bounds__(TPB) void k_c2p(KP P, const int* __restrict__ leaves) {
    const int b = leaves[blockIdx.x / WG_C2P];
    const int t = (blockIdx.x % WG_C2P) * TPB + threadIdx.x;
    if (t >= NC2P) return;
    const int c = P.cell_order[t];
    const unsigned char st = cst_at(P, b, c);
    if (st == 1 || st == CST_DONE) return;
    if (st == CST_FLOOR) {
        // con2prim of the floor state: where it fails, the failure path
        // (c2p_floor) would store this same W and U again
        const unsigned char code = P.flc[(size_t)b * NCELL + c];
        if (code) {
            add_faults(P, flc_faults(code));
            return;
        }
    }
    if ((st & ~1) == CST_FLFAIL) {
        // known failure of a zone already in its floor state: the floor
        // path would store the same W and U again
        add_faults(P, flc_faults(P.flc[(size_t)b * NCELL + c]));
        cst_at(P, b, c) = 0;
        return;
    }
    int nf = 0;
    GPt m = cell_tree(P, b, c);
    if (st == CST_FAIL || st == CST_FAIL + 1) {
        // failure found by the early pass on this same U
        nf = fail_faults(st);
        c2p_floor(P, b, c, m, nf);
        cst_at(P, b, c) = 0;
        add_faults(P, nf);
        return;
    }
    Prim pv;
    int r = c2p_first(P, b, c, m, Q_C2P, pv, nf);
    if (r > 0) {
        storeW_c2p(P, b, c, pv);
        cst_at(P, b, c) = 1;
    } else if (r == 0) {
        c2p_floor(P, b, c, m, nf);
    }
    add_faults(P, nf);
}

// Same-level z-side copies done by the source's early con2prim.  A z-layer
// zone (k < 2 or k >= 14) of leaf b whose neighbour across that side copies
// it into a ghost row (push_to: set where that neighbour is a leaf that no
// prolongation reads; used where the metric bits agree) writes the copy and
// the ghost's con2prim outcome there directly, the values the fill and the
// ghost con2prim would write: U and W copied; on success the memo set; on
// failure rho, eps floored and U recomputed with the shared metric.  This
// replaces the fill's reads of two-cell rows, which waste most of each line.
__device__ __forceinline__ void push_zghost(const KP& P, int b, int i, int j, int k, const GPt& m, const Cons& u,
                                            const Prim& w, unsigned char st) {
    const int lo = k < GH;
    const int B = P.push_to[b * 2 + (lo ? 0 : 1)];
    if (B < 0 || !P.ghsame[B * 6 + (lo ? 5 : 4)]) return;
    const int g = lindx(i, j, lo ? k + BLK : k - BLK);
    if (st == 1) {
        double* up = P.U + (size_t)B * 8 * NCELL + g;
        double* wp = P.W + (size_t)B * 8 * NCELL + g;
        const double ub0 = up[UB0 * NCELL], ub1 = up[UB1 * NCELL], ub2 = up[UB2 * NCELL];
        const double wb0 = wp[WB0 * NCELL], wb1 = wp[WB1 * NCELL], wb2 = wp[WB2 * NCELL];
        up[UD * NCELL] = u.D; up[US0 * NCELL] = u.S[0]; up[US1 * NCELL] = u.S[1]; up[US2 * NCELL] = u.S[2];
        up[UTAU * NCELL] = u.tau;
        wp[WRHO * NCELL] = w.rho; wp[WEPS * NCELL] = w.eps;
        wp[WV0 * NCELL] = w.v[0]; wp[WV1 * NCELL] = w.v[1]; wp[WV2 * NCELL] = w.v[2];
        store_changed(up + UB0 * NCELL, ub0, u.B[0]); store_changed(up + UB1 * NCELL, ub1, u.B[1]);
        store_changed(up + UB2 * NCELL, ub2, u.B[2]);
        store_changed(wp + WB0 * NCELL, wb0, w.B[0]); store_changed(wp + WB1 * NCELL, wb1, w.B[1]);
        store_changed(wp + WB2 * NCELL, wb2, w.B[2]);
        cst_at(P, B, g) = 1;
        return;
    }
    int nf = fail_faults(st);
    const double* fl = P.gfl + (size_t)B * 2 * NCELL;
    if (vb_zero(w) && cst_at(P, B, g) == CST_FLOOR) {
        // the ghost already holds the floor state this path would store
        add_faults(P, nf + (fl[g] <= 0.0 ? 2 : 0));
        return;
    }
    Prim wg = w;
    wg.rho = fl[g];
    wg.eps = fl[NCELL + g];
    storeW(P, B, g, wg);
    storeU_floor(P, B, g, ptoc(m, P.gamma, wg, nf));
    cst_at(P, B, g) = vb_zero(wg) ? CST_FLOOR : 0;
    add_faults(P, nf);
}
__device__ __forceinline__ bool z_layer(int k) { return k < GH || k >= BLK - GH; }
// push_zghost of a zone that keeps its floor state with a known failure
// (CST_FLFAIL): a destination ghost that holds its own floor state would be
// left as it is, so only the faults of the ghost floor path are counted.
__device__ __forceinline__ void push_zfloor(const KP& P, int b, int i, int j, int k, const GPt& m, unsigned char st) {
    const int lo = k < GH;
    const int B = P.push_to[b * 2 + (lo ? 0 : 1)];
    if (B < 0 || !P.ghsame[B * 6 + (lo ? 5 : 4)]) return;
    const int g = lindx(i, j, lo ? k + BLK : k - BLK);
    if (cst_at(P, B, g) == CST_FLOOR) {
        add_faults(P, fail_faults(st) + (P.gfl[(size_t)B * 2 * NCELL + g] <= 0.0 ? 2 : 0));
        return;
    }
    const int c = lindx(i, j, k);
    push_zghost(P, b, i, j, k, m, loadU(P, b, c), loadW(P, b, c), st);
}
// Active zones no ghost fill reads: copies and mirrors read the two layers
// next to a face, prolongation the coarse layers -1, 0, 1 (quirk Q25).
__host__ __device__ __forceinline__ bool deep_zone(int i, int j, int k) {
    return i >= GH && i < BLK - GH && j >= GH && j < BLK - GH && k >= GH && k < BLK - GH;
}

// The con2prim of the active zones of every leaf, run before the ghost fill.
// The fill reads active zones' U (copies, prolongation) and W (copies, whose
// ghost con2prim then replaces it or keeps its v and B on failure), never a
// result of the con2prim phase.  So a success may be stored early (W <-
// ctop(U), U unchanged, flagged current; the con2prim phase skips it), and a
// failure is only recorded (its floor state changes U, so the con2prim phase
// applies it); copied ghosts reuse either outcome.  Zones a boundary mirror
// reads keep their W and are left to the con2prim phase.
__global__ __launch_bounds__(TPB) void k_c2p_pre(KP P, const int* __restrict__ leaves) {
    const int b = leaves[blockIdx.x / WG_INT];
    const int a = (blockIdx.x % WG_INT) * TPB + threadIdx.x;
    const int i = a / 256, j = (a / 16) % 16, k = a % 16;
    const int c = lindx(i, j, k);
    if (mirror_source(P, b, i, j, k)) return;
    const unsigned char st0 = cst_at(P, b, c);
    const GPt m = cell_tree(P, b, c);
    if (st0 == CST_FLOOR) {
        // floor state left by the floors: its con2prim outcome is tabulated
        const unsigned char code = P.flc[(size_t)b * NCELL + c];
        if (code) {
            if (deep_zone(i, j, k)) {
                // the failure path would store this same W and U again
                add_faults(P, flc_faults(code));
                cst_at(P, b, c) = CST_DONE;
                return;
            }
            const unsigned char st = fail_status(flc_c2p_faults(code)) + (CST_FLFAIL - CST_FAIL);
            cst_at(P, b, c) = st;
            if (z_layer(k)) push_zfloor(P, b, i, j, k, m, st);
            return;
        }
    } else if (st0) {
        // outcome already known for this U
        if (z_layer(k)) push_zghost(P, b, i, j, k, m, loadU(P, b, c), loadW(P, b, c), st0);
        return;
    }
    int nf = 0;
    Prim pv;
    const int r = c2p_first(P, b, c, m, Q_PRE, pv, nf);
    if (r > 0) {
        storeW_c2p(P, b, c, pv);
        cst_at(P, b, c) = 1;
        if (z_layer(k)) push_zghost(P, b, i, j, k, m, loadU(P, b, c), pv, 1);
    } else if (r == 0) {
        if (deep_zone(i, j, k)) {
            // the con2prim phase's failure path (floor state with the zone's
            // B), applied now: nothing reads the zone before that phase
            c2p_floor(P, b, c, m, nf);
            cst_at(P, b, c) = CST_DONE;
            add_faults(P, nf);
            return;
        }
        const unsigned char st = fail_status(nf);
        cst_at(P, b, c) = st;
        if (z_layer(k)) push_zghost(P, b, i, j, k, m, loadU(P, b, c), loadW(P, b, c), st);
    }
    // a queued zone is pushed by k_brent
}

// floors semantics after con2prim of an active zone (grid.cpp:342-354)
__device__ __forceinline__ void floors_finish(const KP& P, int b, int c, const GPt& m, bool bhl, const Prim& pv,
                                              bool have_pv, int& nf) {
    unsigned char st = 0;
    if (!bhl && have_pv) {
        bool hl = !(pv.rho <= 0.0 || pv.eps <= 0.0);
        if (hl) { storeW_c2p(P, b, c, pv); st = 1; }
    }
    const double* fl = P.gfl + (size_t)b * 2 * NCELL;
    if (bhl || Wa(P, b, WRHO, c) <= fl[c]) {
        Prim f = floor_prim(fl[c], fl[NCELL + c]);
        f.B[0] = Wa(P, b, WB0, c); f.B[1] = Wa(P, b, WB1, c); f.B[2] = Wa(P, b, WB2, c);
        store_floorW(P, b, c, f);
        storeU_floor(P, b, c, ptoc(m, P.gamma, f, nf));
        st = vb_zero(f) ? CST_FLOOR : 0;
    }
    cst_at(P, b, c) = st;
}

// Brent fallback for the queued cells, then the rest of each caller's
// con2prim sequence.
// Workgroup w serves sub-queue w % NSUBQ together with the other
// BRENT_WG / NSUBQ workgroups of that residue.
#define FOR_QUEUED(i)                                                                 \
    const unsigned int q_ = blockIdx.x % NSUBQ;                                       \
    const unsigned int n_ = P.qn[q_ * QN_STRIDE];                                     \
    const unsigned int part_ = blockIdx.x / NSUBQ, nparts_ = gridDim.x / NSUBQ;       \
    for (unsigned int j_ = part_ * TPB + threadIdx.x; j_ < n_; j_ += nparts_ * TPB) { \
        const size_t i = (size_t)q_ * P.qcap + j_;

// Brent fallback (conserved.cpp:232-285) for the queued cells, from the
// con2prim input their pushers stored, then the rest of each caller's
// con2prim sequence: primitive recovery from the root and what the caller
// does on success or failure.
__global__ __launch_bounds__(TPB) void k_brent(KP P, int pass) {
    if (blockIdx.x < NSUBQ && threadIdx.x == 0) {
        const unsigned int n = P.qn[blockIdx.x * QN_STRIDE];
        if (n) atomicAdd(P.qstat + pass, (unsigned long long)n);
    }
    FOR_QUEUED(i)
        const int b = P.qb[i];
        const int qc = P.qc[i];
        const int c = qc & 0xffff;
        const int kind = qc >> 16;
        int nf = 0;
        Prim pv;
        bool ok = false;
        {
            C2P s;
            queued_setup(P, i, s);
            double xi;
            ok = ctop_brent(P.gamma, s, P.qx[i], xi) && ctop_finish(P.gamma, s, xi, pv, nf);
        }
        if (kind == Q_PRE) {
            // early con2prim pass: store a success, record a failure
            unsigned char st = 1;
            const int ci = c / 400 - GH, cj = (c / 20) % 20 - GH, ck = c % 20 - GH;
            if (ok) {
                storeW_c2p(P, b, c, pv);
            } else if (deep_zone(ci, cj, ck)) {
                c2p_floor(P, b, c, cell_tree(P, b, c), nf);
                cst_at(P, b, c) = CST_DONE;
                add_faults(P, nf);
                continue;
            } else {
                st = fail_status(nf);
            }
            cst_at(P, b, c) = st;
            if (z_layer(ck))
                push_zghost(P, b, ci, cj, ck, cell_tree(P, b, c), loadU(P, b, c),
                            ok ? pv : loadW(P, b, c), st);
            continue;
        }
        const GPt m = cell_tree(P, b, c);
        if (kind == Q_FLOOR) {
            floors_finish(P, b, c, m, !ok, pv, ok, nf);
        } else if (ok) {
            storeW_c2p(P, b, c, pv);
            cst_at(P, b, c) = 1;
        } else if (kind == Q_C2P) {
            c2p_floor(P, b, c, m, nf);
            cst_at(P, b, c) = 0;
        } else if (kind == Q_GHOST) {
            cst_at(P, b, c) = ghost_floor(P, b, c, m, nf);
        } else {
            // prolonged ghost: c2p floor, then the ghost con2prim of the new U
            c2p_floor(P, b, c, m, nf);
            if (ctop(m, P.gamma, loadU(P, b, c), pv, nf)) {
                storeW(P, b, c, pv);
                cst_at(P, b, c) = 1;
            } else {
                cst_at(P, b, c) = ghost_floor(P, b, c, m, nf);
            }
        }
        add_faults(P, nf);
    }
}
#undef FOR_QUEUED
__global__ void k_qreset(KP P) {
    if (threadIdx.x < NSUBQ) P.qn[threadIdx.x * QN_STRIDE] = 0u;
}

// Outcome of the con2prim of every cell's floor state (the full con2prim,
// conserved.cpp:125-229, of ptoc of the floor primitives), for CST_FLOOR.
__global__ __launch_bounds__(TPB) void k_floor_c2p(KP P, unsigned char* __restrict__ out) {
    const int b = blockIdx.x / WG_CELL;
    const int c = (blockIdx.x % WG_CELL) * TPB + threadIdx.x;
    if (c >= NCELL) return;
    const double* fl = P.gfl + (size_t)b * 2 * NCELL;
    const GPt m = cell_tree(P, b, c);
    int nfp = 0, nf = 0;
    const Cons u = ptoc(m, P.gamma, floor_prim(fl[c], fl[NCELL + c]), nfp);
    Prim pv;
    const bool ok = ctop(m, P.gamma, u, pv, nf);
    out[(size_t)b * NCELL + c] = ok ? 0 : (unsigned char)(1 + nf + 8 * nfp);
}

// ln(rho) and ln(eps) of every stored cell, for the log-space reconstruction
// (reconstruction.cpp:49-61 takes the logs of the four stencil cells of
// each face; each cell's value is the same for every face that uses it)
__global__ __launch_bounds__(TPB) void k_logs(KP P) {
    const int b = blockIdx.x / WG_CELL;
    const int c = (blockIdx.x % WG_CELL) * TPB + threadIdx.x;
    if (c >= NCELL) return;
    double* lr = P.lg + (size_t)b * 2 * NCELL;
    const double rho = Wa(P, b, WRHO, c), eps = Wa(P, b, WEPS, c);
    const double l0 = cr_log(rho), l1 = cr_log(eps);
    lr[c] = l0;
    lr[NCELL + c] = l1;
}

// ---------------------------------------------------------------- fluxes
// patch::fluxcomp, grid.cpp:255-329: PLM reconstruction (log rho, log eps,
// linear v and transverse B; reconstruction.cpp:40-97), the normal field
// taken from the face, ptoc of both states at the face and the HLLD flux.
// quirk Q1: the Riemann solver is called with dim 0 for every direction.
// Three wavefronts per SIMD (158 VGPRs, nothing spilled: at four the
// spilled registers went through the memory hierarchy with every face).
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_flux(KP P) {
    const int b = blockIdx.x / WG_FACE;
    const int f = (blockIdx.x % WG_FACE) * TPB + threadIdx.x;
    if (f >= 3 * NFC) return;
    const int dir = f / NFC;
    const int lf = f % NFC;
    int i, j, k;
    if (dir == 0) { i = lf / 256 - 1; j = (lf / 16) % 16; k = lf % 16; }
    else if (dir == 1) { i = lf / 272; j = (lf / 16) % 17 - 1; k = lf % 16; }
    else { i = lf / 272; j = (lf / 17) % 16; k = lf % 17 - 1; }
    const int stride = dir == 0 ? CPN * CPN : (dir == 1 ? CPN : 1);
    const int c0 = lindx(i, j, k);
    const int cm = c0 - stride, cp = c0 + stride, cpp = c0 + 2 * stride;
    int nf = 0;
    Prim L, R;
    {
        double fl, fr;
        // ln(rho), ln(eps) of the four stencil cells, computed once per cell by k_logs
        const double* lr = P.lg + (size_t)b * 2 * NCELL;
        const double* le = lr + NCELL;
        recon_scal(lr[cm], lr[c0], lr[cp], lr[cpp], P.limiter, fl, fr);
        L.rho = cr_exp(fl); R.rho = cr_exp(fr);
        recon_scal(le[cm], le[c0], le[cp], le[cpp], P.limiter, fl, fr);
        L.eps = cr_exp(fl); R.eps = cr_exp(fr);
        for (int d = 0; d < 3; d++) {
            recon_scal(Wa(P, b, WV0 + d, cm), Wa(P, b, WV0 + d, c0), Wa(P, b, WV0 + d, cp), Wa(P, b, WV0 + d, cpp),
                       P.limiter, fl, fr);
            L.v[d] = fl; R.v[d] = fr;
        }
        causal_v(L.v, gcc_at(P, b, 3, c0), gcc_at(P, b, 4, c0), gcc_at(P, b, 5, c0));
        causal_v(R.v, gcc_at(P, b, 3, cp), gcc_at(P, b, 4, cp), gcc_at(P, b, 5, cp));
        for (int d = 0; d < 3; d++) {
            if (d == dir) continue;
            recon_scal(Wa(P, b, WB0 + d, cm), Wa(P, b, WB0 + d, c0), Wa(P, b, WB0 + d, cp), Wa(P, b, WB0 + d, cpp),
                       P.limiter, fl, fr);
            L.B[d] = fl; R.B[d] = fr;
        }
    }
    // normal field from the face (grid.cpp:263-264, 287-288, 311-312); the
    // unrolled select keeps L.B and R.B in registers
    const double* Bf = P.Bf + ((size_t)b * 3 + dir) * NFACE;
    const int bidx = dir == 0 ? ixf(i + 1, j, k) : (dir == 1 ? iyf(i, j + 1, k) : izf(i, j, k + 1));
    const double bn = Bf[bidx];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (d == dir) { L.B[d] = bn; R.B[d] = bn; }
    }
    // tree metric at the face
    const double* gf = P.gfc + ((size_t)b * 3 + dir) * NGFC * NFC + lf;
    GPt m;
    m.sqrtdetg = gf[0]; m.sqrtdetgam = gf[NFC]; m.alpha = gf[2 * NFC]; m.beta0 = gf[3 * NFC];
    m.g11 = gf[4 * NFC]; m.g13 = gf[5 * NFC]; m.g22 = gf[6 * NFC]; m.g33 = gf[7 * NFC];
    m.gi13 = 0.0; m.gi33 = 0.0;
    Cons UL = ptoc(m, P.gamma, L, nf);
    Cons UR = ptoc(m, P.gamma, R, nf);
    double ct;
    Cons Fl = hlld(L, R, UL, UR, m, P.gamma, 0, ct, nf);
    double* Fo = P.F + ((size_t)b * 3 + dir) * 8 * NFC + lf;
    Fo[0] = Fl.D; Fo[NFC] = Fl.S[0]; Fo[2 * NFC] = Fl.S[1]; Fo[3 * NFC] = Fl.S[2];
    Fo[4 * NFC] = Fl.tau;
    Fo[5 * NFC] = Fl.B[0];
    Fo[6 * NFC] = Fl.B[1];
    Fo[7 * NFC] = Fl.B[2];
    add_faults(P, nf);
}

// ---------------------------------------------------- constrained transport
// constrans::emfcomp, transport.cpp:11-36: the edge EMFs as 4-point
// averages of the transverse face fluxes of B (zero outside computed faces,
// quirk Q19).  Ex sits at iex(i, j, k), Ey at izf(i, j, k), Ez at iyf(i, j, k)
// of the reference's EMF arrays.
struct EmfView {
    const double* F;
    __device__ double fB(int dir, int comp, int ii, int jj, int kk) const {
        return face_in(dir, ii, jj, kk) ? F[(dir * 8 + 5 + comp) * NFC + face_lf(dir, ii, jj, kk)] : 0.0;
    }
    __device__ double ex(int i, int j, int k) const {
        return (fB(1, 2, i, j, k) + fB(1, 2, i, j, k + 1) - fB(2, 1, i, j, k) - fB(2, 1, i, j + 1, k)) / 4;
    }
    __device__ double ey(int i, int j, int k) const {
        return (fB(2, 0, i, j, k) + fB(2, 0, i + 1, j, k) - fB(0, 2, i, j, k) - fB(0, 2, i, j, k + 1)) / 4;
    }
    __device__ double ez(int i, int j, int k) const {
        return (fB(0, 1, i, j, k) + fB(0, 1, i, j + 1, k) - fB(1, 0, i, j, k) - fB(1, 0, i + 1, j, k)) / 4;
    }
};

// constrans::Bfupdate, transport.cpp:38-58 (not RK-averaged, quirk Q21), with
// the nine EMFs each face triple needs formed in place from the fluxes
// (emfcomp's EMF arrays are not stored: every EMF read here lies in the
// range emfcomp computes, [-2, 16]^3)
__global__ __launch_bounds__(TPB) void k_bfupd(KP P) {
    const int b = blockIdx.x / WG_BF;
    const int t = (blockIdx.x % WG_BF) * TPB + threadIdx.x;
    if (t >= 18 * 18 * 18) return;
    const int i = t / 324 - 1, j = (t / 18) % 18 - 1, k = t % 18 - 1;
    const double dt = __longlong_as_double((long long)*P.dtbits);
    const double dx = P.dxyz[3 * b], dy = P.dxyz[3 * b + 1], dz = P.dxyz[3 * b + 2];
    const EmfView E{P.F + (size_t)b * 3 * 8 * NFC};
    double* Bx = P.Bf + (size_t)b * 3 * NFACE;
    double* By = Bx + NFACE;
    double* Bz = Bx + 2 * NFACE;
    // all loads ahead of the stores
    const double ez = E.ez(i, j, k), ez_jm = E.ez(i, j - 1, k), ez_im = E.ez(i - 1, j, k);
    const double ey = E.ey(i, j, k), ey_km = E.ey(i, j, k - 1), ey_im = E.ey(i - 1, j, k);
    const double ex = E.ex(i, j, k), ex_km = E.ex(i, j, k - 1), ex_jm = E.ex(i, j - 1, k);
    const double bx = Bx[ixf(i, j, k)], by = By[iyf(i, j, k)], bz = Bz[izf(i, j, k)];
    double dEz_dy = (ez - ez_jm) / dy;
    double dEy_dz = (ey - ey_km) / dz;
    double dEz_dx = (ez - ez_im) / dx;
    double dEx_dz = (ex - ex_km) / dz;
    double dEy_dx = (ey - ey_im) / dx;
    double dEx_dy = (ex - ex_jm) / dy;
    store_changed(Bx + ixf(i, j, k), bx, bx + dt * (dEz_dy - dEy_dz));
    store_changed(By + iyf(i, j, k), by, by + dt * (dEx_dz - dEz_dx));
    store_changed(Bz + izf(i, j, k), bz, bz + dt * (dEy_dx - dEx_dy));
}

// constrans::f2cB, transport.cpp:60-75: cell B = face average; U.B uses the
// cell metric's sqrt(-g) (quirk Q4)
__global__ __launch_bounds__(TPB) void k_f2cb(KP P) {
    const int wg = xcd_local(blockIdx.x, gridDim.x);
    const int b = wg / WG_EDGE;
    const int t = (wg % WG_EDGE) * TPB + threadIdx.x;
    if (t >= 19 * 19 * 19) return;
    const int i = t / 361 - 1, j = (t / 19) % 19 - 1, k = t % 19 - 1;
    const double* Bx = P.Bf + (size_t)b * 3 * NFACE;
    const double* By = Bx + NFACE;
    const double* Bz = Bx + 2 * NFACE;
    const int c = lindx(i, j, k);
    double b0 = (Bx[ixf(i, j, k)] + Bx[ixf(i - 1, j, k)]) / 2;
    double b1 = (By[iyf(i, j, k)] + By[iyf(i, j - 1, k)]) / 2;
    double b2 = (Bz[izf(i, j, k)] + Bz[izf(i, j, k - 1)]) / 2;
    const double sg = cell_sqrtdetg(P, b, c);  // loaded before the stores
    const double w0 = Wa(P, b, WB0, c), w1 = Wa(P, b, WB1, c), w2 = Wa(P, b, WB2, c);
    const double u0 = Ua(P, b, UB0, c), u1 = Ua(P, b, UB1, c), u2 = Ua(P, b, UB2, c);
    const double n0 = sg * b0, n1 = sg * b1, n2 = sg * b2;
    // unchanged B (bit for bit) leaves U, W and the con2prim memo as they are
    const bool same = __double_as_longlong(w0) == __double_as_longlong(b0) &&
                      __double_as_longlong(w1) == __double_as_longlong(b1) &&
                      __double_as_longlong(w2) == __double_as_longlong(b2) &&
                      __double_as_longlong(u0) == __double_as_longlong(n0) &&
                      __double_as_longlong(u1) == __double_as_longlong(n1) &&
                      __double_as_longlong(u2) == __double_as_longlong(n2);
    if (same) return;
    Wa(P, b, WB0, c) = b0; Wa(P, b, WB1, c) = b1; Wa(P, b, WB2, c) = b2;
    Ua(P, b, UB0, c) = n0; Ua(P, b, UB1, c) = n1; Ua(P, b, UB2, c) = n2;
    cst_at(P, b, c) = 0;
}

// ---------------------------------------------------------------- update
// Geometric source term, rk2integrator::source, integrator.cpp:41-153, with
// the cell metric (a=0); spherical components land in Cartesian slots.
__device__ __forceinline__ void source_terms(const KP& P, int b, int c, int a, const Prim& W, double src[4],
                                             int& nf) {
    const double sgc = gcc_at(P, b, 0, c), alpha = gcc_at(P, b, 1, c), beta0 = gcc_at(P, b, 2, c);
    const double g11 = gcc_at(P, b, 3, c), g22 = gcc_at(P, b, 4, c), g33 = gcc_at(P, b, 5, c);
    const double gam[3][3] = {{g11, 0.0, CELL_G13}, {0.0, g22, 0.0}, {CELL_G13, 0.0, g33}};
    const double gi00 = gcc_at(P, b, 6, c), gi01 = gcc_at(P, b, 7, c), gi11 = gcc_at(P, b, 8, c);
    const double gi22 = gcc_at(P, b, 9, c), gi33 = gcc_at(P, b, 10, c);
    const double gi[4][4] = {{gi00, gi01, 0.0, 0.0}, {gi01, gi11, 0.0, 0.0}, {0.0, 0.0, gi22, 0.0}, {0.0, 0.0, 0.0, gi33}};
    const double beta[3] = {beta0, 0.0, 0.0};
    double v2 = 0.0;
    double vlow[3] = {0.0, 0.0, 0.0}, Blow[3] = {0.0, 0.0, 0.0};
    for (int m = 0; m < 3; m++) {
        for (int n = 0; n < 3; n++) {
            vlow[m] += gam[m][n] * W.v[n];
            Blow[m] += gam[m][n] * W.B[n];
        }
        v2 += vlow[m] * W.v[m];
    }
    v2 = smin(v2, 1 - 1e-10);
    double ltz2 = 1.0 / (1.0 - v2);
    double ltz = sqrt(ltz2);
    double Bv = 0.0, Bsq = 0.0;
    for (int m = 0; m < 3; m++) { Bv += Blow[m] * W.v[m]; Bsq += Blow[m] * W.B[m]; }
    double b2 = Bv * Bv + Bsq / ltz2;
    double px = eos_press(P.gamma, W.rho, W.eps, nf);
    double h = eos_enth(P.gamma, W.rho, W.eps, nf);
    double u4[4], b4[4];
    u4[0] = ltz / alpha;
    for (int m = 0; m < 3; m++) u4[m + 1] = ltz * (W.v[m] - beta[m] / alpha);
    b4[0] = ltz * Bv / alpha;
    for (int m = 0; m < 3; m++) b4[m + 1] = (W.B[m] + alpha * b4[0] * u4[m + 1]) / ltz;
    double rhohb2 = W.rho * h + b2;
    double T[4][4];
    for (int m = 0; m < 4; m++)
        for (int n = 0; n < 4; n++) T[m][n] = rhohb2 * u4[m] * u4[n] + px * gi[m][n] - b4[m] * b4[n];
    double G[NCHR];
    const double* gp = P.gchr + (size_t)b * NCHR * NINT + a;
#pragma unroll
    for (int z = 0; z < NCHR; z++) G[z] = gp[(size_t)z * NINT];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        double s = 0.0;
#pragma unroll
        for (int n = 0; n < 4; n++)
#pragma unroll
            for (int l = 0; l < 4; l++) {
                // the 26 structurally zero entries enter as +0.0 (the product
                // still propagates a non-finite T like the reference's)
                const int z = chr_slot(m, n, l);
                s += T[n][l] * (z >= 0 ? G[z] : 0.0);
            }
        src[m] = s * sgc;
    }
}

// Conserved update on active zones of every block (integrator.cpp:238-259;
// quirks Q20, Q22, Q31): dU = div F + S; stage 0 stores U and takes an Euler
// step, stage 1 averages; U.B is reset from the cell-centred field.
__global__ __launch_bounds__(TPB) void k_update(KP P, int stage) {
    const int wg = xcd_local(blockIdx.x, gridDim.x);
    const int b = wg / WG_INT;
    const int a = (wg % WG_INT) * TPB + threadIdx.x;
    const int i = a / 256, j = (a / 16) % 16, k = a % 16;
    const int c = lindx(i, j, k);
    const double dt = __longlong_as_double((long long)*P.dtbits);
    const double dx = P.dxyz[3 * b], dy = P.dxyz[3 * b + 1], dz = P.dxyz[3 * b + 2];
    const double sg = cell_sqrtdetg(P, b, c);
    const double* Fx = P.F + ((size_t)b * 3 + 0) * 8 * NFC;
    const double* Fy = P.F + ((size_t)b * 3 + 1) * 8 * NFC;
    const double* Fz = P.F + ((size_t)b * 3 + 2) * 8 * NFC;
    const int xR = face_lf(0, i, j, k), xL = face_lf(0, i - 1, j, k);
    const int yR = face_lf(1, i, j, k), yL = face_lf(1, i, j - 1, k);
    const int zR = face_lf(2, i, j, k), zL = face_lf(2, i, j, k - 1);
    int nf = 0;
    Prim W = loadW(P, b, c);
    double src[4];
    source_terms(P, b, c, a, W, src, nf);
    // U component v in {D, S0, S1, S2, tau}; source slot: tau <- src0, S_d <- src_{d+1}.
    // Every load precedes the first store (U, Us and F are all double arrays).
    double fxr[5], fxl[5], fyr[5], fyl[5], fzr[5], fzl[5], u[5], us[5];
    double* usp = P.Us + (size_t)b * 5 * NINT + a;
#pragma unroll
    for (int v = 0; v < 5; v++) {
        const size_t o = (size_t)v * NFC;
        fxr[v] = Fx[o + xR]; fxl[v] = Fx[o + xL];
        fyr[v] = Fy[o + yR]; fyl[v] = Fy[o + yL];
        fzr[v] = Fz[o + zR]; fzl[v] = Fz[o + zL];
        u[v] = Ua(P, b, v, c);
        us[v] = stage == 0 ? 0.0 : usp[(size_t)v * NINT];
    }
    double nu[5];
#pragma unroll
    for (int v = 0; v < 5; v++) {
        double div = -((fxr[v] - fxl[v]) / dx + (fyr[v] - fyl[v]) / dy + (fzr[v] - fzl[v]) / dz) / sg;
        double s = v == UD ? 0.0 : (v == UTAU ? src[0] : src[v]);
        double dU = div + s;
        nu[v] = stage == 0 ? u[v] + dU * dt : ((us[v] + u[v]) + dU * dt) * 0.5;
    }
#pragma unroll
    for (int v = 0; v < 5; v++) {
        if (stage == 0) usp[(size_t)v * NINT] = u[v];
        Ua(P, b, v, c) = nu[v];
    }
    // U.B = sqrt(-g) W.B (integrator.cpp:252-255) is already in place: k_f2cb
    // stored exactly these products over [-1, 17]^3 earlier in the stage.
    cst_at(P, b, c) = 0;
    add_faults(P, nf);
}

// ------------------------------------------------------------------- dt
// rk2integrator::dtcomp, integrator.cpp:156-189: min over leaf active zones
// and directions of cfl*dx/|alpha(v +- c_s) - beta| (cell metric, quirk Q23).
// Candidates are positive, so their bit patterns order like their values.
__global__ void k_dtinit(KP P) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *P.dtbits = (unsigned long long)__double_as_longlong(DBL_MAX);
}
__global__ __launch_bounds__(TPB) void k_dtcomp(KP P, const int* __restrict__ leaves, double cfl,
                                                unsigned long long* __restrict__ partial) {
    const int b = leaves[blockIdx.x / WG_INT];
    const int a = (blockIdx.x % WG_INT) * TPB + threadIdx.x;
    const int c = lindx(a / 256, (a / 16) % 16, a % 16);
    int nf = 0;
    const double rho = Wa(P, b, WRHO, c), eps = Wa(P, b, WEPS, c);
    const double cs = sqrt(eos_cs2(P.gamma, rho, eps, nf));
    const double alpha = gcc_at(P, b, 1, c);
    const double beta[3] = {gcc_at(P, b, 2, c), 0.0, 0.0};
    unsigned long long best = (unsigned long long)__double_as_longlong(DBL_MAX);
    for (int d = 0; d < 3; d++) {
        double v = Wa(P, b, WV0 + d, c);
        double ld_p = fabs(alpha * (v + cs) - beta[d]);
        double ld_m = fabs(alpha * (v - cs) - beta[d]);
        double vmax = smax(ld_p, ld_m);
        if (vmax > 1e-14) {
            double cand = cfl * P.dxyz[3 * b + d] / vmax;
            unsigned long long cb = (unsigned long long)__double_as_longlong(cand);
            best = cb < best ? cb : best;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(best, off);
        best = o < best ? o : best;
    }
    __shared__ unsigned long long wmin[TPB / 64];
    if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = wmin[0];
        for (int w = 1; w < TPB / 64; w++) m = wmin[w] < m ? wmin[w] : m;
        partial[blockIdx.x] = m;
    }
    add_faults(P, nf);
}
// min over the per-workgroup candidates (one workgroup)
__global__ __launch_bounds__(TPB) void k_dtreduce(KP P, const unsigned long long* __restrict__ partial, int n) {
    unsigned long long best = (unsigned long long)__double_as_longlong(DBL_MAX);
    for (int i = threadIdx.x; i < n; i += TPB) best = partial[i] < best ? partial[i] : best;
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(best, off);
        best = o < best ? o : best;
    }
    __shared__ unsigned long long wmin[TPB / 64];
    if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = wmin[0];
        for (int w = 1; w < TPB / 64; w++) m = wmin[w] < m ? wmin[w] : m;
        *P.dtbits = m;
    }
}

// --------------------------------------------------------------- floors
// patch::floors on active zones of every block, grid.cpp:331-358
__global__ __launch_bounds__(TPB) void k_floors(KP P) {
    const int b = blockIdx.x / WG_INT;
    const int a = (blockIdx.x % WG_INT) * TPB + threadIdx.x;
    const int c = lindx(a / 256, (a / 16) % 16, a % 16);
    int nf = 0;
    GPt m = cell_tree(P, b, c);
    const double D = Ua(P, b, UD, c), tau = Ua(P, b, UTAU, c);
    bool bhl = !isfinite(D) || !isfinite(tau) || D <= 0.0 || tau <= 0.0;
    Prim pv;
    bool have = false;
    if (!bhl) {
        int r = c2p_first(P, b, c, m, Q_FLOOR, pv, nf);
        if (r < 0) { add_faults(P, nf); return; }
        have = r > 0;
        bhl = r == 0;
    }
    floors_finish(P, b, c, m, bhl, pv, have, nf);
    add_faults(P, nf);
}

// ---------------------------------------------------------- diagnostics
// analysis::analyser quantities that read only the evolved state
// (src/fluid/data/quantitatives.cpp): per active zone of every leaf, the
// terms of mass (:82-86), energy (:87-91), L_pol (:92-96), thermE
// (:129-133) and, on the shells |r - r_s| <= |(dx, dy, dz)| of surf_int
// (:61-79), Edot (:103-107) and Phi_B (:114-118) at the horizon and the
// Hscal numerator and denominator (:199-208) at rs; each term is formed with
// the reference's operations (cell centre, r and volume as patch::cell_init,
// grid.cpp:149-175; sqrt(-g), alpha, beta of the cell metric).  Sums run per
// workgroup, then over workgroups, in a fixed order.
constexpr int NDIAG = 8;
__device__ __forceinline__ void block_sum(double v[NDIAG], double* lds) {
    // fixed-order tree over the workgroup: lds[NDIAG][TPB]
    for (int q = 0; q < NDIAG; q++) lds[q * TPB + threadIdx.x] = v[q];
    __syncthreads();
    for (int h = TPB / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int q = 0; q < NDIAG; q++) lds[q * TPB + threadIdx.x] += lds[q * TPB + threadIdx.x + h];
        __syncthreads();
    }
}
__global__ __launch_bounds__(TPB) void k_diag(KP P, const int* __restrict__ leaves, double hrz, double rs,
                                              double* __restrict__ partial) {
    __shared__ double lds[NDIAG * TPB];
    const int b = leaves[blockIdx.x / WG_INT];
    const int a = (blockIdx.x % WG_INT) * TPB + threadIdx.x;
    const int i = a / 256, j = (a / 16) % 16, k = a % 16;
    const int c = lindx(i, j, k);
    const double dx = P.dxyz[3 * b], dy = P.dxyz[3 * b + 1], dz = P.dxyz[3 * b + 2];
    const double xc = P.org[3 * b] + (i + 0.5) * dx, yc = P.org[3 * b + 1] + (j + 0.5) * dy,
                 zc = P.org[3 * b + 2] + (k + 0.5) * dz;
    const double vol = dx * dy * dz;
    double r = sqrt(xc * xc + yc * yc + zc * zc);
    if (r < 1e-10) r = 1e-10;
    const double shell = sqrt(dx * dx + dy * dy + dz * dz);
    const double sg = gcc_at(P, b, 0, c), alpha = gcc_at(P, b, 1, c), beta0 = gcc_at(P, b, 2, c);
    const double D = Ua(P, b, UD, c), S0 = Ua(P, b, US0, c), S1 = Ua(P, b, US1, c), tau = Ua(P, b, UTAU, c);
    const Prim w = loadW(P, b, c);
    const double vr = (xc * w.v[0] + yc * w.v[1] + zc * w.v[2]) / r;
    const double Br = (xc * w.B[0] + yc * w.B[1] + zc * w.B[2]) / r;
    const bool on_hrz = !(fabs(r - hrz) > shell), on_rs = !(fabs(r - rs) > shell);
    double v[NDIAG];
    v[0] = D * vol;
    v[1] = (tau + D) * vol;
    v[2] = (xc * S1 - yc * S0) * vol;
    // beta^1 = beta^2 = +0.0 in the cell metric
    v[3] = on_hrz ? -((tau + D) * (alpha * vr - (beta0 * xc + 0.0 * yc + 0.0 * zc) / r)) * vol : 0.0;
    v[4] = on_hrz ? fabs(Br) * sg * vol : 0.0;
    v[5] = w.rho * w.eps * sg * vol;
    v[6] = on_rs ? fabs(zc) * w.rho * sg * vol : 0.0;
    v[7] = on_rs ? w.rho * sg * vol : 0.0;
    block_sum(v, lds);
    if (threadIdx.x < NDIAG) partial[(size_t)blockIdx.x * NDIAG + threadIdx.x] = lds[threadIdx.x * TPB];
}
// one workgroup: sum the per-workgroup partials in a fixed order
__global__ __launch_bounds__(TPB) void k_diag_reduce(const double* __restrict__ partial, int n,
                                                     double* __restrict__ out) {
    __shared__ double lds[NDIAG * TPB];
    double v[NDIAG];
    for (int q = 0; q < NDIAG; q++) v[q] = 0.0;
    for (int p = threadIdx.x; p < n; p += TPB)
        for (int q = 0; q < NDIAG; q++) v[q] += partial[(size_t)p * NDIAG + q];
    block_sum(v, lds);
    if (threadIdx.x < NDIAG) out[threadIdx.x] = lds[threadIdx.x * TPB];
}

// ---------------------------------------------------------- initial data
// The set-up of src/fluid/main.cpp:94-159 on the device: the constructor's
// floor state, fm_init, the global rho_max, B_pot_init and ptoc of the active
// zones.  Cell r and theta, the floor state and the metric come from the
// geometry tables; the transcendental functions of the initial data are the
// correctly rounded ones of bh_crtrig.hpp.
struct FmPar { double M, a, Q, Gamma, K, l0, W_in, r_H, rho_scal; };

__device__ __forceinline__ void cell_centre(const KP& P, int b, int i, int j, int k, double& xc, double& yc,
                                            double& zc) {
    xc = P.org[3 * b] + (i + 0.5) * P.dxyz[3 * b];
    yc = P.org[3 * b + 1] + (j + 0.5) * P.dxyz[3 * b + 1];
    zc = P.org[3 * b + 2] + (k + 0.5) * P.dxyz[3 * b + 2];
}
// maximum of non-negative doubles by their bit patterns, one atomic per workgroup
__device__ __forceinline__ void wg_max_bits(unsigned long long v, unsigned long long* out) {
    __shared__ unsigned long long red[TPB];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int h = TPB / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] = red[threadIdx.x] > red[threadIdx.x + h] ? red[threadIdx.x] : red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0] != 0ull) atomicMax(out, red[0]);
}

// amrtree constructor, grid.cpp:177-206: floor state and its conserved form
// in every stored cell
__global__ __launch_bounds__(TPB) void k_ic_atmosphere(KP P) {
    const int b = blockIdx.x / WG_CELL;
    const int c = (blockIdx.x % WG_CELL) * TPB + threadIdx.x;
    if (c >= NCELL) return;
    const double* fl = P.gfl + (size_t)b * 2 * NCELL;
    const Prim w = floor_prim(fl[c], fl[NCELL + c]);
    int nf = 0;
    const Cons u = ptoc(cell_tree(P, b, c), P.gamma, w, nf);
    storeW(P, b, c, w);
    storeU(P, b, c, u);
    add_faults(P, nf);
}
// One active zone of fm_init (initconds.cpp:446-470 and 476-589): false where
// the zone keeps the atmosphere (near the axis or the horizon, or ln h <= 0),
// else the unnormalised torus density and the metric parts at the zone.
__device__ __forceinline__ bool fm_zone(const KP& P, const FmPar& f, int b, int i, int j, int k, double& xc,
                                        double& yc, KsParts& mc, double& rho) {
    double zc;
    cell_centre(P, b, i, j, k, xc, yc, zc);
    double r = sqrt(xc * xc + yc * yc + zc * zc);
    if (r < 1e-10) r = 1e-10;
    const double th = P.gth[(size_t)b * NCELL + lindx(i, j, k)];
    if (!isfinite(r) || !isfinite(th) || r <= 1.05 * f.r_H || fabs(cr_sin(th)) < 0.01) return false;
    mc = ks_parts(f.M, f.a, f.Q, r, th);
    const double lnH = fm_lnh(mc, f.l0, f.W_in);
    if (lnH <= 0.0 || !isfinite(lnH)) return false;
    rho = fm_rho(lnH, f.Gamma, f.K);
    return true;
}
// the first pass of fm_init: maximum finite torus density (initconds.cpp:446-470)
__global__ __launch_bounds__(TPB) void k_fm_rho_max(KP P, FmPar f, unsigned long long* __restrict__ out) {
    const int b = blockIdx.x / WG_INT;
    const int a = (blockIdx.x % WG_INT) * TPB + threadIdx.x;
    double xc, yc, rho = 0.0;
    KsParts mc;
    unsigned long long v = 0ull;
    if (fm_zone(P, f, b, a / 256, (a / 16) % 16, a % 16, xc, yc, mc, rho) && isfinite(rho) && rho > 0.0)
        v = (unsigned long long)__double_as_longlong(rho);
    wg_max_bits(v, out);
}
// the second pass of fm_init: torus state, normalised to rho_max = 1, or the
// atmosphere (initconds.cpp:476-589)
__global__ __launch_bounds__(TPB) void k_fm_write(KP P, FmPar f) {
    const int b = blockIdx.x / WG_INT;
    const int a = (blockIdx.x % WG_INT) * TPB + threadIdx.x;
    const int i = a / 256, j = (a / 16) % 16, k = a % 16;
    const int c = lindx(i, j, k);
    const double* fl = P.gfl + (size_t)b * 2 * NCELL;
    Prim W = floor_prim(fl[c], fl[NCELL + c]);
    double xc, yc, rho;
    KsParts mc;
    if (fm_zone(P, f, b, i, j, k, xc, yc, mc, rho)) {
        W.rho = f.rho_scal * rho;
        const double pp = f.K * cr_pow(W.rho, f.Gamma);
        W.eps = pp / ((f.Gamma - 1.0) * W.rho);
        fm_velocity(mc, f.l0, xc, yc, W.v);
    }
    storeW(P, b, c, W);
}
// global maximum rho over active zones, main.cpp:101-111 (std::max from 0)
__global__ __launch_bounds__(TPB) void k_ic_rho_max(KP P, unsigned long long* __restrict__ out) {
    const int b = blockIdx.x / WG_INT;
    const int a = (blockIdx.x % WG_INT) * TPB + threadIdx.x;
    const double rho = Wa(P, b, WRHO, lindx(a / 256, (a / 16) % 16, a % 16));
    wg_max_bits(rho > 0.0 ? (unsigned long long)__double_as_longlong(rho) : 0ull, out);
}
// B_pot_init's patch rho_max: the global value raised to the patch's active
// maximum (initconds.cpp:597-606); one workgroup per block
__global__ __launch_bounds__(TPB) void k_bpot_max(KP P, double glmx, double* __restrict__ pmax) {
    __shared__ double red[TPB];
    const int b = blockIdx.x;
    double m = glmx;
    for (int a = threadIdx.x; a < NINT; a += TPB) m = smax(m, Wa(P, b, WRHO, lindx(a / 256, (a / 16) % 16, a % 16)));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int h = TPB / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] = smax(red[threadIdx.x], red[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) pmax[b] = red[0];
}
// Face fields from the curl of A at the edges (initconds.cpp:608-655, quirk:
// the z-face takes A_y on x-edges and A_x on y-edges) for i, j, k in [-1, 17]
__global__ __launch_bounds__(TPB) void k_bpot_faces(KP P, const double* __restrict__ pmax) {
    const int b = blockIdx.x / WG_EDGE;
    const int n = (blockIdx.x % WG_EDGE) * TPB + threadIdx.x;
    if (n >= 19 * 19 * 19) return;
    const int i = n / 361 - 1, j = (n / 19) % 19 - 1, k = n % 19 - 1;
    const double rho_max = pmax[b];
    const double dx = P.dxyz[3 * b], dy = P.dxyz[3 * b + 1], dz = P.dxyz[3 * b + 2];
    const int top = BLK + GH - 1;
    auto rho = [&](int ii, int jj, int kk) { return Wa(P, b, WRHO, lindx(ii, jj, kk)); };
    // edge densities, initconds.cpp:25-39
    auto rhox = [&](int ii, int jj, int kk) {
        const int j1 = min(jj + 1, top), k1 = min(kk + 1, top);
        return (rho(ii, jj, kk) + rho(ii, j1, kk) + rho(ii, jj, k1) + rho(ii, j1, k1)) / 4;
    };
    auto rhoy = [&](int ii, int jj, int kk) {
        const int i1 = min(ii + 1, top), k1 = min(kk + 1, top);
        return (rho(ii, jj, kk) + rho(i1, jj, kk) + rho(ii, jj, k1) + rho(i1, jj, k1)) / 4;
    };
    // A at the x-edge (xc, yc + dy/2, zc + dz/2) and y-edge (xc + dx/2, yc, zc + dz/2)
    // of cell (ii, jj, kk), initconds.cpp:41-48
    auto Axe = [&](int ii, int jj, int kk, double& Ax, double& Ay) {
        double xc, yc, zc;
        cell_centre(P, b, ii, jj, kk, xc, yc, zc);
        fm_A_xy(xc, yc + dy / 2.0, zc + dz / 2.0, rhox(ii, jj, kk), rho_max, Ax, Ay);
    };
    auto Aye = [&](int ii, int jj, int kk, double& Ax, double& Ay) {
        double xc, yc, zc;
        cell_centre(P, b, ii, jj, kk, xc, yc, zc);
        fm_A_xy(xc + dx / 2.0, yc, zc + dz / 2.0, rhoy(ii, jj, kk), rho_max, Ax, Ay);
    };
    double xe_x, xe_y, xem_x, xem_y, xei_x, xei_y;  // x-edges at (i,j,k), (i,j,k-1), (i-1,j,k)
    double ye_x, ye_y, yem_x, yem_y, yej_x, yej_y;  // y-edges at (i,j,k), (i,j,k-1), (i,j-1,k)
    Axe(i, j, k, xe_x, xe_y);
    Axe(i, j, k - 1, xem_x, xem_y);
    Axe(i - 1, j, k, xei_x, xei_y);
    Aye(i, j, k, ye_x, ye_y);
    Aye(i, j, k - 1, yem_x, yem_y);
    Aye(i, j - 1, k, yej_x, yej_y);
    const double Az = 0.0;  // A_z on the z-edges
    double* Bf = P.Bf + (size_t)b * 3 * NFACE;
    Bf[ixf(i, j, k)] = (Az - Az) / dy - (ye_y - yem_y) / dz;
    Bf[NFACE + iyf(i, j, k)] = (xe_x - xem_x) / dz - (Az - Az) / dx;
    Bf[2 * NFACE + izf(i, j, k)] = (xe_y - xei_y) / dx - (ye_x - yej_x) / dy;
}
// constrans::f2cB with the cell metric (transport.cpp:60-75) over
// i, j, k in [-1, 17]: W.B the face average, U.B = sqrt(-g) B
__global__ __launch_bounds__(TPB) void k_bpot_f2cb(KP P) {
    const int b = blockIdx.x / WG_EDGE;
    const int n = (blockIdx.x % WG_EDGE) * TPB + threadIdx.x;
    if (n >= 19 * 19 * 19) return;
    const int i = n / 361 - 1, j = (n / 19) % 19 - 1, k = n % 19 - 1;
    const int c = lindx(i, j, k);
    const double* Bf = P.Bf + (size_t)b * 3 * NFACE;
    const double b0 = (Bf[ixf(i, j, k)] + Bf[ixf(i - 1, j, k)]) / 2;
    const double b1 = (Bf[NFACE + iyf(i, j, k)] + Bf[NFACE + iyf(i, j - 1, k)]) / 2;
    const double b2 = (Bf[2 * NFACE + izf(i, j, k)] + Bf[2 * NFACE + izf(i, j, k - 1)]) / 2;
    const double sg = cell_sqrtdetg(P, b, c);
    Wa(P, b, WB0, c) = b0; Wa(P, b, WB1, c) = b1; Wa(P, b, WB2, c) = b2;
    Ua(P, b, UB0, c) = sg * b0; Ua(P, b, UB1, c) = sg * b1; Ua(P, b, UB2, c) = sg * b2;
}
// U = ptoc(W) on active zones, main.cpp:148-159
__global__ __launch_bounds__(TPB) void k_ptoc_interior(KP P) {
    const int b = blockIdx.x / WG_INT;
    const int a = (blockIdx.x % WG_INT) * TPB + threadIdx.x;
    const int c = lindx(a / 256, (a / 16) % 16, a % 16);
    int nf = 0;
    storeU(P, b, c, ptoc(cell_tree(P, b, c), P.gamma, loadW(P, b, c), nf));
    add_faults(P, nf);
}

// ---------------------------------------------------------------- regrid
// The flagging loop of amrtree::regrid (grid.cpp:744-795) over the active
// zones of every leaf: bit 0 of flags[leaf] = some zone asks for refinement
// (r < 3 at level < 4, or a relative rho or B gradient above 0.5 at level
// < 4), bit 1 = some zone forbids flattening (the same, or a relative rho
// gradient above 0.1).  r, dx, dy, dz as patch::cell_init.
__global__ __launch_bounds__(TPB) void k_regrid_flags(KP P, const int* __restrict__ leaves,
                                                      const int* __restrict__ lvl, unsigned int* __restrict__ flags) {
    __shared__ unsigned int wg_flags;
    if (threadIdx.x == 0) wg_flags = 0u;
    __syncthreads();
    const int l = blockIdx.x / WG_INT;
    const int b = leaves[l];
    const int a = (blockIdx.x % WG_INT) * TPB + threadIdx.x;
    const int i = a / 256, j = (a / 16) % 16, k = a % 16;
    const int c = lindx(i, j, k);
    const int level = lvl[b];
    const double dx = P.dxyz[3 * b], dy = P.dxyz[3 * b + 1], dz = P.dxyz[3 * b + 2];
    const double xc = P.org[3 * b] + (i + 0.5) * dx, yc = P.org[3 * b + 1] + (j + 0.5) * dy,
                 zc = P.org[3 * b + 2] + (k + 0.5) * dz;
    double r = sqrt(xc * xc + yc * yc + zc * zc);
    if (r < 1e-10) r = 1e-10;
    // centred difference of word w along an axis, over twice the spacing
    const double h[3] = {dx, dy, dz};
    auto centred = [&](int w, int axis) {
        const int up = axis == 0 ? lindx(i + 1, j, k) : axis == 1 ? lindx(i, j + 1, k) : lindx(i, j, k + 1);
        const int dn = axis == 0 ? lindx(i - 1, j, k) : axis == 1 ? lindx(i, j - 1, k) : lindx(i, j, k - 1);
        return fabs(Wa(P, b, w, up) - Wa(P, b, w, dn)) / (2 * h[axis]);
    };
    const bool deeper = level < 4;
    unsigned int f = 0u;
    if (r < 3.0 && deeper) {
        f = 3u;
    } else {
        const double rho = Wa(P, b, WRHO, c);
        if (rho > 0.0) {
            const double gx = centred(WRHO, 0), gy = centred(WRHO, 1), gz = centred(WRHO, 2);
            const double contrast = sqrt(gx * gx + gy * gy + gz * gz) * dx / rho;
            if (contrast > 0.5 && deeper) f |= 3u;
            if (contrast > 0.1) f |= 2u;
        }
        const double b0 = Wa(P, b, WB0, c), b1 = Wa(P, b, WB1, c), b2 = Wa(P, b, WB2, c);
        const double bmag = sqrt(b0 * b0 + b1 * b1 + b2 * b2);
        if (bmag > 1e-14) {
            const double var = (centred(WB0, 0) * dx + centred(WB1, 1) * dy + centred(WB2, 2) * dz) / bmag;
            if (var > 0.5 && deeper) f |= 3u;
        }
    }
    if (f) atomicOr(&wg_flags, f);
    __syncthreads();
    if (threadIdx.x == 0 && wg_flags) atomicOr(flags + l, wg_flags);
}

// Copy per-block slabs of `words` doubles into their new slots: dst block b
// takes src block map[b] (map[b] < 0: left alone).
__global__ __launch_bounds__(TPB) void k_gather(double* __restrict__ dst, const double* __restrict__ src,
                                                const int* __restrict__ map, long words) {
    const int b = blockIdx.y;
    const int s = map[b];
    if (s < 0) return;
    for (long w = (long)blockIdx.x * TPB + threadIdx.x; w < words; w += (long)gridDim.x * TPB)
        dst[(size_t)b * words + w] = src[(size_t)s * words + w];
}

// floor_or_keep of the regrid's data work: con2prim of the zone's U with
// the tree metric; on failure the floor state keeps the zone's B and U is
// recomputed (grid.cpp:637-647, 705-714).
__device__ __forceinline__ void regrid_c2p(const KP& P, int b, int c, int& nf) {
    GPt m = cell_tree(P, b, c);
    Prim pv;
    if (ctop(m, P.gamma, loadU(P, b, c), pv, nf)) {
        storeW(P, b, c, pv);
    } else {
        const double* fl = P.gfl + (size_t)b * 2 * NCELL;
        Prim f = floor_prim(fl[c], fl[NCELL + c]);
        f.B[0] = Wa(P, b, WB0, c); f.B[1] = Wa(P, b, WB1, c); f.B[2] = Wa(P, b, WB2, c);
        storeW(P, b, c, f);
        storeU(P, b, c, ptoc(m, P.gamma, f, nf));
    }
}
// amrtree::refine's data (grid.cpp:599-653, quirk Q24: the slope uses the +
// neighbour as "minus"): child zone U from the parent's U (old layout, slot
// par[n]) by limited linear prolongation, then floor_or_keep.  One
// workgroup row per (refined parent n, child h); ch[8 n + h] = child slot.
__global__ __launch_bounds__(TPB) void k_refine_zones(KP P, const double* __restrict__ Uold,
                                                      const int* __restrict__ par, const int* __restrict__ ch) {
    const int n = blockIdx.y / 8, h = blockIdx.y % 8;
    const int a = blockIdx.x * TPB + threadIdx.x;
    const int i = a / 256, j = (a / 16) % 16, k = a % 16;
    const int ix = (h >> 2) & 1, iy = (h >> 1) & 1, iz = h & 1;
    const int pi = (i + ix * BLK) / 2, pj = (j + iy * BLK) / 2, pk = (k + iz * BLK) / 2;
    const double ox = (i % 2 == 0) ? -0.25 : 0.25;
    const double oy = (j % 2 == 0) ? -0.25 : 0.25;
    const double oz = (k % 2 == 0) ? -0.25 : 0.25;
    const int pc = lindx(pi, pj, pk);
    const int ppx = lindx(pi + 1, pj, pk), pmx = lindx(pi - 1, pj, pk);
    const int ppy = lindx(pi, pj + 1, pk), pmy = lindx(pi, pj - 1, pk);
    const int ppz = lindx(pi, pj, pk + 1), pmz = lindx(pi, pj, pk - 1);
    const double* up = Uold + (size_t)par[n] * 8 * NCELL;
    const int b = ch[8 * n + h];
    const int cc = lindx(i, j, k);
    double nu[8];
    for (int v = 0; v < 8; v++) {
        const double* u = up + (size_t)v * NCELL;
        const double vc = u[pc];
        double sx = minmod(vc - u[ppx], u[pmx] - vc);
        double sy = minmod(vc - u[ppy], u[pmy] - vc);
        double sz = minmod(vc - u[ppz], u[pmz] - vc);
        nu[v] = vc + sx * ox + sy * oy + sz * oz;
    }
    for (int v = 0; v < 8; v++) Ua(P, b, v, cc) = nu[v];
    int nf = 0;
    regrid_c2p(P, b, cc, nf);
    add_faults(P, nf);
}
// amrtree::ccrstr (grid.cpp:657-718): parent zone U = volume-weighted mean
// of the children's zones it covers (old layout, slots ch[8 n + h], -1 =
// none), then floor_or_keep.  One workgroup row per flattened parent.
__global__ __launch_bounds__(TPB) void k_restrict_zones(KP P, const double* __restrict__ Uold,
                                                        const double* __restrict__ dxyz_old,
                                                        const int* __restrict__ par, const int* __restrict__ ch) {
    const int n = blockIdx.y;
    const int a = blockIdx.x * TPB + threadIdx.x;
    const int pi = a / 256, pj = (a / 16) % 16, pk = a % 16;
    const int b = par[n];
    const int pc = lindx(pi, pj, pk);
    double acc[8];
    for (int v = 0; v < 8; v++) acc[v] = 0.0;
    double volt = 0.0;
    for (int h = 0; h < 8; h++) {
        const int cs = ch[8 * n + h];
        if (cs < 0) continue;
        const int ix = (h >> 2) & 1, iy = (h >> 1) & 1, iz = h & 1;
        const double vol = dxyz_old[3 * cs] * dxyz_old[3 * cs + 1] * dxyz_old[3 * cs + 2];
        const double* uc = Uold + (size_t)cs * 8 * NCELL;
        for (int di = 0; di < 2; di++)
            for (int dj = 0; dj < 2; dj++)
                for (int dk = 0; dk < 2; dk++) {
                    const int fi = 2 * (pi - ix * (BLK / 2)) + di;
                    const int fj = 2 * (pj - iy * (BLK / 2)) + dj;
                    const int fk = 2 * (pk - iz * (BLK / 2)) + dk;
                    if (fi < 0 || fi >= BLK || fj < 0 || fj >= BLK || fk < 0 || fk >= BLK) continue;
                    const int fc = lindx(fi, fj, fk);
                    volt += vol;
                    for (int v = 0; v < 8; v++) acc[v] += vol * uc[(size_t)v * NCELL + fc];
                }
    }
    if (volt > 0.0) {
        const double inv = 1.0 / volt;
        for (int v = 0; v < 8; v++) acc[v] *= inv;
    }
    for (int v = 0; v < 8; v++) Ua(P, b, v, pc) = acc[v];
    int nf = 0;
    regrid_c2p(P, b, pc, nf);
    add_faults(P, nf);
}
// patch::B_init (grid.cpp:208-251, quirk Q28: the z sweep tests i, not k)
// of the listed blocks: face fields from the stored cells' W.B.
__global__ __launch_bounds__(TPB) void k_binit(KP P, const int* __restrict__ blocks) {
    const int b = blocks[blockIdx.y];
    const int t = blockIdx.x * TPB + threadIdx.x;
    if (t >= 3 * NFACE) return;
    const int dir = t / NFACE, f = t % NFACE;
    auto WB = [&](int d, int i, int j, int k) { return Wa(P, b, WB0 + d, lindx(i, j, k)); };
    double v;
    int i, j, k, idx;
    if (dir == 0) {  // i in [-2, 18], j, k in [-2, 18)
        i = f / 400 - GH; j = (f / 20) % 20 - GH; k = f % 20 - GH;
        if (i > -GH && i < BLK + GH) v = (WB(0, i - 1, j, k) + WB(0, i, j, k)) / 2;
        else if (i == -GH) v = WB(0, i, j, k);
        else v = WB(0, i - 1, j, k);
        idx = ixf(i, j, k);
    } else if (dir == 1) {  // j in [-2, 18]
        i = f / 420 - GH; j = (f / 20) % 21 - GH; k = f % 20 - GH;
        if (j > -GH && j < BLK + GH) v = (WB(1, i, j - 1, k) + WB(1, i, j, k)) / 2;
        else if (j == -GH) v = WB(1, i, j, k);
        else v = WB(1, i, j - 1, k);
        idx = iyf(i, j, k);
    } else {  // k in [-2, 18]
        i = f / 420 - GH; j = (f / 21) % 20 - GH; k = f % 21 - GH;
        if (k > -GH && k < BLK + GH) v = (WB(2, i, j, k - 1) + WB(2, i, j, k)) / 2;
        else if (i == -GH) v = WB(2, i, j, k);
        else v = WB(2, i, j, k - 1);
        idx = izf(i, j, k);
    }
    P.Bf[((size_t)b * 3 + dir) * NFACE + idx] = v;
}
// Exact-copy flags of the ghost plan (see mark_exact_copies) from the
// device's tree-metric tables: one workgroup per (leaf, side).
__global__ __launch_bounds__(TPB) void k_mark_same(KP P, const int* __restrict__ leaves,
                                                   unsigned char* __restrict__ same) {
    __shared__ int differ;
    if (threadIdx.x == 0) differ = 0;
    __syncthreads();
    const int b = leaves[blockIdx.x / 6], side = blockIdx.x % 6;
    const int kind = P.ghtab[b * 12 + side * 2], q = P.ghtab[b * 12 + side * 2 + 1];
    if (kind != GH_COPY) {
        if (threadIdx.x == 0) same[b * 6 + side] = 0;
        return;
    }
    const int dim = side / 2, s = side % 2;
    int d = 0;
    for (int t = threadIdx.x; t < GH * BLK * BLK; t += TPB) {
        const int gl = t / (BLK * BLK), t1 = (t / BLK) % BLK, t2 = t % BLK;
        const int c = cell_of(dim, s ? BLK + gl : gl - GH, t1, t2);
        const int sc = cell_of(dim, (s ? 0 : BLK - GH) + gl, t1, t2);
        for (int v = 0; v < NGCT; v++) {
            const double x = P.gct[((size_t)b * NGCT + v) * NCELL + c];
            const double y = P.gct[((size_t)q * NGCT + v) * NCELL + sc];
            if (__double_as_longlong(x) != __double_as_longlong(y)) d = 1;
        }
    }
    if (d) atomicOr(&differ, 1);
    __syncthreads();
    if (threadIdx.x == 0) same[b * 6 + side] = differ ? 0 : 1;
}

// constrans::maxdivB (src/fluid/grid/transport.cpp:77-89) of every listed
// block: max over cells [-1, 17)^3 of |div B| of the face fields, one
// workgroup per block; the maximum keeps the reference's std::max semantics
// (a NaN divergence leaves the running maximum unchanged).
__global__ __launch_bounds__(TPB) void k_maxdivb(KP P, const int* __restrict__ blocks, double* __restrict__ out) {
    __shared__ double sm[TPB];
    const int b = blocks[blockIdx.x];
    const double dx = P.dxyz[3 * b], dy = P.dxyz[3 * b + 1], dz = P.dxyz[3 * b + 2];
    const double* bx = P.Bf + ((size_t)b * 3 + 0) * NFACE;
    const double* by = P.Bf + ((size_t)b * 3 + 1) * NFACE;
    const double* bz = P.Bf + ((size_t)b * 3 + 2) * NFACE;
    constexpr int N = BLK + 2 * GH - 2;  // 18 cells per direction
    double mx = 0.0;
    for (int t = threadIdx.x; t < N * N * N; t += TPB) {
        const int i = t / (N * N) - GH + 1, j = (t / N) % N - GH + 1, k = t % N - GH + 1;
        const double dv = (bx[ixf(i, j, k)] - bx[ixf(i - 1, j, k)]) / dx + (by[iyf(i, j, k)] - by[iyf(i, j - 1, k)]) / dy +
                          (bz[izf(i, j, k)] - bz[izf(i, j, k - 1)]) / dz;
        mx = smax(mx, fabs(dv));
    }
    sm[threadIdx.x] = mx;
    __syncthreads();
    for (int o = TPB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sm[threadIdx.x] = smax(sm[threadIdx.x], sm[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sm[0];
}

// Protect mask of the early con2prim pass (KP::pmask): for every leaf side
// that copies from a same-level block without reusing the source's outcome,
// mark the source's opposite side.
__global__ void k_pmask(KP P, const int* __restrict__ leaves, int nleaves, int* __restrict__ pmask) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nleaves * 6) return;
    const int b = leaves[t / 6], side = t % 6;
    const int kind = P.ghtab[b * 12 + side * 2], q = P.ghtab[b * 12 + side * 2 + 1];
    if (kind != GH_COPY || q < 0 || P.ghsame[b * 6 + side]) return;
    const int dim = side / 2, s = side % 2;
    atomicOr(pmask + q, 1 << (2 * dim + (1 - s)));
}

// ----------------------------------------------------------- state dump
// W of the active zones of leaves l0 .. l0 + n - 1, packed [n][8][4096]
// (k fastest) for the dump file.
__global__ __launch_bounds__(TPB) void k_pack_w(KP P, const int* __restrict__ leaves, int l0,
                                                double* __restrict__ out) {
    const int l = blockIdx.x / WG_INT;
    const int a = (blockIdx.x % WG_INT) * TPB + threadIdx.x;
    const int b = leaves[l0 + l];
    const int c = lindx(a / 256, (a / 16) % 16, a % 16);
    double w[8];
#pragma unroll
    for (int v = 0; v < 8; v++) w[v] = Wa(P, b, v, c);
#pragma unroll
    for (int v = 0; v < 8; v++) out[((size_t)l * 8 + v) * NINT + a] = w[v];
}

// ---------------------------------------------------- pointwise batches
__global__ void k_ptoc_batch(int n, double gm, const double* g, const double* W, double* U,
                             unsigned long long* faults) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    GPt m;
    m.sqrtdetg = g[10 * t]; m.sqrtdetgam = g[10 * t + 1]; m.alpha = g[10 * t + 2]; m.beta0 = g[10 * t + 3];
    m.g11 = g[10 * t + 4]; m.g13 = g[10 * t + 5]; m.g22 = g[10 * t + 6]; m.g33 = g[10 * t + 7];
    m.gi13 = g[10 * t + 8]; m.gi33 = g[10 * t + 9];
    Prim w;
    w.rho = W[8 * t]; w.eps = W[8 * t + 1];
    for (int d = 0; d < 3; d++) { w.v[d] = W[8 * t + 2 + d]; w.B[d] = W[8 * t + 5 + d]; }
    int nf = 0;
    Cons q = ptoc(m, gm, w, nf);
    U[8 * t] = q.D; U[8 * t + 1] = q.S[0]; U[8 * t + 2] = q.S[1]; U[8 * t + 3] = q.S[2];
    U[8 * t + 4] = q.tau; U[8 * t + 5] = q.B[0]; U[8 * t + 6] = q.B[1]; U[8 * t + 7] = q.B[2];
    if (nf) atomicAdd(faults, (unsigned long long)nf);
}
__global__ void k_ctop_batch(int n, double gm, const double* g, const double* U, double* W, int* ok,
                             unsigned long long* faults) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    GPt m;
    m.sqrtdetg = g[10 * t]; m.sqrtdetgam = g[10 * t + 1]; m.alpha = g[10 * t + 2]; m.beta0 = g[10 * t + 3];
    m.g11 = g[10 * t + 4]; m.g13 = g[10 * t + 5]; m.g22 = g[10 * t + 6]; m.g33 = g[10 * t + 7];
    m.gi13 = g[10 * t + 8]; m.gi33 = g[10 * t + 9];
    Cons q;
    q.D = U[8 * t]; q.S[0] = U[8 * t + 1]; q.S[1] = U[8 * t + 2]; q.S[2] = U[8 * t + 3];
    q.tau = U[8 * t + 4]; q.B[0] = U[8 * t + 5]; q.B[1] = U[8 * t + 6]; q.B[2] = U[8 * t + 7];
    int nf = 0;
    Prim w;
    for (int v = 0; v < 8; v++) W[8 * t + v] = 0.0;
    bool r = ctop(m, gm, q, w, nf);
    if (r) {
        W[8 * t] = w.rho; W[8 * t + 1] = w.eps;
        for (int d = 0; d < 3; d++) { W[8 * t + 2 + d] = w.v[d]; W[8 * t + 5 + d] = w.B[d]; }
    }
    ok[t] = r ? 1 : 0;
    if (nf) atomicAdd(faults, (unsigned long long)nf);
}
__global__ void k_hlld_batch(int n, double gm, int dim, const double* g, const double* WL, const double* WR,
                             const double* UL, const double* UR, double* F, double* ct,
                             unsigned long long* faults) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    GPt m;
    m.sqrtdetg = g[10 * t]; m.sqrtdetgam = g[10 * t + 1]; m.alpha = g[10 * t + 2]; m.beta0 = g[10 * t + 3];
    m.g11 = g[10 * t + 4]; m.g13 = g[10 * t + 5]; m.g22 = g[10 * t + 6]; m.g33 = g[10 * t + 7];
    m.gi13 = g[10 * t + 8]; m.gi33 = g[10 * t + 9];
    auto ldW = [&](const double* A) {
        Prim w;
        w.rho = A[8 * t]; w.eps = A[8 * t + 1];
        for (int d = 0; d < 3; d++) { w.v[d] = A[8 * t + 2 + d]; w.B[d] = A[8 * t + 5 + d]; }
        return w;
    };
    auto ldU = [&](const double* A) {
        Cons q;
        q.D = A[8 * t]; q.S[0] = A[8 * t + 1]; q.S[1] = A[8 * t + 2]; q.S[2] = A[8 * t + 3];
        q.tau = A[8 * t + 4]; q.B[0] = A[8 * t + 5]; q.B[1] = A[8 * t + 6]; q.B[2] = A[8 * t + 7];
        return q;
    };
    int nf = 0;
    double c;
    Cons q = hlld(ldW(WL), ldW(WR), ldU(UL), ldU(UR), m, gm, dim, c, nf);
    F[8 * t] = q.D; F[8 * t + 1] = q.S[0]; F[8 * t + 2] = q.S[1]; F[8 * t + 3] = q.S[2];
    F[8 * t + 4] = q.tau; F[8 * t + 5] = q.B[0]; F[8 * t + 6] = q.B[1]; F[8 * t + 7] = q.B[2];
    ct[t] = c;
    if (nf) atomicAdd(faults, (unsigned long long)nf);
}
// Plain device copy of a known size: the calibration for the HBM traffic
// counters (FETCH_SIZE / WRITE_SIZE per dispatch).
__global__ __launch_bounds__(TPB) void k_copy(const double4* __restrict__ a, double4* __restrict__ b, size_t n) {
    for (size_t i = (size_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (size_t)gridDim.x * TPB) b[i] = a[i];
}
// Proven sign of the con2prim residual next to the residual itself:
// in[i] = (xi, Ssq, Bsq, SB, D, tau), sign[i] = c2p_f_sign, f[i] = c2p_f.
__global__ void k_c2p_sign_batch(int n, double gm, const double* in, int* sign, double* f) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double* a = in + 6 * (size_t)t;
    sign[t] = c2p_f_sign(gm, a[0], a[1], a[2], a[3], a[4], a[5]);
    f[t] = c2p_f(gm, a[0], a[1], a[2], a[3], a[4], a[5]);
}
BH_HD double math1(int op, double x) {
    switch (op) {
        case 0: return cr_log(x);
        case 1: return cr_exp(x);
        case 2: return cr_sin(x);
        case 3: return cr_cos(x);
        default: return cr_acos(x);
    }
}
BH_HD double math2(int op, double x, double y) { return op == 0 ? cr_pow(x, y) : cr_atan2(x, y); }
__global__ void k_math_batch(int op, int n, const double* x, double* y) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    y[t] = math1(op, x[t]);
}
__global__ void k_math2_batch(int op, int n, const double* x, const double* y, double* out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    out[t] = math2(op, x[t], y[t]);
}

// ------------------------------------------------------------ device tree
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t count) {
        release();
        if (count) HIPCHK(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void upload(const T* h, size_t count, hipStream_t s) {
        HIPCHK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T* h, size_t count, hipStream_t s) const {
        HIPCHK(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
    ~DBuf() { release(); }
};

class DeviceTree {
public:
    int nb = 0;
    double gamma = 5.0 / 3.0;
    int limiter = LIM_VANLEER;
    hipStream_t stream = nullptr;
    GhostPlan plan;
    DBuf<double> U, W, Bf, Us, F, gct, gcc, gfl, gfc, gchr, dxyz, org, gth;
    DBuf<double> diagpart;          // per-workgroup diagnostic partials, then the 8 sums
    DBuf<int> ghtab, leaves, waveblk, lvl, pushto;
    DBuf<unsigned char> ghsame;
    DBuf<int> pmask;                // [nb] sides read by copies without outcome reuse
    // Reuse of a copied ghost's con2prim outcome where the metric bits agree
    // (off: every copied ghost runs its own con2prim; a testing switch)
    bool exact_copies = true;
    std::vector<int> wave_off;  // offsets into waveblk
    DBuf<short> cell_order;
    DBuf<unsigned long long> scal;  // [0] dt bits, [1] faults
    DBuf<double> dts;               // per-step dt of a multi-step run
    DBuf<unsigned char> cst;        // [nb][8000] con2prim-current flags
    DBuf<unsigned char> flc;        // [nb][8000] con2prim outcome of each cell's floor state
    DBuf<unsigned long long> dtpart;  // per-workgroup dt candidates
    DBuf<double> lg;                // [nb][2][8000] log(rho), log(eps)
    DBuf<unsigned int> qn;          // Brent sub-queue lengths
    unsigned int qcap = 0;          // capacity of each sub-queue
    DBuf<int> qb, qc;
    DBuf<double> qx, qs;
    // cells queued for the Brent pass, per pass
    enum { QS_PRE = 0, QS_GHOST = 1, QS_C2P = 2, QS_FLOOR = 3, QS_N = 4 };
    DBuf<unsigned long long> qstat;

    // Per-kernel timing with HIP events on this tree's stream.
    bool profile = false;
    std::string profile_only;  // time only the kernels of this name ("" = all)
    // The step as a HIP graph: captured on first use for the current block
    // layout and CFL number, replayed per step (no host work between
    // kernels).  The step issues kernels only (queue counters are cleared
    // by k_qreset, not by memset nodes).
    bool use_graph = true;
    int layout_gen = 0;
    hipGraphExec_t step_exec = nullptr;
    int step_gen = -1;
    double step_cfl = 0.0;
    struct Rec { const char* name; hipEvent_t a, b; };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    struct Acc { std::string name; double ms = 0; long n = 0; };
    std::vector<Acc> acc;

    hipEvent_t ev() {
        if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        return e;
    }
    template <class Kern, class... Args>
    void launch(const char* name, Kern k, dim3 g, Args... args) {
        if (profile && (profile_only.empty() || profile_only == name)) {
            Rec r{name, ev(), ev()};
            HIPCHK(hipEventRecord(r.a, stream));
            hipLaunchKernelGGL(k, g, dim3(TPB), 0, stream, args...);
            HIPCHK(hipEventRecord(r.b, stream));
            recs.push_back(r);
        } else {
            hipLaunchKernelGGL(k, g, dim3(TPB), 0, stream, args...);
        }
    }
    void collect() {
        if (recs.empty()) return;
        HIPCHK(hipStreamSynchronize(stream));
        for (auto& r : recs) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
            auto it = std::find_if(acc.begin(), acc.end(), [&](const Acc& x) { return x.name == r.name; });
            if (it == acc.end()) { acc.push_back({r.name}); it = acc.end() - 1; }
            it->ms += ms;
            it->n += 1;
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        recs.clear();
    }

    DeviceTree() { HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking)); }
    void graph_release() {
        if (step_exec) (void)hipGraphExecDestroy(step_exec);
        step_exec = nullptr;
        step_gen = -1;
    }
    ~DeviceTree() {
        graph_release();
        for (auto& r : recs) { pool.push_back(r.a); pool.push_back(r.b); }
        for (auto e : pool) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }

    KP kp() {
        KP P;
        P.gamma = gamma; P.limiter = limiter;
        P.U = U.p; P.W = W.p; P.Bf = Bf.p; P.Us = Us.p; P.F = F.p;
        P.gct = gct.p; P.gcc = gcc.p; P.gfl = gfl.p; P.gfc = gfc.p; P.gchr = gchr.p; P.dxyz = dxyz.p; P.org = org.p;
        P.gth = gth.p;
        P.ghtab = ghtab.p; P.ghsame = ghsame.p; P.pmask = pmask.p; P.push_to = pushto.p;
        P.cell_order = cell_order.p;
        P.dtbits = scal.p; P.faults = scal.p + 1;
        P.cst = cst.p;
        P.flc = flc.p;
        P.lg = lg.p;
        P.qn = qn.p; P.qcap = qcap; P.qb = qb.p; P.qc = qc.p; P.qx = qx.p; P.qs = qs.p;
        P.qstat = qstat.p;
        return P;
    }

    // Per-block scratch arrays (sized by nb), cleared.
    void alloc_scratch() {
        Us.alloc((size_t)nb * 5 * NINT);
        F.alloc((size_t)nb * 3 * 8 * NFC);
        HIPCHK(hipMemsetAsync(F.p, 0, F.n * sizeof(double), stream));
        HIPCHK(hipMemsetAsync(Us.p, 0, Us.n * sizeof(double), stream));
        cst.alloc((size_t)nb * NCELL);
        HIPCHK(hipMemsetAsync(cst.p, 0, cst.n, stream));
        lg.alloc((size_t)nb * 2 * NCELL);
    }
    // Everything derived from the block structure: ghost plan, leaf list,
    // ordering waves, queues and reduction scratch.  The exact-copy flags
    // come from the host geometry when given, else from the device tables.
    void tables(const HostTree& t, const GeomHost* g) {
        build_ghost_plan(t, plan);
        ghtab.alloc(plan.tab.size()); ghtab.upload(plan.tab.data(), plan.tab.size(), stream);
        leaves.alloc(std::max<size_t>(1, plan.leaves.size()));
        if (!plan.leaves.empty()) leaves.upload(plan.leaves.data(), plan.leaves.size(), stream);
        ghsame.alloc(plan.same.size());
        if (!exact_copies) {
            HIPCHK(hipMemsetAsync(ghsame.p, 0, ghsame.n, stream));
        } else if (g) {
            mark_exact_copies(*g, plan);
            ghsame.upload(plan.same.data(), plan.same.size(), stream);
        } else {
            HIPCHK(hipMemsetAsync(ghsame.p, 0, ghsame.n, stream));
            if (!plan.leaves.empty())
                hipLaunchKernelGGL(k_mark_same, dim3(plan.leaves.size() * 6), dim3(TPB), 0, stream, kp(),
                                   (const int*)leaves.p, ghsame.p);
            HIPCHK(hipGetLastError());
        }
        pmask.alloc(std::max<size_t>(1, (size_t)nb));
        HIPCHK(hipMemsetAsync(pmask.p, 0, pmask.n * sizeof(int), stream));
        if (!plan.leaves.empty()) {
            const int n = (int)plan.leaves.size() * 6;
            hipLaunchKernelGGL(k_pmask, dim3((n + TPB - 1) / TPB), dim3(TPB), 0, stream, kp(), (const int*)leaves.p,
                               (int)plan.leaves.size(), pmask.p);
            HIPCHK(hipGetLastError());
        }
        {
            // z-side copy pushes: destination leaves that no prolongation reads
            std::vector<char> isleaf(nb, 0), psrc(nb, 0);
            for (int L : plan.leaves) isleaf[L] = 1;
            for (int L : plan.leaves)
                for (int sd = 0; sd < 6; sd++)
                    if (plan.tab[(size_t)L * 12 + sd * 2] == GH_PROLONG) psrc[plan.tab[(size_t)L * 12 + sd * 2 + 1]] = 1;
            std::vector<int> pt((size_t)nb * 2, -1);
            for (int B : plan.leaves)
                for (int s = 0; s < 2; s++) {
                    const int kind = plan.tab[(size_t)B * 12 + (4 + s) * 2], N = plan.tab[(size_t)B * 12 + (4 + s) * 2 + 1];
                    if (kind == GH_COPY && N >= 0 && N != B && isleaf[N] && !psrc[B]) pt[(size_t)N * 2 + (s ? 0 : 1)] = B;
                }
            pushto.alloc(std::max<size_t>(2, pt.size()));
            if (!pt.empty()) pushto.upload(pt.data(), pt.size(), stream);
        }
        std::vector<int> wb;
        wave_off.assign(1, 0);
        for (auto& w : plan.waves) {
            wb.insert(wb.end(), w.begin(), w.end());
            wave_off.push_back((int)wb.size());
        }
        waveblk.alloc(std::max<size_t>(1, wb.size()));
        if (!wb.empty()) waveblk.upload(wb.data(), wb.size(), stream);
        std::vector<int> lv(nb);
        for (int b = 0; b < nb; b++) lv[b] = t.quilt[b]->lvl;
        lvl.alloc(std::max<size_t>(1, lv.size()));
        if (!lv.empty()) lvl.upload(lv.data(), lv.size(), stream);
        diagpart.alloc(std::max<size_t>(1, plan.leaves.size() * WG_INT) * NDIAG + NDIAG);
        dtpart.alloc(std::max<size_t>(1, plan.leaves.size() * WG_INT));
        {
            // the largest pushing grid, in workgroups: c2p over leaves, floors over all blocks
            const size_t wg = std::max<size_t>({(size_t)1, plan.leaves.size() * WG_CELL, (size_t)nb * WG_INT});
            qcap = (unsigned int)(((wg + NSUBQ - 1) / NSUBQ) * TPB);
            const size_t total = (size_t)NSUBQ * qcap;
            qn.alloc((size_t)NSUBQ * QN_STRIDE);
            HIPCHK(hipMemsetAsync(qn.p, 0, qn.n * sizeof(unsigned int), stream));
            qb.alloc(total); qc.alloc(total); qx.alloc(total); qs.alloc(total * QS_WORDS);
        }
        // floor-state con2prim outcomes (geometry and floors are in place)
        flc.alloc((size_t)std::max(nb, 1) * NCELL);
        if (nb > 0) {
            hipLaunchKernelGGL(k_floor_c2p, dim3(nb * WG_CELL), dim3(TPB), 0, stream, kp(), flc.p);
            HIPCHK(hipGetLastError());
        }
    }

    // (Re)build for the host tree's current block structure.
    void build(const HostTree& t) {
        layout_gen++;
        gamma = t.prm.gamma;
        nb = t.nblocks();
        GeomHost g;
        build_geometry(t, g);
        U.alloc((size_t)nb * 8 * NCELL);
        W.alloc((size_t)nb * 8 * NCELL);
        Bf.alloc((size_t)nb * 3 * NFACE);
        alloc_scratch();
        gct.alloc(g.cell_tree.size()); gct.upload(g.cell_tree.data(), g.cell_tree.size(), stream);
        gcc.alloc(g.cell_cell.size()); gcc.upload(g.cell_cell.data(), g.cell_cell.size(), stream);
        gfl.alloc(g.floors.size()); gfl.upload(g.floors.data(), g.floors.size(), stream);
        gfc.alloc(g.face.size()); gfc.upload(g.face.data(), g.face.size(), stream);
        gchr.alloc(g.chr.size()); gchr.upload(g.chr.data(), g.chr.size(), stream);
        dxyz.alloc(g.dxyz.size()); dxyz.upload(g.dxyz.data(), g.dxyz.size(), stream);
        gth.alloc(g.th.size()); gth.upload(g.th.data(), g.th.size(), stream);
        {
            std::vector<double> o((size_t)nb * 3);
            for (int b = 0; b < nb; b++) {
                const Patch& p = *t.quilt[b];
                o[3 * b] = p.xe[0]; o[3 * b + 1] = p.ye[0]; o[3 * b + 2] = p.ze[0];
            }
            org.alloc(o.size()); org.upload(o.data(), o.size(), stream);
        }
        tables(t, &g);
        // ghost cells in storage order, then the border and the deep active zones
        std::vector<short> order;
        for (int c = 0; c < NCELL; c++) {
            int i = c / 400 - GH, j = (c / 20) % 20 - GH, k = c % 20 - GH;
            if (i < 0 || i >= BLK || j < 0 || j >= BLK || k < 0 || k >= BLK) order.push_back((short)c);
        }
        if ((int)order.size() != NGHOST) throw DevError("ghost cell table size");
        for (int pass = 0; pass < 2; pass++)
            for (int a = 0; a < NINT; a++) {
                const int i = a / 256, j = (a / 16) % 16, k = a % 16;
                if (deep_zone(i, j, k) == (pass == 1)) order.push_back((short)lindx(i, j, k));
            }
        if ((int)order.size() != NCELL) throw DevError("cell order table size");
        cell_order.alloc(order.size()); cell_order.upload(order.data(), order.size(), stream);
        scal.alloc(2);
        HIPCHK(hipMemsetAsync(scal.p, 0, 2 * sizeof(unsigned long long), stream));
        qstat.alloc(QS_N);
        HIPCHK(hipMemsetAsync(qstat.p, 0, QS_N * sizeof(unsigned long long), stream));
        HIPCHK(hipStreamSynchronize(stream));
    }

    // Re-lay a per-block array for the new block order: new slot b takes old
    // slot map_old[b], or (for a block that did not exist) slot map_new[b]
    // of `fresh` (packed host data of the new blocks), else zeros.
    void relayout(DBuf<double>& X, size_t words, int nb_new, const int* map_old_d, const int* map_new_d,
                  const std::vector<double>* fresh) {
        DBuf<double> Y;
        Y.alloc((size_t)nb_new * words);
        HIPCHK(hipMemsetAsync(Y.p, 0, Y.n * sizeof(double), stream));
        const int chunks = (int)std::min<size_t>(64, (words + TPB - 1) / TPB);
        if (nb_new > 0) {
            hipLaunchKernelGGL(k_gather, dim3(chunks, nb_new), dim3(TPB), 0, stream, Y.p, (const double*)X.p,
                               map_old_d, (long)words);
            if (fresh && !fresh->empty()) {
                DBuf<double> tmp;
                tmp.alloc(fresh->size());
                tmp.upload(fresh->data(), fresh->size(), stream);
                hipLaunchKernelGGL(k_gather, dim3(chunks, nb_new), dim3(TPB), 0, stream, Y.p, (const double*)tmp.p,
                                   map_new_d, (long)words);
                HIPCHK(hipStreamSynchronize(stream));
            }
        }
        HIPCHK(hipGetLastError());
        std::swap(X.p, Y.p);
        std::swap(X.n, Y.n);
    }

    // amrtree::regrid with the state on the device: flags from the device
    // state, the tree changes on the host, the data work on the device.
    // Returns the number of blocks refined plus flattened.
    int regrid(HostTree& t) {
        layout_gen++;
        const int nl = (int)plan.leaves.size();
        for (int b = 0; b < nb; b++) t.quilt[b]->dev = b;
        std::vector<unsigned char> lf(nl, 0);
        if (nl) {
            DBuf<unsigned int> fl;
            fl.alloc(nl);
            HIPCHK(hipMemsetAsync(fl.p, 0, fl.n * sizeof(unsigned int), stream));
            hipLaunchKernelGGL(k_regrid_flags, dim3(nl * WG_INT), dim3(TPB), 0, stream, kp(), (const int*)leaves.p,
                               (const int*)lvl.p, fl.p);
            HIPCHK(hipGetLastError());
            std::vector<unsigned int> f(nl);
            HIPCHK(hipMemcpyAsync(f.data(), fl.p, nl * sizeof(unsigned int), hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
            for (int l = 0; l < nl; l++) {
                const int lv = t.quilt[plan.leaves[l]]->lvl;
                const bool refine = (f[l] & 1u) && lv < 4;
                const bool erase = !(f[l] & 2u) && lv > 0;
                lf[l] = (refine ? 1 : 0) | (erase ? 2 : 0);
            }
        }
        RegridLog log;
        t.regrid_structure(lf, log);
        if (log.refined.empty() && log.flattened.empty()) return 0;
        // new layout
        const int nb_new = t.nblocks();
        std::vector<int> map_old(nb_new), map_new(nb_new, -1), fresh_blocks;
        std::unordered_map<const Patch*, int> slot;
        for (int b = 0; b < nb_new; b++) {
            const Patch* p = t.quilt[b].get();
            slot[p] = b;
            map_old[b] = p->dev;
            if (p->dev < 0) { map_new[b] = (int)fresh_blocks.size(); fresh_blocks.push_back(b); }
        }
        GeomHost g;
        build_geometry_blocks(t, fresh_blocks, g);
        std::vector<double> o((size_t)fresh_blocks.size() * 3);
        for (size_t n = 0; n < fresh_blocks.size(); n++) {
            const Patch& p = *t.quilt[fresh_blocks[n]];
            o[3 * n] = p.xe[0]; o[3 * n + 1] = p.ye[0]; o[3 * n + 2] = p.ze[0];
        }
        DBuf<int> mo, mn;
        mo.alloc(nb_new); mo.upload(map_old.data(), nb_new, stream);
        mn.alloc(nb_new); mn.upload(map_new.data(), nb_new, stream);
        // the old state stays the source of the data work
        DBuf<double> Uold, dxyz_old;
        Uold.alloc(U.n);
        HIPCHK(hipMemcpyAsync(Uold.p, U.p, U.n * sizeof(double), hipMemcpyDeviceToDevice, stream));
        dxyz_old.alloc(dxyz.n);
        HIPCHK(hipMemcpyAsync(dxyz_old.p, dxyz.p, dxyz.n * sizeof(double), hipMemcpyDeviceToDevice, stream));
        relayout(U, (size_t)8 * NCELL, nb_new, mo.p, mn.p, nullptr);
        relayout(W, (size_t)8 * NCELL, nb_new, mo.p, mn.p, nullptr);
        relayout(Bf, (size_t)3 * NFACE, nb_new, mo.p, mn.p, nullptr);
        relayout(gct, (size_t)NGCT * NCELL, nb_new, mo.p, mn.p, &g.cell_tree);
        relayout(gcc, (size_t)NGCC * NCELL, nb_new, mo.p, mn.p, &g.cell_cell);
        relayout(gfl, (size_t)2 * NCELL, nb_new, mo.p, mn.p, &g.floors);
        relayout(gfc, (size_t)3 * NGFC * NFC, nb_new, mo.p, mn.p, &g.face);
        relayout(gchr, (size_t)NCHR * NINT, nb_new, mo.p, mn.p, &g.chr);
        relayout(dxyz, 3, nb_new, mo.p, mn.p, &g.dxyz);
        relayout(gth, (size_t)NCELL, nb_new, mo.p, mn.p, &g.th);
        relayout(org, 3, nb_new, mo.p, mn.p, &o);
        nb = nb_new;
        alloc_scratch();
        KP P = kp();
        std::vector<int> binit;
        if (!log.refined.empty()) {
            std::vector<int> par, ch;
            for (Patch* p : log.refined) {
                par.push_back(p->dev);
                for (int h = 0; h < 8; h++) {
                    ch.push_back(slot.at(p->children[h]));
                    binit.push_back(ch.back());
                }
            }
            DBuf<int> dp, dc;
            dp.alloc(par.size()); dp.upload(par.data(), par.size(), stream);
            dc.alloc(ch.size()); dc.upload(ch.data(), ch.size(), stream);
            hipLaunchKernelGGL(k_refine_zones, dim3(NINT / TPB, (unsigned)(8 * par.size())), dim3(TPB), 0, stream, P,
                               (const double*)Uold.p, (const int*)dp.p, (const int*)dc.p);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(stream));
        }
        if (!log.flattened.empty()) {
            std::vector<int> par, ch;
            for (auto& fp : log.flattened) {
                par.push_back(slot.at(fp.first));
                binit.push_back(par.back());
                for (int h = 0; h < 8; h++) ch.push_back(fp.second[h]);
            }
            DBuf<int> dp, dc;
            dp.alloc(par.size()); dp.upload(par.data(), par.size(), stream);
            dc.alloc(ch.size()); dc.upload(ch.data(), ch.size(), stream);
            hipLaunchKernelGGL(k_restrict_zones, dim3(NINT / TPB, (unsigned)par.size()), dim3(TPB), 0, stream, P,
                               (const double*)Uold.p, (const double*)dxyz_old.p, (const int*)dp.p, (const int*)dc.p);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(stream));
        }
        if (!binit.empty()) {
            DBuf<int> db;
            db.alloc(binit.size()); db.upload(binit.data(), binit.size(), stream);
            hipLaunchKernelGGL(k_binit, dim3((3 * NFACE + TPB - 1) / TPB, (unsigned)binit.size()), dim3(TPB), 0, stream,
                               P, (const int*)db.p);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(stream));
        }
        tables(t, nullptr);
        HIPCHK(hipStreamSynchronize(stream));
        return (int)(log.refined.size() + log.flattened.size());
    }

    void upload_state(const HostTree& t) {
        std::vector<double> u((size_t)nb * 8 * NCELL), w((size_t)nb * 8 * NCELL), bf((size_t)nb * 3 * NFACE);
        t.export_state(u.data(), w.data(), bf.data());
        U.upload(u.data(), u.size(), stream);
        W.upload(w.data(), w.size(), stream);
        Bf.upload(bf.data(), bf.size(), stream);
        HIPCHK(hipMemsetAsync(cst.p, 0, cst.n, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }
    void download_state(HostTree& t) {
        std::vector<double> u((size_t)nb * 8 * NCELL), w((size_t)nb * 8 * NCELL), bf((size_t)nb * 3 * NFACE);
        U.download(u.data(), u.size(), stream);
        W.download(w.data(), w.size(), stream);
        Bf.download(bf.data(), bf.size(), stream);
        HIPCHK(hipStreamSynchronize(stream));
        t.import_state(u.data(), w.data(), bf.data());
    }

    void launch_dtcomp(double cfl) {
        KP P = kp();
        int nl = (int)plan.leaves.size();
        if (!nl) {
            launch("k_dtinit", k_dtinit, dim3(1), P);
            return;
        }
        launch("k_dtcomp", k_dtcomp, dim3(nl * WG_INT), P, (const int*)leaves.p, cfl, dtpart.p);
        launch("k_dtreduce", k_dtreduce, dim3(1), P, (const unsigned long long*)dtpart.p, nl * WG_INT);
    }
    void set_dt(double dt) {
        unsigned long long bits;
        std::memcpy(&bits, &dt, sizeof bits);
        HIPCHK(hipMemcpyAsync(scal.p, &bits, sizeof bits, hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }
    void launch_rkstg(int stage) {
        KP P = kp();
        int nl = (int)plan.leaves.size();
        if (nl) {
            launch("k_qreset", k_qreset, dim3(1), P);
            launch("k_c2p_pre", k_c2p_pre, dim3(nl * WG_INT), P, (const int*)leaves.p);
            launch("k_brent", k_brent, dim3(BRENT_WG), P, (int)QS_PRE);
        }
        for (size_t w = 0; w + 1 < wave_off.size(); w++) {
            int n = wave_off[w + 1] - wave_off[w];
            if (!n) continue;
            launch("k_qreset", k_qreset, dim3(1), P);
            launch("k_ghost", k_ghost, dim3(n * WG_GHOST), P, (const int*)(waveblk.p + wave_off[w]));
            launch("k_brent", k_brent, dim3(BRENT_WG), P, (int)QS_GHOST);
        }
        if (nl) {
            launch("k_qreset", k_qreset, dim3(1), P);
            launch("k_c2p", k_c2p, dim3(nl * WG_C2P), P, (const int*)leaves.p);
            launch("k_brent", k_brent, dim3(BRENT_WG), P, (int)QS_C2P);
        }
        launch("k_logs", k_logs, dim3(nb * WG_CELL), P);
        launch("k_flux", k_flux, dim3(nb * WG_FACE), P);
        launch("k_bfupd", k_bfupd, dim3(nb * WG_BF), P);
        launch("k_f2cb", k_f2cb, dim3(nb * WG_EDGE), P);
        launch("k_update", k_update, dim3(nb * WG_INT), P, stage);
    }
    void launch_floors() {
        KP P = kp();
        launch("k_qreset", k_qreset, dim3(1), P);
        launch("k_floors", k_floors, dim3(nb * WG_INT), P);
        launch("k_brent", k_brent, dim3(BRENT_WG), P, (int)QS_FLOOR);
    }
    void launch_step(double cfl) {
        launch_dtcomp(cfl);
        launch_rkstg(0);
        launch_rkstg(1);
        launch_floors();
    }
    // one step, from the captured graph where allowed (not while kernels are
    // timed one by one)
    void step(double cfl) {
        if (profile || !use_graph) {
            launch_step(cfl);
            return;
        }
        if (!step_exec || step_gen != layout_gen || step_cfl != cfl) {
            graph_release();
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            launch_step(cfl);
            HIPCHK(hipStreamEndCapture(stream, &g));
            HIPCHK(hipGraphInstantiate(&step_exec, g, nullptr, nullptr, 0));
            HIPCHK(hipGraphDestroy(g));
            step_gen = layout_gen;
            step_cfl = cfl;
        }
        HIPCHK(hipGraphLaunch(step_exec, stream));
    }
    void run_steps(double cfl, int nsteps) {
        if ((int)dts.n < nsteps) dts.alloc(nsteps);
        for (int s = 0; s < nsteps; s++) {
            step(cfl);
            HIPCHK(hipMemcpyAsync(dts.p + s, scal.p, sizeof(double), hipMemcpyDeviceToDevice, stream));
        }
    }
    // mass, energy, L_pol, Edot, Phi_B, thermE, Hscal (see k_diag)
    void diagnostics(double M, double a, double Q, double out[7]) {
        const double hrz = M + std::sqrt(M * M - a * a - Q * Q);
        const double rs = 10.0 * M;
        KP P = kp();
        const int nl = (int)plan.leaves.size();
        double s[NDIAG] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (nl) {
            double* fin = diagpart.p + (size_t)nl * WG_INT * NDIAG;
            launch("k_diag", k_diag, dim3(nl * WG_INT), P, (const int*)leaves.p, hrz, rs, diagpart.p);
            launch("k_diag_reduce", k_diag_reduce, dim3(1), (const double*)diagpart.p, nl * WG_INT, fin);
            HIPCHK(hipMemcpyAsync(s, fin, sizeof s, hipMemcpyDeviceToHost, stream));
            HIPCHK(hipStreamSynchronize(stream));
        }
        out[0] = s[0]; out[1] = s[1]; out[2] = s[2]; out[3] = s[3];
        out[4] = 0.5 * s[4];
        out[5] = s[5];
        out[6] = (s[7] > 1e-20) ? (s[6] / s[7]) / rs : 0.0;
    }
    // constrans::maxdivB of every block (per_block[nb]) and its maximum over
    // the leaves, reduced on the device
    void maxdivb(double* per_block, double* max_leaf) {
        std::vector<int> all(nb);
        for (int b = 0; b < nb; b++) all[b] = b;
        DBuf<int> lst;
        lst.alloc(std::max(1, nb));
        lst.upload(all.data(), all.size(), stream);
        DBuf<double> res;
        res.alloc(std::max(1, nb));
        KP P = kp();
        if (nb) hipLaunchKernelGGL(k_maxdivb, dim3(nb), dim3(TPB), 0, stream, P, (const int*)lst.p, res.p);
        HIPCHK(hipGetLastError());
        std::vector<double> h(nb);
        if (nb) HIPCHK(hipMemcpyAsync(h.data(), res.p, nb * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        double m = 0.0;
        for (int L : plan.leaves) m = smax(m, h[L]);
        if (per_block) std::copy(h.begin(), h.end(), per_block);
        if (max_leaf) *max_leaf = m;
    }
    // packed W of leaves [l0, l0 + n) into host memory h ([n][8][4096])
    void pack_w(int l0, int n, double* dbuf, double* h) {
        if (n <= 0) return;
        KP P = kp();
        hipLaunchKernelGGL(k_pack_w, dim3(n * WG_INT), dim3(TPB), 0, stream, P, (const int*)leaves.p, l0, dbuf);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(h, dbuf, (size_t)n * 8 * NINT * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }
    double read_dt() {
        unsigned long long bits = 0;
        HIPCHK(hipMemcpyAsync(&bits, scal.p, sizeof bits, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        double dt;
        std::memcpy(&dt, &bits, sizeof dt);
        return dt;
    }
    // The initial data of main.cpp:94-159 written on the device over every
    // block: floor state, fm_init, global rho_max, B_pot_init, ptoc of the
    // active zones.  Returns 0, 1 where fm_init's L0_comp has a vanishing
    // denominator (nothing written), 2 where fm_init places no torus (the
    // rest is done), 3 where the global rho_max is below 1e-14 (main.cpp
    // stops there; floor state and torus written).
    int init_torus(const TreeParams& prm, double& rho_max, std::string& msg) {
        FmScalars s;
        int fm = fm_scalars(prm.M, prm.a, prm.Q, prm.gamma, s);
        if (fm == 1) {
            msg = "LNRF angular momentum: denominator is zero (r_max inside the ergosphere)";
            return 1;
        }
        KP P = kp();
        HIPCHK(hipMemsetAsync(Bf.p, 0, Bf.n * sizeof(double), stream));
        // every cell gets a new state: no con2prim outcome is known for it
        HIPCHK(hipMemsetAsync(cst.p, 0, cst.n, stream));
        launch("k_ic_atmosphere", k_ic_atmosphere, dim3(nb * WG_CELL), P);
        DBuf<unsigned long long> red;
        red.alloc(2);
        HIPCHK(hipMemsetAsync(red.p, 0, 2 * sizeof(unsigned long long), stream));
        unsigned long long bits[2] = {0ull, 0ull};
        FmPar f{prm.M, prm.a, prm.Q, prm.gamma, s.K, s.l0, s.W_in, s.r_H, 0.0};
        if (fm == 0) {
            launch("k_fm_rho_max", k_fm_rho_max, dim3(nb * WG_INT), P, f, red.p);
            red.download(bits, 1, stream);
            HIPCHK(hipStreamSynchronize(stream));
            double fm_max;
            std::memcpy(&fm_max, &bits[0], sizeof fm_max);
            if (fm_max < 1e-14) {
                fm = 2;
                msg = "no torus cells found";
            } else {
                f.rho_scal = 1.0 / fm_max;
                launch("k_fm_write", k_fm_write, dim3(nb * WG_INT), P, f);
            }
        } else {
            msg = "ln(h) <= 0 at r = r_max: the torus cannot be initialised";
        }
        launch("k_ic_rho_max", k_ic_rho_max, dim3(nb * WG_INT), P, red.p + 1);
        red.download(bits, 2, stream);
        HIPCHK(hipStreamSynchronize(stream));
        std::memcpy(&rho_max, &bits[1], sizeof rho_max);
        if (rho_max < 1e-14) {
            msg = "rho_max is zero after the torus initialisation";
            return 3;
        }
        // B_pot_init returns at once for rho_max <= 1e-10 (initconds.cpp:597-599)
        if (rho_max > 1e-10) {
            DBuf<double> pmax;
            pmax.alloc(nb);
            hipLaunchKernelGGL(k_bpot_max, dim3(nb), dim3(TPB), 0, stream, P, rho_max, pmax.p);
            launch("k_bpot_faces", k_bpot_faces, dim3(nb * WG_EDGE), P, (const double*)pmax.p);
            launch("k_bpot_f2cb", k_bpot_f2cb, dim3(nb * WG_EDGE), P);
            HIPCHK(hipStreamSynchronize(stream));
        }
        launch("k_ptoc_interior", k_ptoc_interior, dim3(nb * WG_INT), P);
        sync();
        return fm;
    }
    long long take_faults() {
        unsigned long long f = 0;
        HIPCHK(hipMemcpyAsync(&f, scal.p + 1, sizeof f, hipMemcpyDeviceToHost, stream));
        unsigned long long zero = 0;
        HIPCHK(hipMemcpyAsync(scal.p + 1, &zero, sizeof zero, hipMemcpyHostToDevice, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return (long long)f;
    }
    void sync() {
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream));
    }
};

}  // namespace bh

// =================================================================== C ABI
using namespace bh;

struct bh_tree {
    std::unique_ptr<HostTree> host;
    std::unique_ptr<DeviceTree> dev;
    bh_params params{BH_LIM_VANLEER, {0, 0, 0, 0, 0, 0, 0}};
    bool host_valid = true;   // host copy holds the current state
    bool dev_valid = false;   // device copy holds the current state
    bool layout_dirty = true; // block structure changed since the device build
    long long faults = 0;
    std::string err;
};

namespace {
thread_local std::string g_create_err;

int fail(bh_tree* t, int code, const std::string& msg) {
    if (t) t->err = msg;
    else g_create_err = msg;
    return code;
}
void need_host(bh_tree* t) {
    if (!t->host_valid) {
        t->dev->download_state(*t->host);
        t->host_valid = true;
    }
}
void host_changed(bh_tree* t) {
    t->host_valid = true;
    t->dev_valid = false;
}
void need_device(bh_tree* t) {
    if (!t->dev) t->dev = std::make_unique<DeviceTree>();
    if (t->dev->limiter != t->params.limiter) {
        t->dev->limiter = t->params.limiter;
        t->dev->graph_release();  // the captured step holds the old kernel arguments
    }
    if (t->layout_dirty) {
        need_host(t);
        t->dev->build(*t->host);
        t->layout_dirty = false;
        t->dev_valid = false;
    }
    if (!t->dev_valid) {
        t->dev->upload_state(*t->host);
        t->dev_valid = true;
    }
}
void device_changed(bh_tree* t) {
    t->dev_valid = true;
    t->host_valid = false;
}
int finish_faults(bh_tree* t, long long add) {
    t->faults += add;
    return add > 0 ? BH_EDOMAIN : BH_OK;
}

template <class F>
int guarded(bh_tree* t, F&& f) {
    if (!t) return fail(nullptr, BH_EINVAL, "null tree");
    try {
        return f();
    } catch (const DevError& e) {
        return fail(t, BH_EDEVICE, e.msg);
    } catch (const std::exception& e) {
        return fail(t, BH_ESTATE, e.what());
    }
}

std::vector<double> point_geometry(double M, double a, double Q, int n, const double* r, const double* th) {
    std::vector<double> g((size_t)10 * n);
    for (int i = 0; i < n; i++) {
        GPt m = gpt_from(metric_comp(M, a, Q, r[i], th[i]));
        double v[10] = {m.sqrtdetg, m.sqrtdetgam, m.alpha, m.beta0, m.g11, m.g13, m.g22, m.g33, m.gi13, m.gi33};
        for (int k = 0; k < 10; k++) g[(size_t)10 * i + k] = v[k];
    }
    return g;
}

struct Scratch {
    std::vector<void*> ptrs;
    template <class T>
    T* put(const T* h, size_t n) {
        T* p = nullptr;
        HIPCHK(hipMalloc(&p, std::max<size_t>(1, n) * sizeof(T)));
        ptrs.push_back(p);
        if (h && n) HIPCHK(hipMemcpy(p, h, n * sizeof(T), hipMemcpyHostToDevice));
        return p;
    }
    ~Scratch() {
        for (void* p : ptrs) (void)hipFree(p);
    }
};
}  // namespace

extern "C" {

int bh_tree_create(const double dom_l[3], const double dom_h[3], int nqlt, double M, double a, double Q, double gm,
                   bh_tree** out) {
    if (!out || !dom_l || !dom_h || nqlt < 1) return fail(nullptr, BH_EINVAL, "invalid arguments");
    try {
        auto* t = new bh_tree;
        t->host = std::make_unique<HostTree>(dom_l, dom_h, nqlt, M, a, Q, gm);
        t->faults = t->host->faults;
        *out = t;
        return BH_OK;
    } catch (const std::exception& e) {
        return fail(nullptr, BH_ESTATE, e.what());
    }
}
void bh_tree_destroy(bh_tree* t) { delete t; }
const char* bh_last_error(const bh_tree* t) { return t ? t->err.c_str() : g_create_err.c_str(); }

int bh_tree_fm_init(bh_tree* t) {
    return guarded(t, [&] {
        need_host(t);
        int rc = t->host->fm_init();
        host_changed(t);
        if (rc == 1) return fail(t, BH_ESTATE, t->host->last_msg);
        if (rc == 2) { t->err = t->host->last_msg; return 2; }
        return BH_OK;
    });
}
int bh_tree_rho_max(bh_tree* t, double* rho_max) {
    return guarded(t, [&] {
        need_host(t);
        *rho_max = t->host->rho_max();
        return BH_OK;
    });
}
int bh_tree_B_pot_init(bh_tree* t, double rho_max) {
    return guarded(t, [&] {
        need_host(t);
        for (auto& p : t->host->quilt) t->host->B_pot_init(*p, rho_max);
        host_changed(t);
        return BH_OK;
    });
}
int bh_tree_ptoc_interior(bh_tree* t) {
    return guarded(t, [&] {
        need_host(t);
        long long f0 = t->host->faults;
        t->host->ptoc_interior();
        host_changed(t);
        return finish_faults(t, t->host->faults - f0);
    });
}
int bh_tree_regrid(bh_tree* t) {
    return guarded(t, [&] {
        need_host(t);
        long long f0 = t->host->faults;
        t->host->regrid();
        host_changed(t);
        t->layout_dirty = true;
        return finish_faults(t, t->host->faults - f0);
    });
}
int bh_tree_set_params(bh_tree* t, const bh_params* p) {
    if (!t || !p) return fail(t, BH_EINVAL, "invalid arguments");
    if (p->limiter < BH_LIM_MINMOD || p->limiter > BH_LIM_MC) return fail(t, BH_EINVAL, "limiter must be 0, 1 or 2");
    for (int q = 0; q < 7; q++)
        if (p->reserved[q]) return fail(t, BH_EINVAL, "reserved parameter words must be zero");
    t->params = *p;
    return BH_OK;
}
int bh_tree_get_params(const bh_tree* t, bh_params* p) {
    if (!t || !p) return BH_EINVAL;
    *p = t->params;
    return BH_OK;
}
int bh_tree_set_blocks(bh_tree* t, int nb, const int* meta) {
    if (!t || nb <= 0 || !meta) return fail(t, BH_EINVAL, "invalid arguments");
    return guarded(t, [&] {
        need_host(t);
        t->host->set_blocks(nb, meta);
        host_changed(t);
        t->layout_dirty = true;
        return BH_OK;
    });
}
int bh_tree_regrid_device(bh_tree* t, int* changed) {
    return guarded(t, [&] {
        need_device(t);
        const int n = t->dev->regrid(*t->host);
        if (changed) *changed = n;
        // the device holds the state of the new layout; the host copy of
        // every block is now stale
        t->host_valid = false;
        t->dev_valid = true;
        t->layout_dirty = false;
        return finish_faults(t, t->dev->take_faults());
    });
}
int bh_tree_init_torus_device(bh_tree* t, double* rho_max) {
    return guarded(t, [&] {
        if (!t->dev) t->dev = std::make_unique<DeviceTree>();
        if (t->layout_dirty) {
            t->dev->build(*t->host);
            t->layout_dirty = false;
        }
        double rm = 0.0;
        std::string msg;
        const int rc = t->dev->init_torus(t->host->prm, rm, msg);
        if (rc == 1) return fail(t, BH_ESTATE, msg);
        // the device now holds the whole state
        device_changed(t);
        if (rho_max) *rho_max = rm;
        if (rc == 3) return fail(t, BH_ESTATE, msg);
        const long long nf = t->dev->take_faults();
        if (nf > 0) return finish_faults(t, nf);
        if (rc == 2) { t->err = msg; return 2; }
        return BH_OK;
    });
}
int bh_tree_nblocks(const bh_tree* t, int* nb) {
    if (!t || !nb) return BH_EINVAL;
    *nb = t->host->nblocks();
    return BH_OK;
}
int bh_tree_get_blocks(const bh_tree* t, int* meta, double* edges) {
    if (!t || !meta || !edges) return BH_EINVAL;
    t->host->export_blocks(meta, edges);
    return BH_OK;
}
int bh_tree_get_state(bh_tree* t, double* U, double* W, double* Bf) {
    return guarded(t, [&] {
        need_host(t);
        t->host->export_state(U, W, Bf);
        return BH_OK;
    });
}
int bh_tree_set_state(bh_tree* t, const double* U, const double* W, const double* Bf) {
    return guarded(t, [&] {
        need_host(t);
        t->host->import_state(U, W, Bf);
        host_changed(t);
        return BH_OK;
    });
}

int bh_rk2_dtcomp(bh_tree* t, double cfl, double* dt) {
    return guarded(t, [&] {
        need_device(t);
        t->dev->launch_dtcomp(cfl);
        t->dev->sync();
        *dt = t->dev->read_dt();
        return finish_faults(t, t->dev->take_faults());
    });
}
int bh_rk2_rkstg(bh_tree* t, double dt, int stage) {
    if (stage != 0 && stage != 1) return fail(t, BH_EINVAL, "stage must be 0 or 1");
    return guarded(t, [&] {
        need_device(t);
        t->dev->set_dt(dt);
        t->dev->launch_rkstg(stage);
        t->dev->sync();
        device_changed(t);
        return finish_faults(t, t->dev->take_faults());
    });
}
int bh_tree_floors(bh_tree* t) {
    return guarded(t, [&] {
        need_device(t);
        t->dev->launch_floors();
        t->dev->sync();
        device_changed(t);
        return finish_faults(t, t->dev->take_faults());
    });
}
int bh_rk2_step(bh_tree* t, double cfl, double* dt) {
    return guarded(t, [&] {
        need_device(t);
        t->dev->step(cfl);
        t->dev->sync();
        device_changed(t);
        if (dt) *dt = t->dev->read_dt();
        return finish_faults(t, t->dev->take_faults());
    });
}
int bh_rk2_advance(bh_tree* t, double cfl, int nsteps, double t_end, double* t_io, int* steps_out) {
    return guarded(t, [&] {
        need_device(t);
        double tt = t_io ? *t_io : 0.0;
        int n = 0;
        long long nf = 0;
        // the reference's time loop ends at the first equation-of-state
        // fault (std::domain_error out of step); here the faulting step
        // completes, is counted, and the loop stops after it
        while (tt < t_end && n < nsteps) {
            t->dev->step(cfl);
            t->dev->sync();
            tt += t->dev->read_dt();
            n++;
            nf += t->dev->take_faults();
            if (nf) break;
        }
        device_changed(t);
        if (t_io) *t_io = tt;
        if (steps_out) *steps_out = n;
        return finish_faults(t, nf);
    });
}
int bh_profile_only(bh_tree* t, const char* name) {
    return guarded(t, [&] {
        if (!t->dev) t->dev = std::make_unique<DeviceTree>();
        t->dev->profile_only = name ? name : "";
        return BH_OK;
    });
}
int bh_tree_set_exact_copies(bh_tree* t, int enable) {
    return guarded(t, [&] {
        if (!t->dev) t->dev = std::make_unique<DeviceTree>();
        if ((enable != 0) != t->dev->exact_copies) {
            if (t->dev_valid) need_host(t);
            t->dev->exact_copies = enable != 0;
            t->dev->graph_release();
            t->layout_dirty = true;
        }
        return BH_OK;
    });
}
int bh_tree_use_graph(bh_tree* t, int enable) {
    return guarded(t, [&] {
        if (!t->dev) t->dev = std::make_unique<DeviceTree>();
        t->dev->use_graph = enable != 0;
        if (!enable) t->dev->graph_release();
        return BH_OK;
    });
}
int bh_rk2_steps(bh_tree* t, double cfl, int nsteps, double* dt_out) {
    if (nsteps < 0) return fail(t, BH_EINVAL, "nsteps < 0");
    return guarded(t, [&] {
        need_device(t);
        t->dev->run_steps(cfl, nsteps);
        t->dev->sync();
        device_changed(t);
        if (dt_out && nsteps) t->dev->dts.download(dt_out, nsteps, t->dev->stream);
        t->dev->sync();
        return finish_faults(t, t->dev->take_faults());
    });
}
int bh_tree_prepare(bh_tree* t) {
    return guarded(t, [&] {
        need_device(t);
        t->dev->sync();
        return BH_OK;
    });
}
int bh_profile(bh_tree* t, int enable) {
    return guarded(t, [&] {
        if (!t->dev) t->dev = std::make_unique<DeviceTree>();
        t->dev->collect();
        t->dev->profile = enable != 0;
        if (enable) {
            t->dev->acc.clear();
            if (t->dev->qstat.p)
                HIPCHK(hipMemsetAsync(t->dev->qstat.p, 0, DeviceTree::QS_N * sizeof(unsigned long long),
                                      t->dev->stream));
        }
        return BH_OK;
    });
}
int bh_profile_report(bh_tree* t, char* buf, int len) {
    return guarded(t, [&] {
        std::string out;
        if (t->dev) {
            t->dev->collect();
            char line[160];
            for (auto& a : t->dev->acc) {
                std::snprintf(line, sizeof line, "%s %ld %.6f\n", a.name.c_str(), a.n, a.ms);
                out += line;
            }
            // cells queued for the Brent pass since profiling was enabled
            unsigned long long q[DeviceTree::QS_N] = {0, 0, 0, 0};
            if (t->dev->qstat.p) HIPCHK(hipMemcpy(q, t->dev->qstat.p, sizeof q, hipMemcpyDeviceToHost));
            const char* qn[DeviceTree::QS_N] = {"pre", "ghost", "c2p", "floor"};
            for (int i = 0; i < DeviceTree::QS_N; i++) {
                std::snprintf(line, sizeof line, "queue:%s %llu 0\n", qn[i], q[i]);
                out += line;
            }
        }
        if (buf && len > 0) std::snprintf(buf, len, "%s", out.c_str());
        return (int)out.size() < len ? BH_OK : BH_EINVAL;
    });
}
int bh_device_sync(void) { return hipDeviceSynchronize() == hipSuccess ? BH_OK : BH_EDEVICE; }
int bh_tree_maxdivB_device(bh_tree* t, double* per_block, double* max_leaf) {
    return guarded(t, [&] {
        need_device(t);
        t->dev->maxdivb(per_block, max_leaf);
        return BH_OK;
    });
}
int bh_tree_diagnostics(bh_tree* t, double* out) {
    if (!out) return fail(t, BH_EINVAL, "null output");
    return guarded(t, [&] {
        need_device(t);
        const TreeParams& m = t->host->prm;
        t->dev->diagnostics(m.M, m.a, m.Q, out);
        return BH_OK;
    });
}
// Dump file, version 1 (little-endian):
//   header, 128 bytes: "BHFLUIDW", u32 version, u32 header bytes, i32 leaf
//     count, i32 variables (8), i32 zones per block (4096), i32 0, f64 M, a,
//     Q, gamma, time, i64 step, zero padding
//   block table, 64 bytes per leaf: i32 level, i32 i, j, k, f64 lo[3], hi[3]
//   data: per leaf, per variable (rho, eps, v0, v1, v2, B0, B1, B2), the
//     4096 active zones, k fastest
// Leaves in quilt order.
int bh_tree_write_dump(bh_tree* t, const char* path, double time, long long step) {
    if (!path) return fail(t, BH_EINVAL, "null path");
    return guarded(t, [&] {
        std::vector<int> leaves;
        for (int b = 0; b < t->host->nblocks(); b++)
            if (t->host->quilt[b]->leaf) leaves.push_back(b);
        const int nl = (int)leaves.size();
        FILE* f = std::fopen(path, "wb");
        if (!f) return fail(t, BH_EINVAL, std::string("cannot open ") + path);
        std::unique_ptr<FILE, int (*)(FILE*)> guard(f, std::fclose);
        unsigned char hdr[128];
        std::memset(hdr, 0, sizeof hdr);
        std::memcpy(hdr, "BHFLUIDW", 8);
        const uint32_t ver = 1, hb = 128;
        const int32_t ints[4] = {nl, 8, NINT, 0};
        const TreeParams& m = t->host->prm;
        const double dbl[5] = {m.M, m.a, m.Q, m.gamma, time};
        const int64_t st = step;
        std::memcpy(hdr + 8, &ver, 4);
        std::memcpy(hdr + 12, &hb, 4);
        std::memcpy(hdr + 16, ints, 16);
        std::memcpy(hdr + 32, dbl, 40);
        std::memcpy(hdr + 72, &st, 8);
        bool ok = std::fwrite(hdr, 1, sizeof hdr, f) == sizeof hdr;
        for (int b : leaves) {
            const Patch& p = *t->host->quilt[b];
            unsigned char e[64];
            const int32_t li[4] = {p.lvl, p.ijk[0], p.ijk[1], p.ijk[2]};
            const double lohi[6] = {p.xe[0], p.ye[0], p.ze[0], p.xe[BLK], p.ye[BLK], p.ze[BLK]};
            std::memcpy(e, li, 16);
            std::memcpy(e + 16, lohi, 48);
            ok = ok && std::fwrite(e, 1, sizeof e, f) == sizeof e;
        }
        // data in chunks of up to 256 leaves, from whichever copy is current
        const int chunk = 256;
        std::vector<double> buf((size_t)std::min(nl, chunk) * 8 * NINT);
        const bool from_device = !t->host_valid;
        DBuf<double> dbuf;
        if (from_device) dbuf.alloc(buf.size());
        for (int l0 = 0; l0 < nl && ok; l0 += chunk) {
            const int n = std::min(chunk, nl - l0);
            if (from_device) {
                t->dev->pack_w(l0, n, dbuf.p, buf.data());
            } else {
                for (int l = 0; l < n; l++) {
                    Patch& p = *t->host->quilt[leaves[l0 + l]];
                    for (int v = 0; v < 8; v++)
                        for (int a = 0; a < NINT; a++)
                            buf[((size_t)l * 8 + v) * NINT + a] = p.w(v, lindx(a / 256, (a / 16) % 16, a % 16));
                }
            }
            const size_t cnt = (size_t)n * 8 * NINT;
            ok = std::fwrite(buf.data(), sizeof(double), cnt, f) == cnt;
        }
        if (!ok) return fail(t, BH_EINVAL, std::string("write failed: ") + path);
        return BH_OK;
    });
}
int bh_tree_take_faults(bh_tree* t, long long* n) {
    if (!t || !n) return BH_EINVAL;
    *n = t->faults;
    t->faults = 0;
    return BH_OK;
}
int bh_tree_maxdivB(bh_tree* t, int b, double* out) {
    return guarded(t, [&] {
        need_host(t);
        if (b < 0 || b >= t->host->nblocks()) return fail(t, BH_EINVAL, "block index out of range");
        Patch& p = *t->host->quilt[b];
        double dx = p.dx(), dy = p.dy(), dz = p.dz(), mx = 0.0;
        for (int i = -GH + 1; i < BLK + GH - 1; i++)
            for (int j = -GH + 1; j < BLK + GH - 1; j++)
                for (int k = -GH + 1; k < BLK + GH - 1; k++) {
                    double dv = (p.bfx()[ixf(i, j, k)] - p.bfx()[ixf(i - 1, j, k)]) / dx +
                                (p.bfy()[iyf(i, j, k)] - p.bfy()[iyf(i, j - 1, k)]) / dy +
                                (p.bfz()[izf(i, j, k)] - p.bfz()[izf(i, j, k - 1)]) / dz;
                    mx = smax(mx, fabs(dv));
                }
        *out = mx;
        return BH_OK;
    });
}

int bh_ptoc_batch(double M, double a, double Q, double gm, int n, const double* r, const double* th,
                  const double* W, double* U) {
    if (n < 0) return BH_EINVAL;
    try {
        std::vector<double> g = point_geometry(M, a, Q, n, r, th);
        Scratch s;
        double* dg = s.put(g.data(), g.size());
        double* dW = s.put(W, (size_t)8 * n);
        double* dU = s.put<double>(nullptr, (size_t)8 * n);
        unsigned long long zero = 0;
        unsigned long long* df = s.put(&zero, 1);
        if (n) hipLaunchKernelGGL(k_ptoc_batch, dim3((n + 255) / 256), dim3(256), 0, 0, n, gm, dg, dW, dU, df);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(U, dU, (size_t)8 * n * sizeof(double), hipMemcpyDeviceToHost));
        unsigned long long f = 0;
        HIPCHK(hipMemcpy(&f, df, sizeof f, hipMemcpyDeviceToHost));
        return f ? BH_EDOMAIN : BH_OK;
    } catch (const DevError& e) {
        return fail(nullptr, BH_EDEVICE, e.msg);
    }
}
int bh_ctop_batch(double M, double a, double Q, double gm, int n, const double* r, const double* th,
                  const double* U, double* W, int* ok) {
    if (n < 0) return BH_EINVAL;
    try {
        std::vector<double> g = point_geometry(M, a, Q, n, r, th);
        Scratch s;
        double* dg = s.put(g.data(), g.size());
        double* dU = s.put(U, (size_t)8 * n);
        double* dW = s.put<double>(nullptr, (size_t)8 * n);
        int* dok = s.put<int>(nullptr, n);
        unsigned long long zero = 0;
        unsigned long long* df = s.put(&zero, 1);
        if (n) hipLaunchKernelGGL(k_ctop_batch, dim3((n + 255) / 256), dim3(256), 0, 0, n, gm, dg, dU, dW, dok, df);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(W, dW, (size_t)8 * n * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(ok, dok, (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
        unsigned long long f = 0;
        HIPCHK(hipMemcpy(&f, df, sizeof f, hipMemcpyDeviceToHost));
        return f ? BH_EDOMAIN : BH_OK;
    } catch (const DevError& e) {
        return fail(nullptr, BH_EDEVICE, e.msg);
    }
}
int bh_hlld_batch(double M, double a, double Q, double gm, int n, const double* r, const double* th, int dim,
                  const double* WL, const double* WR, const double* UL, const double* UR, double* F,
                  double* ctspeed) {
    if (n < 0 || dim < 0 || dim > 2) return BH_EINVAL;
    try {
        std::vector<double> g = point_geometry(M, a, Q, n, r, th);
        Scratch s;
        double* dg = s.put(g.data(), g.size());
        double* dWL = s.put(WL, (size_t)8 * n);
        double* dWR = s.put(WR, (size_t)8 * n);
        double* dUL = s.put(UL, (size_t)8 * n);
        double* dUR = s.put(UR, (size_t)8 * n);
        double* dF = s.put<double>(nullptr, (size_t)8 * n);
        double* dc = s.put<double>(nullptr, n);
        unsigned long long zero = 0;
        unsigned long long* df = s.put(&zero, 1);
        if (n)
            hipLaunchKernelGGL(k_hlld_batch, dim3((n + 255) / 256), dim3(256), 0, 0, n, gm, dim, dg, dWL, dWR, dUL,
                               dUR, dF, dc, df);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(F, dF, (size_t)8 * n * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(ctspeed, dc, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
        unsigned long long f = 0;
        HIPCHK(hipMemcpy(&f, df, sizeof f, hipMemcpyDeviceToHost));
        return f ? BH_EDOMAIN : BH_OK;
    } catch (const DevError& e) {
        return fail(nullptr, BH_EDEVICE, e.msg);
    }
}
int bh_math_batch(int op, int n, const double* x, double* y, int on_device) {
    if (n < 0 || op < 0 || op > 4) return BH_EINVAL;
    if (!on_device) {
        for (int i = 0; i < n; i++) y[i] = math1(op, x[i]);
        return BH_OK;
    }
    try {
        Scratch s;
        double* dx = s.put(x, n);
        double* dy = s.put<double>(nullptr, n);
        if (n) hipLaunchKernelGGL(k_math_batch, dim3((n + 255) / 256), dim3(256), 0, 0, op, n, dx, dy);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(y, dy, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
        return BH_OK;
    } catch (const DevError& e) {
        return fail(nullptr, BH_EDEVICE, e.msg);
    }
}
int bh_math2_batch(int op, int n, const double* x, const double* y, double* out, int on_device) {
    if (n < 0 || op < 0 || op > 1) return BH_EINVAL;
    if (!on_device) {
        for (int i = 0; i < n; i++) out[i] = math2(op, x[i], y[i]);
        return BH_OK;
    }
    try {
        Scratch s;
        double* dx = s.put(x, n);
        double* dy = s.put(y, n);
        double* dout = s.put<double>(nullptr, n);
        if (n) hipLaunchKernelGGL(k_math2_batch, dim3((n + 255) / 256), dim3(256), 0, 0, op, n, dx, dy, dout);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(out, dout, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
        return BH_OK;
    } catch (const DevError& e) {
        return fail(nullptr, BH_EDEVICE, e.msg);
    }
}
int bh_c2p_sign_batch(double gm, int n, const double* in, int* sign, double* f, int on_device) {
    if (n < 0 || (n && (!in || !sign || !f))) return BH_EINVAL;
    if (!on_device) {
        for (int i = 0; i < n; i++) {
            const double* a = in + 6 * (size_t)i;
            sign[i] = c2p_f_sign(gm, a[0], a[1], a[2], a[3], a[4], a[5]);
            f[i] = c2p_f(gm, a[0], a[1], a[2], a[3], a[4], a[5]);
        }
        return BH_OK;
    }
    try {
        Scratch s;
        double* din = s.put(in, (size_t)6 * n);
        int* ds = s.put<int>(nullptr, n);
        double* df = s.put<double>(nullptr, n);
        if (n) hipLaunchKernelGGL(k_c2p_sign_batch, dim3((n + 255) / 256), dim3(256), 0, 0, n, gm, din, ds, df);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpy(sign, ds, (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(f, df, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
        return BH_OK;
    } catch (const DevError& e) {
        return fail(nullptr, BH_EDEVICE, e.msg);
    }
}
int bh_copy_probe(long long nbytes, int reps, double* ms_per_copy) {
    if (nbytes < 32 || reps < 1) return BH_EINVAL;
    try {
        size_t n = (size_t)nbytes / sizeof(double4);
        double4 *a = nullptr, *b = nullptr;
        HIPCHK(hipMalloc(&a, n * sizeof(double4)));
        HIPCHK(hipMalloc(&b, n * sizeof(double4)));
        HIPCHK(hipMemset(a, 0, n * sizeof(double4)));
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        hipLaunchKernelGGL(k_copy, dim3(4096), dim3(TPB), 0, 0, a, b, n);
        HIPCHK(hipEventRecord(e0, 0));
        for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy, dim3(4096), dim3(TPB), 0, 0, a, b, n);
        HIPCHK(hipEventRecord(e1, 0));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        if (ms_per_copy) *ms_per_copy = ms / reps;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipFree(a);
        (void)hipFree(b);
        return BH_OK;
    } catch (const DevError& e) {
        return fail(nullptr, BH_EDEVICE, e.msg);
    }
}
int bh_device_info(char* buf, int len) {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        if (buf && len > 0) std::snprintf(buf, len, "no HIP device");
        return BH_EDEVICE;
    }
    if (buf && len > 0)
        std::snprintf(buf, len, "%s %s CUs=%d HBM=%.1fGiB", prop.name, prop.gcnArchName, prop.multiProcessorCount,
                      prop.totalGlobalMem / 1073741824.0);
    return BH_OK;
}

}  // extern "C"
