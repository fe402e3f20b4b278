// Host-side block tree: the product's counterpart of grid::amrtree
// (src/fluid/grid.hpp:109-144).  It owns the block table and the host copy of
// the state, builds the initial torus and performs regrids; the per-step work
// runs on the GPU (bh_device.hip) over flat copies of these arrays.
#pragma once

#include <array>
#include <memory>
#include <string>
#include <vector>

#include "bh_common.hpp"
#include "bh_physics.hpp"

namespace bh {

// Kerr-Newman-Schild metric components at (r, theta), metriccomp of
// src/fluid/metric.hpp:11-21.
struct Metric {
    double g[4][4], gi[4][4], G[4][4][4], sqrtdetg, gam[3][3], gami[3][3], sqrtdetgam, alpha, beta[3];
};
Metric metric_comp(double M, double a, double Q, double r, double th);
GPt gpt_from(const Metric& m);
GCell gcell_from(const Metric& m);

struct Patch {
    int lvl = 0;
    std::array<int, 3> ijk{};
    Patch* parent = nullptr;
    std::array<Patch*, 8> children{};
    bool leaf = true;
    bool refine = false;
    bool erase = false;
    int dev = -1;  // slot of this block in the device arrays before a device regrid
    double xe[BLK + 1], ye[BLK + 1], ze[BLK + 1];
    std::vector<double> U, W, Bf;  // [8][8000], [8][8000], [3][8400]
    std::vector<double> r, th, xc, yc, zc, vol;  // per stored cell

    Patch(int lvl, std::array<int, 3> ijk, std::array<double, 3> c1, std::array<double, 3> c2, Patch* parent);
    double dx() const { return xe[1] - xe[0]; }
    double dy() const { return ye[1] - ye[0]; }
    double dz() const { return ze[1] - ze[0]; }
    double& u(int v, int c) { return U[(size_t)v * NCELL + c]; }
    double& w(int v, int c) { return W[(size_t)v * NCELL + c]; }
    double* bfx() { return &Bf[0]; }
    double* bfy() { return &Bf[NFACE]; }
    double* bfz() { return &Bf[2 * NFACE]; }
    Prim getW(int c) const;
    void setW(int c, const Prim& p);
    Cons getU(int c) const;
    void setU(int c, const Cons& q);
    void B_init();
};

struct TreeParams {
    double M, a, Q, gamma;
};

// Structural record of a regrid whose data work runs elsewhere (on the
// device): the refined leaves in refine order and, per flattened parent, the
// slots its removed children held.
struct RegridLog {
    std::vector<Patch*> refined;
    std::vector<std::pair<Patch*, std::array<int, 8>>> flattened;
};

class HostTree {
public:
    TreeParams prm;
    std::vector<std::unique_ptr<Patch>> quilt;
    long long faults = 0;     // equation-of-state calls with rho <= 0 on the host
    std::string last_msg;     // diagnostics of the last initialisation call

    HostTree(const double lo[3], const double hi[3], int nqlt, double M, double a, double Q, double gm);

    // initial data, src/fluid/solver/initconds.cpp and main.cpp:98-159
    int fm_init();
    double rho_max() const;
    void B_pot_init(Patch& p, double glmx_rho);
    void ptoc_interior();
    // mesh adaptation, grid.cpp:574-835
    void regrid();
    void set_blocks(int nb, const int* meta);
    struct LeafVerdict {
        bool refine;    // asks to be refined
        bool erasable;  // may be merged back into its parent
    };
    LeafVerdict leaf_verdict(Patch* p) const;
    void append_children(Patch* p);
    void refine(Patch* p);
    void ccrstr(Patch* p);
    void flatten(Patch* p);
    // regrid's tree changes alone, from per-leaf flags computed elsewhere
    // (leaf order = quilt order; bit 0 refine, bit 1 erase); the data work of
    // refine and ccrstr is left to the caller, who reads it from `log`
    void regrid_structure(const std::vector<unsigned char>& leaf_flags, RegridLog& log);

    int nblocks() const { return (int)quilt.size(); }
    int index_of(const Patch* p) const;
    void export_blocks(int* meta, double* edges) const;
    void export_state(double* U, double* W, double* Bf) const;
    void import_state(const double* U, const double* W, const double* Bf);

    GPt tree_point(double r, double th) const;
    Prim pvfs(double r) const;
    void floor_or_keep(Patch& p, int c, const Cons& q);

private:
    std::vector<Patch*> erased_;
};

// Static per-block geometry consumed by the kernels (the metric does not
// change during a run), computed with the same libm calls as the reference.
constexpr int NGCT = 8;   // tree metric at cells: sqrtdetgam, alpha, g11, g13, g22, g33, gi13, gi33
constexpr int NGCC = 11;  // cell metric at cells: GCell fields
constexpr int NGFC = 8;   // tree metric at faces: sqrtdetg, sqrtdetgam, alpha, beta0, g11, g13, g22, g33
// Lowered Christoffels Gamma_{m n l} = g_{m s} Gamma^s_{n l} of the cell metric
// (a = 0): only these 14 (m, n<=l) entries can be non-zero; the other 26 are
// exactly +0.0 (sums of zero products started from +0.0), which
// build_geometry verifies cell by cell.
constexpr int NCHR = 14;
constexpr int kChrIdx[NCHR][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 2, 2}, {0, 3, 3}, {1, 0, 0}, {1, 0, 1},
                                  {1, 1, 1}, {1, 2, 2}, {1, 3, 3}, {2, 1, 2}, {2, 3, 3}, {3, 1, 3}, {3, 2, 3}};
struct GeomHost {
    std::vector<double> cell_tree;  // [nb][NGCT][8000]
    std::vector<double> cell_cell;  // [nb][NGCC][8000]
    std::vector<double> floors;     // [nb][2][8000]  rho_floor, eps_floor
    std::vector<double> face;       // [nb][3][NGFC][4352]
    std::vector<double> chr;        // [nb][NCHR][4096] the non-zero lowered Christoffels
    std::vector<double> dxyz;       // [nb][3]
    std::vector<double> th;         // [nb][8000] polar angle of each cell centre (patch::th)
};
void build_geometry(const HostTree& t, GeomHost& g);
// the same tables for the listed blocks only, packed in list order
void build_geometry_blocks(const HostTree& t, const std::vector<int>& blocks, GeomHost& g);
BH_HD int sym4(int n, int l) { return n <= l ? n * (7 - n) / 2 + l : l * (7 - l) / 2 + n; }

// Ghost-fill plan: for every leaf and face side the neighbour found by
// amrtree::nbhd (grid.cpp:361-386) and the fill kind, plus the launch waves
// that keep the reference's serial fill order where a coarse-to-fine
// prolongation reads ghost cells another leaf fills in the same phase.
enum { GH_COPY = 0, GH_PROLONG = 1, GH_BNDY = 2 };
struct GhostPlan {
    std::vector<int> leaves;              // leaf block ids in quilt order
    std::vector<int> tab;                 // [nb][6][2] (kind, neighbour), side = 2*dim + (side>0)
    std::vector<std::vector<int>> waves;  // leaf ids per wave
    // [nb][6]: 1 when every ghost cell a same-level copy fills on that side
    // has the same tree-metric bits as its source cell, so that the ghost's
    // con2prim equals the source's (see mark_exact_copies)
    std::vector<unsigned char> same;
};
void build_ghost_plan(const HostTree& t, GhostPlan& g);
void mark_exact_copies(const GeomHost& geo, GhostPlan& g);

}  // namespace bh
