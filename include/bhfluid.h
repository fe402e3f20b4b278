/* bhfluid: C ABI of the MI355X (gfx950) GRMHD fluid step.
 *
 * Drop-in for the reference's fluid hot path (Fjoerde/black-hole-simulator,
 * src/fluid).  The reference exposes this path only as C++ classes; each
 * entry point below names the reference member it replaces.  Plain pointers
 * and sizes only; every call returns 0 on success, BH_EDOMAIN (1) when the
 * step completed but the equation of state saw rho <= 0 (where the reference
 * throws std::domain_error, src/fluid/physics/thermo.cpp:11), or a negative
 * code with the reason in bh_last_error().
 *
 * State layout (structure of arrays, one block = one 16^3 patch with two
 * ghost layers, cell index (i+2)*400 + (j+2)*20 + (k+2), k fastest):
 *   U [nb][8][8000]  D, S0, S1, S2, tau, B0, B1, B2      (cons, src/fluid/cons.hpp:14-59)
 *   W [nb][8][8000]  rho, eps, v0, v1, v2, B0, B1, B2    (prim, src/fluid/prim.hpp:14-27)
 *   Bf[nb][3][8400]  face fields Bfx, Bfy, Bfz in the patch's face indexing
 *                    (src/fluid/grid/grid.cpp:40-42, 72-80)
 * Block table: meta[nb][14] = lvl, i, j, k, leaf, parent, children[8]
 * (-1 = none); edges[nb][3][17] = the patch's x, y, z cell-edge coordinates.
 * Block order is the reference's quilt order.
 */
#ifndef BHFLUID_H
#define BHFLUID_H

#ifdef __cplusplus
extern "C" {
#endif

#define BH_OK 0
#define BH_EDOMAIN 1
#define BH_EINVAL (-1)
#define BH_EDEVICE (-2)
#define BH_ESTATE (-3)

typedef struct bh_tree bh_tree;

/* grid::amrtree::amrtree(dom_l, dom_h, nqlt, M, a, Q, gm), src/fluid/grid/grid.cpp:177-206:
 * nqlt^3 root patches, floor state in every cell. */
int bh_tree_create(const double dom_l[3], const double dom_h[3], int nqlt, double M, double a, double Q, double gm,
                   bh_tree** out);
void bh_tree_destroy(bh_tree* t);
/* Message of the last failing call on t (or of the last failing bh_tree_create if t is NULL). */
const char* bh_last_error(const bh_tree* t);

/* torus::init::fm_init(tree), src/fluid/solver/initconds.cpp:228-591.
 * Returns 0, or 2 where the reference prints an error and returns. */
int bh_tree_fm_init(bh_tree* t);
/* Maximum rho over active zones, src/fluid/main.cpp:101-111. */
int bh_tree_rho_max(bh_tree* t, double* rho_max);
/* torus::init::B_pot_init(patch, mtr, rho_max) on every patch in quilt order,
 * src/fluid/solver/initconds.cpp:594-668 as called at src/fluid/main.cpp:142-145. */
int bh_tree_B_pot_init(bh_tree* t, double rho_max);
/* U = conserved::ptoc(W) on active zones, src/fluid/main.cpp:148-159. */
int bh_tree_ptoc_interior(bh_tree* t);
/* grid::amrtree::regrid(), src/fluid/grid/grid.cpp:739-835. */
int bh_tree_regrid(bh_tree* t);
/* Numerical options of the step, which the reference fixes in its source.
 * limiter: the PLM slope limiter of reconfp (src/fluid/recon.hpp:16-34):
 * BH_LIM_MINMOD, BH_LIM_VANLEER (the one the reference passes,
 * src/fluid/grid/grid.cpp:261,285,309; the default) or BH_LIM_MC.  The
 * reserved words must be zero.  Takes effect from the next step. */
#define BH_LIM_MINMOD 0
#define BH_LIM_VANLEER 1
#define BH_LIM_MC 2
typedef struct {
    int limiter;
    int reserved[7];
} bh_params;
int bh_tree_set_params(bh_tree* t, const bh_params* p);
int bh_tree_get_params(const bh_tree* t, bh_params* p);
/* Load a caller-owned block table in place of the tree's: the quilt of a
 * grid::amrtree (src/fluid/grid.hpp:111-120), meta[nb][14] in the layout of
 * bh_tree_get_blocks (level, i, j, k, leaf, parent, children[8]) and in quilt
 * order.  Blocks without parent must be roots of the grid the tree was
 * created with; every other block must follow its parent and sit in the
 * octant its ijk names.  Patch boxes are derived as amrtree::refine derives
 * them (grid.cpp:574-600), so edges and cell geometry equal the caller's
 * tree bit for bit.  Every zone is reset to the floor state: set the state
 * next (bh_tree_set_state).  BH_EINVAL (with the offending block in
 * bh_last_error) for an inconsistent table. */
int bh_tree_set_blocks(bh_tree* t, int nb, const int* meta);
/* The same regrid (grid.cpp:739-835) with the state on the GPU: the
 * refinement flags are computed on the device, the block table changes on
 * the host, and prolongation (refine), restriction (ccrstr), their con2prim
 * and floor fallback and patch::B_init run on the device, which keeps the
 * state; only the geometry of new blocks is built on the host.  Results are
 * identical to bh_tree_regrid.  *changed (may be NULL) = blocks refined plus
 * parents flattened. */
int bh_tree_regrid_device(bh_tree* t, int* changed);
/* The whole initial data of src/fluid/main.cpp:94-159 computed on the GPU, the
 * state left on the device: the floor state of grid::amrtree's constructor
 * (grid.cpp:177-206) in every stored cell of every block, fm_init
 * (solver/initconds.cpp:228-591), the global rho_max over active zones
 * (main.cpp:101-111), B_pot_init on every patch (initconds.cpp:594-668) and
 * U = ptoc(W) on active zones (main.cpp:148-159).  For a tree fresh from
 * bh_tree_create this is the sequence bh_tree_fm_init, bh_tree_rho_max,
 * bh_tree_B_pot_init, bh_tree_ptoc_interior, with the initial data's
 * transcendental functions (sin, cos, pow, exp, log, atan2, acos) correctly
 * rounded instead of the C library's.  *rho_max (may be NULL) = the global
 * rho_max.  Returns 0; 2 where fm_init places no torus (the rest is done, as
 * main.cpp continues); BH_ESTATE where L0_comp's denominator vanishes (nothing
 * written) or rho_max < 1e-14 (main.cpp stops there). */
int bh_tree_init_torus_device(bh_tree* t, double* rho_max);

int bh_tree_nblocks(const bh_tree* t, int* nb);
int bh_tree_get_blocks(const bh_tree* t, int* meta, double* edges);
/* Copy the state out of / into the tree (any pointer may be NULL). */
int bh_tree_get_state(bh_tree* t, double* U, double* W, double* Bf);
int bh_tree_set_state(bh_tree* t, const double* U, const double* W, const double* Bf);

/* integ::rk2integrator::step(tree), src/fluid/solver/integrator.cpp:262-275:
 * dtcomp, two RK stages, floors; dt returned. */
int bh_rk2_step(bh_tree* t, double cfl, double* dt);
/* rk2integrator::dtcomp, integrator.cpp:156-189 */
int bh_rk2_dtcomp(bh_tree* t, double cfl, double* dt);
/* rk2integrator::rkstg(tree, dt, stage), integrator.cpp:191-260 */
int bh_rk2_rkstg(bh_tree* t, double dt, int stage);
/* patch::floors on every patch, as at integrator.cpp:269-273 */
int bh_tree_floors(bh_tree* t);
/* The time loop of src/fluid/main.cpp:169-173: step while t < t_end and
 * steps < nsteps; *t_io is advanced, *steps_out counts steps taken.  Where the
 * reference's loop ends with std::domain_error (the first step whose equation
 * of state sees rho <= 0), this loop completes that step, counts it in
 * *t_io and *steps_out, stops, and returns BH_EDOMAIN. */
int bh_rk2_advance(bh_tree* t, double cfl, int nsteps, double t_end, double* t_io, int* steps_out);
/* nsteps consecutive steps enqueued back to back with a single host
 * synchronisation at the end (the step sequence of bh_rk2_advance without a
 * t_end test); dt_out[nsteps] receives each step's dt.  No counterpart in
 * the reference; used by the benchmark. */
int bh_rk2_steps(bh_tree* t, double cfl, int nsteps, double* dt_out);
/* Run steps from a captured HIP graph of the whole step (default 1) or launch the kernels
 * one by one (0).  Results are identical; no reference counterpart.  Timing with
 * bh_profile always launches one by one. */
int bh_tree_use_graph(bh_tree* t, int enable);
/* Testing switch without a reference counterpart: enable 0 makes every copied
 * ghost cell run its own con2prim instead of reusing the source zone's outcome
 * where the two cells' metric bits agree (the default, enable 1).  Results are
 * identical either way; the switch exercises the non-reuse ghost path. */
int bh_tree_set_exact_copies(bh_tree* t, int enable);
/* Build the device copy (geometry, ghost plan, state upload) ahead of the
 * first step, so that set-up is not timed with it. */
int bh_tree_prepare(bh_tree* t);
/* Per-kernel timing with HIP events on the tree's stream: enable, then read
 * "name launches total_ms" lines, followed by "queue:<pass> cells 0" lines
 * counting the cells each con2prim pass handed to the Brent fallback.
 * Measurement only. */
int bh_profile(bh_tree* t, int enable);
/* Restrict the timing to the kernels of one name ("" or NULL: all kernels). */
int bh_profile_only(bh_tree* t, const char* name);
int bh_profile_report(bh_tree* t, char* buf, int len);
int bh_device_sync(void);
/* Equation-of-state domain faults (rho <= 0) since the last call. */
int bh_tree_take_faults(bh_tree* t, long long* n);
/* The analysis::analyser quantities that read only the evolved state,
 * src/fluid/data/quantitatives.cpp, as device reductions over the leaf active
 * zones: out[7] = mass (:82-86), energy (:87-91), L_pol (:92-96), Edot and
 * Phi_B at the horizon (:103-107, :114-118), thermE (:129-133), Hscal at
 * 10 M (:199-208).  Each zone's term is the reference's; the sums run in a
 * fixed order of their own (results agree with the reference's sequential
 * sums to rounding and are identical from run to run). */
int bh_tree_diagnostics(bh_tree* t, double* out);
/* Write W of every leaf's active zones to a versioned binary file for the
 * renderer (the reference leaves this format as a stub, src/fluid/data.hpp):
 * a 128-byte header ("BHFLUIDW", version 1, leaf count, 8 variables, 4096
 * zones per leaf, M, a, Q, gamma, time, step), a 64-byte entry per leaf
 * (level, i, j, k, lower and upper corner), then per leaf rho, eps, v0..2,
 * B0..2 over the active zones (k fastest), leaves in quilt order.  The data
 * are packed on the GPU when the device holds the current state. */
int bh_tree_write_dump(bh_tree* t, const char* path, double time, long long step);
/* constrans::maxdivB of block b, src/fluid/grid/transport.cpp:77-89 */
int bh_tree_maxdivB(bh_tree* t, int b, double* out);
/* constrans::maxdivB (transport.cpp:77-89) of every block computed on the
 * device from the device state (per_block[nb], may be NULL) and its maximum
 * over the leaves (*max_leaf, may be NULL); the reference defines the
 * function but never calls it. */
int bh_tree_maxdivB_device(bh_tree* t, double* per_block, double* max_leaf);

/* Pointwise device entry points, for parity tests.  Inputs are host arrays
 * of n points; the metric (M, a, Q) is evaluated at each (r, th).
 *   conserved::ptoc  src/fluid/physics/conserved.cpp:17-73
 *   conserved::ctop  src/fluid/physics/conserved.cpp:135-230 (ok[i] = 1 on success)
 *   hlld             src/fluid/solver/hlld.cpp:230-312 */
int bh_ptoc_batch(double M, double a, double Q, double gm, int n, const double* r, const double* th,
                  const double* W, double* U);
int bh_ctop_batch(double M, double a, double Q, double gm, int n, const double* r, const double* th,
                  const double* U, double* W, int* ok);
int bh_hlld_batch(double M, double a, double Q, double gm, int n, const double* r, const double* th, int dim,
                  const double* WL, const double* WR, const double* UL, const double* UR, double* F,
                  double* ctspeed);
/* Correctly rounded functions of the step and the initial data, for parity tests (no
 * reference counterpart): op 0 log, 1 exp, 2 sin, 3 cos, 4 acos of x[i]; on_device 0
 * evaluates the same code on the host. */
int bh_math_batch(int op, int n, const double* x, double* y, int on_device);
/* op 0: out[i] = pow(x[i], y[i]); op 1: out[i] = atan2(x[i], y[i]); correctly rounded. */
int bh_math2_batch(int op, int n, const double* x, const double* y, double* out, int on_device);
/* The con2prim energy residual (conserved::fres, src/fluid/physics/conserved.cpp:80-123)
 * at in[i] = (xi, Ssq, Bsq, SB, D, tau): f[i] = the residual, sign[i] = its sign as
 * proved by the division-free evaluation the Brent and bracket loops use (+1, -1, or 0
 * when unproved).  For parity tests of that evaluation; no reference counterpart. */
int bh_c2p_sign_batch(double gm, int n, const double* in, int* sign, double* f, int on_device);

/* Measurement helper without a reference counterpart: times a plain device
 * copy of nbytes (read nbytes, write nbytes), the calibration kernel for the
 * HBM traffic counters. */
int bh_copy_probe(long long nbytes, int reps, double* ms_per_copy);
/* Device and build information. */
int bh_device_info(char* buf, int len);

#ifdef __cplusplus
}
#endif
#endif
