/* bhsph: C ABI of the spherical Kerr-Schild engine (MI355X, gfx950).
 *
 * The reference's fluid hot path is rk2integrator::step on an amrtree of
 * Cartesian patches (src/fluid/rk2.hpp:24, src/fluid/solver/integrator.cpp:
 * 262-275; entry point bh_rk2_step in bhfluid.h).  This engine runs the same
 * step structure -- ghost fill, con2prim, PLM + Riemann fluxes, constrained
 * transport, flux divergence + geometric sources, RK2, floors -- on blocks of a
 * logarithmic-r x theta x phi grid in Kerr-Schild coordinates with
 * self-consistent physics, for any number of independent runs advanced in
 * shared launches.  Each entry point names the reference member it stands in
 * for; plain pointers and sizes only; every call returns 0 or a negative code
 * with the reason in bhs_last_error().
 *
 * Layout of one block (16 x 16 x bz active zones in r, theta, phi; bz = 16,
 * or 1 for an axisymmetric engine; two ghost layers):
 *   cell index  ((k+2)*20 + (j+2))*20 + (i+2),  sz = bz + 4 stored phi layers
 *   U [8][20*20*sz]  sqrt(gamma) x (D, S_r, S_theta, S_phi, tau, B^r, B^theta, B^phi)
 *   W [8][20*20*sz]  rho, eps, W v^r, W v^theta, W v^phi, B^r, B^theta, B^phi
 *   Bf[3][nf]        sqrt(gamma) B^d on the faces of direction d (face index
 *                    as the cell index with 21 entries along d);
 *                    nf = max(21*20*sz, 20*20*(sz+1))
 * A run's state arrays hold its blocks in the order of its block table.
 * Block table: blocks[n][4] = level, bi, bj, bk (block coordinates at that
 * level; level l has 2^l times the level-0 zones in each direction).
 */
#ifndef BHSPH_H
#define BHSPH_H

#ifdef __cplusplus
extern "C" {
#endif

#define BHS_OK 0
#define BHS_EINVAL (-1)
#define BHS_EDEVICE (-2)
#define BHS_ESTATE (-3)
#define BHS_ENOMEM (-4)

typedef struct bhs_engine bhs_engine;

/* Spacetime, gas and floors of one run: the reference's compile-time
 * parameters (src/fluid/main.cpp:61-81) and floor constants
 * (src/fluid/grid.hpp:140-143): rho_floor * r^-3/2, eps_floor * r^-3/2, plus
 * a magnetisation ceiling b^2/rho <= sigma_max and a Lorentz-factor ceiling;
 * the Riemann solver (src/fluid/solver/hlld.cpp). */
typedef struct {
    double a, Q, gamma, cfl;
    double rho_floor, eps_floor, sigma_max, gamma_max;
    /* Riemann solver: BHS_HLLD (the reference's five-wave structure,
     * hlld.cpp:230-312, with relativistic intermediate states in the face's
     * orthonormal frame) or BHS_HLL (its two-wave fallback). */
    int riemann;
    int reserved;
} bhs_physics;
#define BHS_HLL 0
#define BHS_HLLD 1

/* Level-0 grid: nr x nth x nph zones over r in [r_in, r_out] (logarithmic),
 * theta in [0, pi], phi in [0, 2 pi) (nph ignored by an axisymmetric engine);
 * nlevels refinement levels. */
typedef struct {
    int nr, nth, nph, nlevels;
    double r_in, r_out;
} bhs_grid;

/* Fishbone-Moncrief torus (torus::init::fm_init and B_pot_init,
 * src/fluid/solver/initconds.cpp:121-668): inner edge, pressure maximum,
 * polytropic constant; A_phi = max(rho/rho_max - a_cut, 0), the field scaled
 * so that max p_gas / max p_mag = beta; pressure times 1 + perturbation (2u-1)
 * with u output n of numpy's PCG64 stream of the given 128-bit state and
 * increment, n = the zone's index in the run (level by level, phi slowest,
 * r fastest). */
typedef struct {
    double r_in, r_max, kappa, beta, a_cut, perturbation;
    unsigned long long pcg_state_hi, pcg_state_lo, pcg_inc_hi, pcg_inc_lo;
} bhs_torus;

/* A device engine with room for max_blocks blocks of shape 16 x 16 x bz
 * (grid::amrtree's storage, src/fluid/grid.hpp:109-144). */
int bhs_create(int bz, long long max_blocks, bhs_engine** out);
void bhs_destroy(bhs_engine* e);
const char* bhs_last_error(const bhs_engine* e);
/* Device bytes one block slot takes (state, RK store, face fields, fluxes). */
long long bhs_block_bytes(int bz);

/* Add a run with a caller-owned block table (grid::amrtree's quilt,
 * src/fluid/grid.hpp:119, after regrid, grid.cpp:739-835).  Levels must be
 * nested and 2:1 balanced.  *run receives its id. */
int bhs_run_add(bhs_engine* e, const bhs_physics* phys, const bhs_grid* grid, int nblocks, const int* blocks,
                int* run);
/* Free a run's block slots and tables. */
int bhs_run_remove(bhs_engine* e, int run);
int bhs_run_nblocks(bhs_engine* e, int run, int* nblocks);
/* Initial data on the device (main.cpp:94-159 with the torus above).
 * info[4] = rho_max (unnormalised), max p_gas, max p_mag before scaling, the
 * field scale factor. */
int bhs_run_init_torus(bhs_engine* e, int run, const bhs_torus* torus, double* info);
/* A run's state: the active zones of U and W and the own faces of Bf (a
 * block owns the faces of its zones, both ends along each direction).
 * Ghost cells and ghost faces are derived data, refilled every stage as far
 * as the step reads them; get returns them as 0 and set may pass anything
 * there.  Any pointer may be NULL. */
int bhs_run_get_state(bhs_engine* e, int run, double* U, double* W, double* Bf);
int bhs_run_set_state(bhs_engine* e, int run, const double* U, const double* W, const double* Bf);

/* nsteps of rk2integrator::step (integrator.cpp:262-275) for the listed
 * runs together, each with its own CFL dt; dt_out[nsteps][nruns] (may be
 * NULL). */
int bhs_steps(bhs_engine* e, int nruns, const int* runs, int nsteps, double* dt_out);
/* bhs_steps with a HIP event recorded on the engine's stream between
 * consecutive steps: ms_per_step[nsteps] = each step's device time (the
 * steps replay the captured graph).  No reference counterpart. */
int bhs_steps_timed(bhs_engine* e, int nruns, const int* runs, int nsteps, double* ms_per_step);
/* Per-kernel timing with HIP events (measurement only; launches one by
 * one): enable, optionally restricted to one kernel name; the report is
 * "name launches total_ms" lines; returns the report's length. */
int bhs_profile(bhs_engine* e, int enable, const char* only);
int bhs_profile_report(bhs_engine* e, char* buf, int len);
/* con2prim with floors of a run's active zones (W of the current U;
 * integrator.cpp:201-222). */
int bhs_c2p(bhs_engine* e, int run);
/* out[8] = rest mass, Killing energy (sum of alpha (tau + D) - beta^j S_j,
 * densitised), max |div B| over the active zones (constrans::maxdivB,
 * src/fluid/grid/transport.cpp:77-89), the divergence stencil's scale,
 * con2prim failures and floored zones since the run was added, the
 * accretion rate Mdot and the horizon magnetic flux Phi_B on the zone row
 * holding the outer horizon (analyser Mdot / Phi_B,
 * src/fluid/data/quantitatives.cpp:99-118). */
int bhs_run_diagnostics(bhs_engine* e, int run, double* out);
/* The analyser quantities that read derived primitives
 * (src/fluid/data/quantitatives.cpp:109-193), in covariant form from u^mu and
 * b^mu: out[10] = thermal energy int rho eps, magnetic energy int b^2/2 (both
 * over sqrt(-g) dV), Ldot = -int T^r_phi and L_BZ = -int (b^2 u^r u_t -
 * b^r b_t) over the horizon row (sqrt(-g) dtheta dphi), eta_BZ = L_BZ/|Mdot|,
 * phi_BH = Phi_B/sqrt(|Mdot|) (the magnetically-arrested measure; the
 * reference's MAD is Phi_B/Mdot), P_beta = int p / int b^2/2, alpha_SS =
 * int (rho h u^r u_phi - b^r b_phi) / int (p + b^2/2) and the Maxwell to
 * Reynolds stress ratio, both over zones with rho >= rho_min, and Mdot. */
int bhs_run_analysis(bhs_engine* e, int run, double rho_min, double* out);
/* bhs_run_diagnostics for callers that cannot allocate (code under graph
 * capture, tight serving loops, harnesses that forbid hipMalloc); no
 * counterpart in the reference.  Same out[8], computed without any device or
 * host allocation: the per-block partials and their reduction go to a
 * static device array of 65536 blocks x 6 words (3 MB per GPU, held until
 * the process ends); runs of more blocks are rejected (BHS_EINVAL).  The
 * engine's block tables must already be on the device (after any step or
 * initial data of a run since the last run was added), else BHS_ESTATE.
 * Calls that overlap on one GPU share the array and race. */
int bhs_run_diagnostics_static(bhs_engine* e, int run, double* out);

/* Launch configuration (no reference counterpart): mixed 1 = FP32
 * reconstruction, Riemann fluxes and EMFs (con2prim, sources, update and
 * reductions FP64), rounded as the CPU rounds them; 2 = the same with FP32
 * division and square root from the hardware reciprocal and square-root
 * instructions (within a few ulp; faster, not bit-identical to a CPU).
 * Schedule ("persistent"): 0 = one kernel per sub-phase;
 * 1 = one queue-driven kernel per RK stage with wg_per_cu workgroups per
 * compute unit (0 = as many as fit); 2 = slab: the sub-phase kernels with the
 * fluxes and the hydro update fused (a block's phi layers walked by 2^(flux
 * variant mod 5) workgroups, the fluxes kept in LDS); 3 = schedule 0 with
 * stage 1's con2prim done by stage 0's update kernel, zone by zone; 4 = one
 * launch per RK stage (the rkstg of integrator.cpp:191-260) whose phases --
 * con2prim, ghost fill, fluxes, face update, zone update -- are work queues
 * over every block of every level, pulled by resident workgroups and
 * separated by grid barriers (wg_per_cu as for 1; the flux variant's
 * occupancy hint applies).  Graphs on by default.  Every schedule gives the
 * same bits. */
int bhs_set_precision(bhs_engine* e, int mixed);
int bhs_set_schedule(bhs_engine* e, int persistent, int wg_per_cu);
int bhs_use_graph(bhs_engine* e, int enable);
/* Launch variants of the flux, update and con2prim kernels (0 .. 5; flux
 * also 6 .. 11: the theta and phi faces computed by lines of faces, one
 * thread per line): workgroup size and minimum wavefronts per SIMD of each
 * are given by bhs_variant_info.  Results do not depend on the variant; the
 * autotuner (bhfluid.tune) times them.  Defaults: flux 0, update 0,
 * con2prim 1.  Under the slab schedule the flux variant v sets the
 * workgroups per block to 2^(v mod 5) (at most bz). */
int bhs_set_variants(bhs_engine* e, int flux, int update, int c2p);
int bhs_variant_info(int variant, int* tpb, int* min_waves);
/* Variant of the face-field update (constrained transport, the spherical
 * form of transport.cpp:11-75): 0 = one own face per thread, its four EMFs
 * formed from the flux words; 1 .. 4 = each edge's EMF formed once into LDS
 * by 2, 4, 8 or 16 workgroups per block (at most bz), then the faces updated
 * from LDS (the default is 2).  Results do not depend on the variant. */
int bhs_set_face_variant(bhs_engine* e, int face);
/* Number the workgroups of the multi-workgroup kernels so that each of the
 * eight XCDs (own L2) takes a contiguous range of blocks (1, the default) or
 * in dispatch order (0). */
int bhs_set_xcd_remap(bhs_engine* e, int enable);
/* Polls of a grid-barrier wait in schedules 1 and 4 before the barrier gives
 * up (0: the default, 2^26).  A launch whose workgroups are not all resident
 * then ends with every workgroup exiting and the next bhs_steps returning
 * BHS_EDEVICE with "grid barrier timed out"; a small value forces that path
 * (for tests of the error handling).  No reference counterpart. */
int bhs_set_barrier_spin_limit(bhs_engine* e, unsigned int polls);
int bhs_sync(bhs_engine* e);

#ifdef __cplusplus
}
#endif
#endif
