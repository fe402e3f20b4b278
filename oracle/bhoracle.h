/* Test infrastructure: the oracle.  A CPU restatement of the reference's
 * GRMHD fluid step (Fjoerde/black-hole-simulator, src/fluid) used only by
 * tests/, the smoke check and the CPU-baseline leg of bench.py, as the
 * checker.  The product library never links or loads it.
 *
 * State layout (identical to the product's):
 *   U [nb][8][8000]  D, S0, S1, S2, tau, B0, B1, B2     (cons, cons.hpp:14-59)
 *   W [nb][8][8000]  rho, eps, v0, v1, v2, B0, B1, B2   (prim, prim.hpp:14-27)
 *   Bf[nb][3][8400]  Bfx, Bfy, Bfz                      (grid.cpp:40-42,72-80)
 *   dU[nb][8][8000]  RK accumulator                     (cell.hpp:20)
 * cell index = (i+2)*400 + (j+2)*20 + (k+2)             (grid.cpp:51-53)
 * Block table: meta[nb][14] = lvl, ijk[3], leaf, parent, children[8];
 *              edges[nb][3][17] = the patch's xedge, yedge, zedge (grid.cpp:21-29)
 *
 * libm mode 0 uses the C library's log/exp (what the reference calls);
 * mode 1 uses correctly rounded log/exp (quad precision, rounded once),
 * which is what the product's device code implements.
 */
#ifndef BHORACLE_H
#define BHORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

void* or_tree_new(int nb, const int* meta, const double* edges, double M, double a, double Q, double gm,
                  int libm_mode);
void or_tree_free(void* t);
/* The initial data of main.cpp:94-159 on a fresh tree: amrtree's constructor
 * state (grid.cpp:177-206), fm_init, global rho_max, B_pot_init, ptoc of the
 * active zones.  libm_mode chooses the functions the initial data calls
 * (0 = C library, 1 = correctly rounded); the geometry is the C library's.
 * Returns 0, 1 (L0_comp throws), 2 (no torus placed), 3 (rho_max < 1e-14);
 * faults counted as in or_step. */
int or_init_torus(void* t, int libm_mode, double* rho_max);
void or_set_state(void* t, const double* U, const double* W, const double* Bf, const double* dU);
void or_get_state(void* t, double* U, double* W, double* Bf, double* dU);
/* Each returns the number of equation-of-state calls that saw rho <= 0 (the
 * reference throws std::domain_error there, thermo.cpp:11). */
long long or_step(void* t, double cfl, double* dt);
long long or_dtcomp(void* t, double cfl, double* dt);
long long or_rkstg(void* t, double dt, int stage);
long long or_floors(void* t);
long long or_ghosts_all(void* t);
long long or_c2p_leaves(void* t);
long long or_flux_ct_patch(void* t, int b, double dt, double* F, double* E);
long long or_update_terms(void* t, int b, double* dF, double* S);
double or_maxdivB(void* t, int b);
void or_cell_geom(void* t, int b, double* out);
/* out[7] = mass, energy, L_pol, Edot, Phi_B, thermE, Hscal (quantitatives.cpp);
 * abs[8] = sum of |term| of each integral (Hscal: numerator [6], denominator [7]). */
void or_diagnostics(void* t, double* out, double* abs);
void or_set_threads(int n);
void or_set_limiter(void* t, int lim);

/* pointwise */
void or_metric(double M, double a, double Q, double r, double th, double* out120);
int or_ptoc(double M, double a, double Q, double gm, const double* w, double r, double th, double* u,
            double* der4);
int or_ctop(double M, double a, double Q, double gm, const double* u, double r, double th, double* w,
            double* der5, int libm_mode);
int or_hlld(double M, double a, double Q, double gm, const double* WL, const double* WR, const double* UL,
            const double* UR, double r, double th, int dim, double* F, double* ct);
void or_reconfp(void* t, int b, int i, int j, int k, int dim, int lim, double* L, double* R);
double or_log(double x, int libm_mode);
double or_exp(double x, int libm_mode);
/* the initial data's functions: op 0 log, 1 exp, 2 sin, 3 cos, 4 acos of x,
 * 5 pow(x, y), 6 atan2(x, y); mode as or_init_torus */
double or_ic_fn(int op, double x, double y, int libm_mode);

#ifdef __cplusplus
}
#endif
#endif
