/*
 * hfv_oracle.h -- CPU oracle for the MUSCL-Hancock finite-volume path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of hydroFV's
 * MUSCL-Hancock update, used by tests/, by the smoke check and by the CPU
 * baseline leg of bench.py to check results.  The product (hydrofv_amd)
 * never links, loads or calls it.
 *
 * Every function follows the arithmetic order of the reference function it
 * cites so that results are bitwise equal to the reference built with
 * -ffp-contract=off (x86-64 without -march has no FMA).  Parity status of each
 * function (see tests/test_oracle_*.py):
 *   pinned against oracle/_ref (the reference compiled from its sources) on
 *   random and edge-case inputs, and against committed golden fixtures:
 *     hfo_limit, hfo_sound_speed, hfo_guess_p, hfo_flux_hllc_1d/2d,
 *     hfo_vacuum_1d/2d, hfo_exact_1d, hfo_flux_exact_1d, hfo_geometry,
 *     hfo_p2c, hfo_c2p, hfo_bc, hfo_dt_1d/2d, hfo_gravity_2d,
 *     hfo_geometry_source_1d, hfo_mh1d_run, hfo_mh2d_run (HLLC),
 *     hfo_god1d_run (HLLC, exact), hfo_god2d_run (HLLC), hfo_van_der_corput,
 *     hfo_rc1d_run
 *   pinned by composition and closed forms only (no reference counterpart):
 *     hfo_flux_exact_2d (normal components equal hfo_flux_exact_1d bitwise;
 *     transverse momentum flux = mass flux x upwinded transverse velocity),
 *     hfo_mh2d_run with the exact solver (mirror symmetry, Sedov-Taylor radius).
 */
#ifndef HFV_ORACLE_H
#define HFV_ORACLE_H

#include <stdint.h>

/* Flags use the reference's numeric values (slope_limiter.h, riemann_solver.h,
 * boundary.h, system.h). */
enum { HFO_MINMOD = 1, HFO_VAN_LEER = 2, HFO_MC = 3 };
enum { HFO_EXACT = 1, HFO_HLLC = 2 };
enum { HFO_BC_NONE = 1, HFO_BC_REFLECTIVE = 2, HFO_BC_TRANSMISSIVE = 3, HFO_BC_PERIODIC = 4, HFO_BC_CUSTOM = 5 };

/* custom boundary hooks, with the reference's signatures (boundary.h:17-33) */
typedef void (*hfo_bc1d_fn)(double *rho, double *u, double *p, int num_ghost_cells_side, int num_cells_x);
typedef void (*hfo_bc2d_fn)(double *rho, double *u, double *v, double *p, int num_ghost_cells_side,
                            int num_cells_x, int num_cells_y);
enum { HFO_CART1D = 1, HFO_CART2D = 2, HFO_CYL1D = 4, HFO_SPH1D = 5 };
/* error codes (error.h:15-22) */
enum { HFO_OK = 0, HFO_FAILURE = 1, HFO_VALUE_ERROR = 2 };

typedef struct hfo_params {
    int coord;          /* HFO_CART1D ... */
    int nx, ny, g;      /* interior cells and ghost width */
    double gamma;
    double gravity;
    double dx, dy;
    int limiter;        /* HFO_MINMOD ... */
    int solver;         /* HFO_EXACT / HFO_HLLC */
    double tol;
    int bc[4];          /* x_min, x_max, y_min, y_max */
    int nthreads;       /* OpenMP threads for the 2D sweeps (1 = serial) */
    void *hook[4];      /* HFO_BC_CUSTOM sides: hfo_bc1d_fn / hfo_bc2d_fn */
} hfo_params;

typedef struct hfo_run {
    double cfl;
    double shrink_factor;
    int shrink_steps;
    double tf;           /* run while t < tf */
    int64_t max_steps;   /* stop after this many steps (<0: unlimited) */
    double fixed_dt;     /* >0: use this dt instead of the CFL dt */
    /* outputs */
    int64_t num_steps;
    double t;
    double dt;
    int error;           /* HFO_OK or the first error code */
    int64_t error_cell;
} hfo_run;

/* ---- scalar kernels ---------------------------------------------------- */
double hfo_limit(int limiter, double a, double b);
double hfo_sound_speed(double gamma, double rho, double p);
double hfo_guess_p(double gamma, double rho_l, double u_l, double p_l, double a_l,
                   double rho_r, double u_r, double p_r, double a_r, double tol);
void hfo_vacuum_1d(double gamma, const double wl[3], double a_l, const double wr[3],
                   double a_r, double speed, double sol[3]);
void hfo_vacuum_2d(double gamma, const double wl[4], double a_l, const double wr[4],
                   double a_r, double speed, double sol[4]);
int hfo_flux_hllc_1d(double gamma, double tol, const double wl[3], const double wr[3], double f[3]);
int hfo_flux_hllc_2d(double gamma, double tol, const double wl[4], const double wr[4], double f[4]);
int hfo_exact_1d(double gamma, double tol, const double wl[3], const double wr[3],
                 double speed, double sol[3]);
int hfo_flux_exact_1d(double gamma, double tol, const double wl[3], const double wr[3], double f[3]);
int hfo_flux_exact_2d(double gamma, double tol, const double wl[4], const double wr[4], double f[4]);

/* batched wrappers (n independent evaluations; arrays are [n][3] or [n][4]) */
void hfo_limit_batch(int limiter, int64_t n, const double *a, const double *b, double *out);
void hfo_flux_batch(int dim, int solver, double gamma, double tol, int64_t n,
                    const double *wl, const double *wr, double *f, int *err);

/* ---- grid kernels (layout idx = j*(nx+2g) + i, x fastest) -------------- */
void hfo_geometry(const hfo_params *p, double x_min, double x_max, double y_min, double y_max,
                  double *dx, double *dy, double *mid_x, double *mid_y,
                  double *area_x, double *area_y, double *volume);
/* w, u are arrays of 3 (1D) or 4 (2D) field pointers: rho,vx,[vy],p / m,mx,[my],E */
void hfo_p2c(const hfo_params *p, const double *volume, double *const *w, double *const *u);
void hfo_c2p(const hfo_params *p, const double *volume, double *const *u, double *const *w);
void hfo_bc(const hfo_params *p, double *const *w);
double hfo_dt_1d(const hfo_params *p, double *const *w, double cfl);
double hfo_dt_2d(const hfo_params *p, double *const *w, double cfl);
void hfo_gravity_2d(const hfo_params *p, double dt, double *const *u);
void hfo_geometry_source_1d(const hfo_params *p, double dt, const double *mid_x,
                            const double *volume, double *const *w, double *const *u);

/* One MUSCL-Hancock step (fluxes, update, c2p, BC, source terms), in place.
 * Returns HFO_OK or an error code; *err_cell receives the flat index of the
 * first failing interface's left cell. */
int hfo_mh2d_step(const hfo_params *p, double dt, const double *area_x, const double *area_y,
                  const double *volume, double *const *w, double *const *u, int64_t *err_cell);
int hfo_mh1d_step(const hfo_params *p, double dt, const double *area_x, const double *mid_x,
                  const double *volume, double *const *w, double *const *u, int64_t *err_cell);

/* One first-order Godunov step (integrator_godunov_first_order.c), in place;
 * the limiter is not used. */
int hfo_god2d_step(const hfo_params *p, double dt, const double *area_x, const double *area_y,
                   const double *volume, double *const *w, double *const *u, int64_t *err_cell);
int hfo_god1d_step(const hfo_params *p, double dt, const double *area_x, const double *mid_x,
                   const double *volume, double *const *w, double *const *u, int64_t *err_cell);

/* The time loop of muscl_hancock_{1d,2d} and godunov_first_order_{1d,2d}
 * (no storing, no progress bar). */
void hfo_mh2d_run(const hfo_params *p, hfo_run *r, const double *area_x, const double *area_y,
                  const double *volume, double *const *w, double *const *u);
void hfo_mh1d_run(const hfo_params *p, hfo_run *r, const double *area_x, const double *mid_x,
                  const double *volume, double *const *w, double *const *u);
void hfo_god2d_run(const hfo_params *p, hfo_run *r, const double *area_x, const double *area_y,
                   const double *volume, double *const *w, double *const *u);
void hfo_god1d_run(const hfo_params *p, hfo_run *r, const double *area_x, const double *mid_x,
                   const double *volume, double *const *w, double *const *u);

/* random_choice_1d (integrator_random_choice_1d.c): the van der Corput
 * sequence, one step (step = steps taken before it), the time loop. */
double hfo_van_der_corput(int64_t n);
int hfo_rc1d_step(const hfo_params *p, double dt, int64_t step, const double *mid_x, const double *volume,
                  double *const *w, double *const *u, int64_t *err_cell);
void hfo_rc1d_run(const hfo_params *p, hfo_run *r, const double *area_x, const double *mid_x,
                  const double *volume, double *const *w, double *const *u);

#endif
