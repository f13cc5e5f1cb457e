/*
 * hfv_gpu.h -- MI355X entry points of libhydrofv_amd.so beyond the hydroFV API.
 *
 * The hydroFV API itself (System, launch_simulation, muscl_hancock_{1d,2d}, ...)
 * is declared in hydro.h with the reference's layouts.  The functions here are
 * the lower-level device interface those entry points are built on; they take
 * plain pointers, sizes and a hipStream_t passed as void*, and return HFV_OK or
 * a negative HFV_E* code.  Reference interfaces they implement are cited per
 * function.
 */
#ifndef HYDROFV_HFV_GPU_H
#define HYDROFV_HFV_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HFV_OK 0
#define HFV_EINVAL (-1)   /* unsupported or inconsistent configuration */
#define HFV_ENOMEM (-2)   /* device allocation failed */
#define HFV_EDEVICE (-3)  /* a HIP runtime call failed */
#define HFV_ETOOBIG (-4)  /* the grid needs more than the fixed scratch of hfv_mh2d_run_noalloc */

/* Device failure kinds reported by a run (see hfv_run_status.err_kind). */
#define HFV_ERR_NONE 0
#define HFV_ERR_DT_ZERO 1        /* "dt is zero." (integrator_muscl_hancock.c:545-549) */
#define HFV_ERR_HLLC_SPEED 2     /* riemann_solver_hllc.c:372-375 */
#define HFV_ERR_EXACT_BRACKET 3  /* riemann_solver_exact.c:702-707 */
#define HFV_ERR_EXACT_BISECT 4   /* riemann_solver_exact.c:745-754 */

/* Configuration of one MUSCL-Hancock run on a uniform Cartesian 2D grid.
 * Flags use the reference's values: slope_limiter 1/2/3 (slope_limiter.h),
 * riemann_solver 1 exact / 2 HLLC (riemann_solver.h), bc[] 1 none /
 * 2 reflective / 3 transmissive / 4 periodic (boundary.h), in the order
 * x_min, x_max, y_min, y_max. */
typedef struct hfv_mh2d_config {
    int num_cells_x, num_cells_y, num_ghost_cells_side;
    double gamma, gravity;
    double dx, dy;                 /* cell widths used by the CFL dt (utils.c:94) */
    double area_x, area_y, volume; /* uniform face areas and cell volume */
    int slope_limiter, riemann_solver;
    double tol;
    int bc[4];
    double cfl, cfl_initial_shrink_factor;
    int cfl_initial_shrink_num_steps;
    double tf;                     /* run while t < tf */
    long long max_steps;           /* stop after this many steps; < 0: no limit */
    double fixed_dt;               /* > 0: constant dt instead of the CFL dt */
    int strip_rows;                /* rows per wavefront strip; 0: the library's choice (28-40 rows,
                                      fitted to the device's resident wavefronts; fewer for small grids) */
    int hllc_form;                 /* HFV_HLLC_SELECTS (0) or HFV_HLLC_BRANCHES (1), below */
    int first_order;               /* 1: first-order Godunov (godunov_first_order_2d,
                                      integrator_godunov_first_order.c:270-570): no
                                      reconstruction, slope_limiter ignored, 1 ghost
                                      layer suffices; 0: MUSCL-Hancock */
    int full_tiles;                /* 1: every strip computes each of its rows; 0 (default): a strip
                                      whose stencil holds one state bit for bit computes one row
                                      and stores its results in the others (the same bits) */
} hfv_mh2d_config;

/* Two code forms of the HLLC flux with identical results: SELECTS evaluates
 * every branch of the solver and selects, interleaving a cell's x and y
 * faces (fastest where neighbouring faces take different branches, as in
 * shear layers); BRANCHES branches per wavefront (fastest where large
 * regions are uniform).  The ensemble's autotuner times both. */
#define HFV_HLLC_SELECTS 0
#define HFV_HLLC_BRANCHES 1

typedef struct hfv_run_status {
    double t, dt, t_prev;
    long long num_steps;
    int err_kind;
    int cur;
    long long err_cell;
    double smax[2];
    int finished;
    /* exact-solver solves that failed after 5000 bisection iterations where
     * the reference's loop goes on past the failure (1D MUSCL-Hancock faces,
     * random-choice cells of the failing step); 0 on the 2D paths.  The host
     * driver prints the reference's verbose warnings for them. */
    long long bisect_fails;
} hfv_run_status;

typedef struct hfv_mh2d hfv_mh2d;

/* Allocate device state for one 2D run (replaces the per-step scratch of
 * muscl_hancock_2d, integrator_muscl_hancock.c:442-484). */
int hfv_mh2d_create(const hfv_mh2d_config *cfg, hfv_mh2d **out);
void hfv_mh2d_destroy(hfv_mh2d *h);
/* Copy the caller's primitive and conserved arrays (reference layout, ghost
 * cells included, x fastest) to the device. */
int hfv_mh2d_upload(hfv_mh2d *h, const double *const W[4], const double *const U[4], void *stream);
/* Reset time control, compute the first dt from W and enqueue the first step. */
int hfv_mh2d_start(hfv_mh2d *h, void *stream);
/* Enqueue further step launches; each one executes a step unless the run has
 * finished (t >= tf, max_steps reached, or a failure).  use_graph replays a
 * captured hipGraph of 48 launches. */
int hfv_mh2d_advance(hfv_mh2d *h, long long launches, int use_graph, void *stream);
/* Pause for a snapshot (storing.c via integrator_muscl_hancock.c:1147-1155):
 * once a step after step resume_n ends at t >= t_pause, further launches
 * only forward the state until the pause is moved.  Call between batches
 * (after hfv_mh2d_status); t_pause = INFINITY lifts it. */
int hfv_mh2d_set_pause(hfv_mh2d *h, double t_pause, long long resume_n);
/* Synchronise the stream and read the time-control state. */
int hfv_mh2d_status(hfv_mh2d *h, hfv_run_status *st, void *stream);
/* Primitive (c2p + ghost fill) and conserved state back to host arrays;
 * either pointer array may be NULL (with W NULL the device's primitive
 * buffer is left as it is).  After a failed flux (err_kind 2-4) this is the
 * state at the start of the failing step. */
int hfv_mh2d_download(hfv_mh2d *h, double *const W[4], double *const U[4], void *stream);
/* Host-side boundary conditions (custom hooks, boundary.h:17-33): the
 * primitive state of every cell, c2p of its own conserved state with no
 * boundary fill (system.c:763-802), copied to the host; and a primitive
 * state, ghost cells filled by the caller, that the next step reads in place
 * of the device's own fill, its CFL maxima included. */
int hfv_mh2d_primitive(hfv_mh2d *h, double *const W[4], void *stream);
int hfv_mh2d_set_primitive(hfv_mh2d *h, const double *const W[4], void *stream);
/* A custom boundary hook in device form: the reference's
 * set_boundary_condition_cartesian_2d_func (boundary.h:25-33) given the
 * primitive arrays in device memory, rows `pitch` doubles apart (pitch >=
 * num_cells_x + 2 * num_ghost_cells_side), the hipStream_t of the run and the
 * caller's pointer.  It enqueues its own kernels on that stream and returns,
 * without synchronising the device or freeing the arrays.  It can be called
 * again on an unchanged state (a paused or finished run), so its result
 * should depend on the arrays alone. */
#ifndef HFV_DEVICE_BC_2D_FUNC_DEFINED
#define HFV_DEVICE_BC_2D_FUNC_DEFINED
typedef void (*hfv_device_bc_2d_func)(double *density, double *velocity_x, double *velocity_y, double *pressure,
                                      int num_ghost_cells_side, int num_cells_x, int num_cells_y, long long pitch,
                                      void *stream, void *user);
#endif
/* Fill the custom side `side` (0..3 = x_min, x_max, y_min, y_max; cfg->bc of
 * it must be 5, custom) with a device hook (the custom branches of
 * boundary.c:513-777).  Once every custom side has one, each launch of
 * hfv_mh2d_advance first refreshes the primitive buffer on the device: c2p
 * of every cell (system.c:763-802), then the four sides in the reference's
 * order, built-in fills and hooks, then the CFL maxima; the step reads that
 * buffer.  hfv_mh2d_download returns the same filled state.  No graphs are
 * used.  fn NULL removes the side's hook. */
int hfv_mh2d_set_device_bc(hfv_mh2d *h, int side, hfv_device_bc_2d_func fn, void *user);
/* Device addresses of the two conserved buffers and the primitive buffer. */
int hfv_mh2d_fields(hfv_mh2d *h, double **U0, double **U1, double **W, long long *pitch);

/* ---- a 2D run that allocates nothing ------------------------------------ */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  The run of hfv_mh2d_create + upload + start + advance +
 * download, with the same kernels and the same results, on the caller's own
 * device arrays: W[4] (primitive state, ghost cells filled, as the first step
 * reads it) and U[4] (conserved state), rows `pitch` doubles apart (pitch >=
 * num_cells_x + 2 g), each num_cells_y + 2 g rows.  It enqueues `launches`
 * step launches (>= 1; each executes a step unless the run has finished: t >=
 * tf, max_steps reached, or a failure); when the enqueued work ends, U holds
 * the final conserved state and W its c2p plus boundary fill.  The library's
 * working memory (the second conserved buffer and the time-control slots)
 * lives in a static device array of HFV_NOALLOC_SCRATCH_BYTES on each GPU it
 * runs on, held until the process ends; a grid with 4 (num_cells_y + 2 g)
 * pitch doubles beyond HFV_NOALLOC_SCRATCH_BYTES - 256 returns HFV_ETOOBIG.
 * Only the HLLC solver (riemann_solver 2) and the built-in boundary
 * conditions 1-4; strip_rows and hllc_form as in hfv_mh2d_create.  Nothing
 * is allocated, freed or captured: plain kernel launches on `stream` (NULL:
 * the null stream).  st != NULL synchronises the stream and reports the
 * run's state; st == NULL synchronises nothing, so the call can be captured.
 * The first call on a device looks up that device's scratch address
 * (hipGetSymbolAddress) and must not itself be under capture.  Calls share
 * their GPU's scratch and keep the last call's contents there: calls that
 * overlap on one GPU (other streams, other host threads) race. */
#define HFV_NOALLOC_SCRATCH_BYTES (64ll << 20)
int hfv_mh2d_run_noalloc(const hfv_mh2d_config *cfg, double *const W[4], double *const U[4], long long pitch,
                         long long launches, hfv_run_status *st, void *stream);

/* ---- batched ensemble (no reference counterpart) ---------------------- */
/* Independent 2D problems of uneven size advanced together, one launch per
 * step covering every strip of every instance; each instance keeps its own
 * CFL time control and gets bitwise the result of a single-grid run.  All
 * instances share the limiter and the Riemann solver. */
typedef struct hfv_ens hfv_ens;
int hfv_ens_create(int count, const hfv_mh2d_config *cfgs, hfv_ens **out);
void hfv_ens_destroy(hfv_ens *e);
/* Strip height per instance (the tile shape the autotuner picks). */
int hfv_ens_set_strip_rows(hfv_ens *e, const int *rows);
/* HLLC code form of every instance (HFV_HLLC_*; initially cfgs[0].hllc_form). */
int hfv_ens_set_hllc_form(hfv_ens *e, int form);
int hfv_ens_num_tiles(hfv_ens *e);
int hfv_ens_upload(hfv_ens *e, int inst, const double *const W[4], const double *const U[4], void *stream);
/* Reset every instance's time control, first dt from W, first step. */
int hfv_ens_start(hfv_ens *e, void *stream);
int hfv_ens_advance(hfv_ens *e, long long launches, int use_graph, void *stream);
/* st[count]: state of every instance. */
int hfv_ens_status(hfv_ens *e, hfv_run_status *st, void *stream);
/* restore = 0: keep a device copy of every instance's state and time control
 * (allocated at the first call, as large as the ensemble's own buffers);
 * restore = 1: put that copy back (HFV_EINVAL before any save).  Both are
 * stream-ordered device-to-device copies; the autotuner starts every
 * candidate from the same state this way. */
int hfv_ens_checkpoint(hfv_ens *e, int restore, void *stream);
int hfv_ens_download(hfv_ens *e, int inst, double *const W[4], double *const U[4], void *stream);

/* ---- 1D: muscl_hancock_1d (integrator_muscl_hancock.c:26-383) ---------- */
typedef struct hfv_mh1d_config {
    int num_cells_x, num_ghost_cells_side;
    int coord_sys_flag;            /* 1 cartesian, 4 cylindrical, 5 spherical (system.h) */
    double gamma;
    double dx;
    int slope_limiter, riemann_solver;
    double tol;
    int bc[2];
    double cfl, cfl_initial_shrink_factor;
    int cfl_initial_shrink_num_steps;
    double tf;
    long long max_steps;
    double fixed_dt;
    int first_order;               /* 1: first-order Godunov (godunov_first_order_1d,
                                      integrator_godunov_first_order.c:26-268), which
                                      reports a failed flux; 0: MUSCL-Hancock */
    int random_choice;             /* 1: random_choice_1d (integrator_random_choice_1d.c):
                                      Glimm's method, exact Riemann sampling at van der
                                      Corput positions; riemann_solver and slope_limiter
                                      ignored, 1 ghost layer suffices */
    /* Resuming and pausing, for snapshots: the run starts at time t_start
     * after steps_start steps (0, 0: a fresh run; W must then be the state
     * the run left), and with t_pause > 0 it stops after the first step
     * that ends at t >= t_pause. */
    double t_start;
    long long steps_start;
    double t_pause;
} hfv_mh1d_config;

/* Run the whole 1D time loop on the device in one workgroup.  W, U are the
 * caller's host arrays (3 fields each: rho,u,p / m,mx,E) of nx+2g cells,
 * updated in place; area has nx+2g+1 entries, mid and volume nx+2g. */
int hfv_mh1d_run(const hfv_mh1d_config *cfg, double *const W[3], double *const U[3], const double *area,
                 const double *mid, const double *volume, hfv_run_status *st, void *stream);

/* ---- fixed-dt 2D step and its discrete adjoint (no reference symbol) ---- */
/* The 2D MUSCL-Hancock + HLLC step of integrator_muscl_hancock.c:552-1141
 * with a given dt on a doubly periodic Cartesian grid, as a stateless
 * operation on caller-owned device arrays (4 conserved fields, each
 * (nx+2g) columns x (ny+2g) rows with row pitch `pitch` doubles), and the
 * transposed Jacobian of that step.  The state is the interior; ghost
 * entries are neither read nor given gradient.  Neither call allocates; both
 * are asynchronous on `stream` (0: the null stream) and OR HFV_ADJ_* bits
 * into the device int *flags, which the caller clears and reads. */
typedef struct hfv_adj2d_config {
    int num_cells_x, num_cells_y, num_ghost_cells_side;
    long long pitch;               /* row pitch of every field, in doubles (>= nx + 2g) */
    double gamma, dt, tol;
    double area_x, area_y, volume; /* uniform Cartesian geometry (system.c:400-415, 490-503) */
    int slope_limiter;             /* slope_limiter.h flags: minmod 1, Van Leer 2, MC 3 */
} hfv_adj2d_config;

enum {
    HFV_ADJ_FLUX_ERROR = 1,  /* forward: an HLLC flux failed (invalid wave speeds) */
    HFV_ADJ_VACUUM = 2,      /* adjoint: a face took the vacuum branch; its derivative is taken as 0 */
    HFV_ADJ_INVALID = 4      /* adjoint: invalid HLLC wave speeds at a face */
};

/* V = step(U); U and V must not overlap. */
int hfv_fwd2d_step(const hfv_adj2d_config *cfg, double *const U[4], double *const V[4], int *flags, void *stream);
/* lam_out = (d step(U) / dU)^T lam_in on the interior (lam_in: the
 * cotangent of step(U), lam_out: that of U; the step's identity part makes
 * this lam_in plus the flux terms' transposed Jacobian); lam_out must not overlap U or
 * lam_in, and its ghost entries are left untouched. */
int hfv_adj2d_step(const hfv_adj2d_config *cfg, double *const U[4], double *const lam_in[4],
                   double *const lam_out[4], int *flags, void *stream);

/* ---- batched evaluation of the device physics (parity probes) ---------- */
/* f[n][nf] = flux of (wl[n][nf], wr[n][nf]); nf 3 (1D) or 4 (2D); err[n].  The 2D HLLC
 * probe evaluates the kernels' split form and checks it against the plain one:
 * err 99 marks a state where the two disagree. */
int hfv_probe_flux(int nf, int solver, double gamma, double tol, long long n, const double *wl,
                   const double *wr, double *f, int *err);
/* out[n] = x[n]^y[n]: impl 0 the device library pow, impl 1 the kernels'
 * pow of the two-rarefaction guess (riemann_solver.c:191-199). */
int hfv_probe_pow(int impl, long long n, const double *x, const double *y, double *out);
/* out[n] = sqrt(x[n]) as the kernels take it: impl 0 the device library sqrt, impl 1 the
 * kernels' grouped root (library sequence without its range scaling when every argument of
 * the wavefront is in [2^-767, +inf)), impl 2 that unscaled core alone (for x[n] in range).
 * The kernels' roots of sound speeds and HLLC wave-speed factors (utils.c:23-26,
 * riemann_solver_hllc.c:30-40) must equal impl 0 bit for bit. */
int hfv_probe_sqrt(int impl, long long n, const double *x, double *out);
/* out[n] = a[n] / b[n] as the kernels take it: impl 0 the device's division, impl 1 the
 * kernels' grouped form (the division's refinement alone, rcp_core / div_core, when every
 * operand of the wavefront has magnitude in [2^-300, 2^300]; the division otherwise),
 * impl 2 that core alone (for operands in range).  The kernels' divisions of primitive
 * conversions and predictors (system.c:792-799, integrator_muscl_hancock.c:692-740,
 * utils.c:23-26) must equal impl 0 bit for bit. */
int hfv_probe_div(int impl, long long n, const double *a, const double *b, double *out);
/* out[n] = limit_slope(a[n], b[n]) for limiter 1/2/3 (slope_limiter.c:93-115). */
int hfv_probe_limit(int limiter, long long n, const double *a, const double *b, double *out);

/* The exact solver's pressure iteration (riemann_solver_exact.c:604-763) for the left and right
 * states wl[n][4], wr[n][4] (rho, u, v, p): Newton from the guess for cap (0-100) iterations,
 * then, for a face not converged, its finish (bracket scan and bisection) as the kernels take
 * it: impl 0 serially per face, impl 1 by a whole wavefront per face (the set-aside faces of
 * the 2D step).  impl 2: Newton for cap iterations serially, the rest of the loop (to 100
 * iterations) by four lanes per face as the 2D step's set-aside faces take it, then the finish
 * serially.  p[n] = p*, rc[n] = the failure kind (HFV_ERR_*), -1 for a face that converged
 * (p: its iterate), -3 for a vacuum face.  impl 1 must equal impl 0 bit for bit at the same
 * cap, impl 2 must equal impl 0 at cap 100. */
int hfv_probe_exact_finish(int impl, double gamma, double tol, int cap, long long n, const double *wl,
                           const double *wr, double *p, int *rc);

/* Text of the most recent HIP runtime error (for error messages). */
const char *hfv_device_error_string(void);

#ifdef __cplusplus
}
#endif

#endif
