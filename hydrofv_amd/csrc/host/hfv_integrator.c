/*
 * hfv_integrator.c -- muscl_hancock_1d / muscl_hancock_2d (integrator.h:145-176)
 * godunov_first_order_1d / godunov_first_order_2d (integrator.h:99-130) and
 * random_choice_1d (integrator.h:76-84) running on the MI355X.
 *
 * The reference's time loop (integrator_muscl_hancock.c:124-349, 530-1160)
 * becomes: copy the System's arrays to the device once, run the steps there
 * (time control included, see hfv_ctrl.hpp), copy the final primitive and
 * conserved arrays back.  SimulationStatus and the error codes/messages follow
 * the reference.
 */
#include <pthread.h>
#include <stdio.h>
#include <string.h>

#include "hydrofv/hfv_gpu.h"
#include "hydrofv/hydro.h"

const char *hfv_device_error_string(void);

static ErrorStatus device_error(int rc)
{
    char msg[256];
    if (rc == HFV_ENOMEM) return WRAP_RAISE_ERROR(MEMORY_ERROR, "Failed to allocate device memory.");
    if (rc == HFV_EINVAL) return WRAP_RAISE_ERROR(VALUE_ERROR, "Configuration not supported by the device integrator.");
    snprintf(msg, sizeof msg, "HIP runtime error: %s", hfv_device_error_string());
    return WRAP_RAISE_ERROR(FAILURE, msg);
}

/* The reference's verbose warnings (riemann_solver_exact.c:736-744) for
 * `solves` exact solves that failed after 5000 bisection iterations: each
 * warned at every 500th iteration before failing.  A bisection that stops
 * short of 5000 iterations takes fewer than 500 unless p* lies below 2^-448
 * of the bracket's width, so the failed solves are the ones that warn. */
static void bisect_warnings(long long solves, int verbose)
{
    if (verbose <= 0) return;
    char buf[128];
    for (long long i = 0; i < solves; i++) {
        for (int c = 500; c <= 5000; c += 500) {
            snprintf(buf, sizeof buf, "Bisection method did not converge after %d iterations for exact riemann solver.",
                     c);
            WRAP_RAISE_WARNING(buf);
        }
    }
}

/* A device failure as the reference reports it; a failed exact solve warns
 * first when verbose > 0 (the run stops at the first failure, so one solve
 * warned), unless the device has counted it among st.bisect_fails. */
static ErrorStatus run_error(int kind, int verbose)
{
    if (kind == HFV_ERR_EXACT_BISECT) bisect_warnings(1, verbose);
    switch (kind) {
    case HFV_ERR_DT_ZERO:
        return WRAP_RAISE_ERROR(VALUE_ERROR, "dt is zero.");
    case HFV_ERR_HLLC_SPEED:
        return WRAP_RAISE_ERROR(VALUE_ERROR,
                                "Invalid wave speed for the HLLC riemann solver. Possible fix: use a smaller CFL.");
    case HFV_ERR_EXACT_BRACKET:
        return WRAP_RAISE_ERROR(FAILURE, "Failed to bracket the root for bisection method.");
    case HFV_ERR_EXACT_BISECT:
        return WRAP_RAISE_ERROR(FAILURE,
                                "Bisection method did not converge after 5001 iterations for exact riemann solver.");
    default:
        return WRAP_RAISE_ERROR(UNKNOWN_ERROR, "Unknown device failure.");
    }
}

static void fill_status(SimulationStatus *out, const hfv_run_status *st)
{
    out->num_steps = st->num_steps;
    out->t = st->t;
    out->dt = st->dt;
    if (st->err_kind == HFV_ERR_DT_ZERO) {
        out->dt = 0.0;
    } else if (st->err_kind != HFV_ERR_NONE) {
        /* the failing step is not counted (integrator_muscl_hancock.c:831-834) */
        out->num_steps = st->num_steps - 1;
        out->t = st->t_prev;
    }
}

static int uniform(const double *a, long n)
{
    for (long k = 1; k < n; k++)
        if (a[k] != a[0]) return 0;
    return 1;
}

static ErrorStatus common_checks(const System *system, const BoundaryConditionParam *bc, const StoringParam *sp,
                                 int two_d)
{
    (void)sp;
    const int flags[4] = {bc->boundary_condition_flag_x_min_, bc->boundary_condition_flag_x_max_,
                          bc->boundary_condition_flag_y_min_, bc->boundary_condition_flag_y_max_};
    for (int k = 0; k < (two_d ? 4 : 2); k++) {
        if (flags[k] < BOUNDARY_CONDITION_NONE || flags[k] > BOUNDARY_CONDITION_CUSTOM)
            return WRAP_RAISE_ERROR(VALUE_ERROR, "Boundary condition flag not recognized.");
    }
    const int g = system->num_ghost_cells_side;
    if (system->num_cells_x < g || (two_d && system->num_cells_y < g))
        return WRAP_RAISE_ERROR(VALUE_ERROR, "The device integrator needs at least as many cells per direction as ghost cells.");
    return make_success_error_status();
}

static int has_custom_bc(const BoundaryConditionParam *bc, int two_d)
{
    return bc->boundary_condition_flag_x_min_ == BOUNDARY_CONDITION_CUSTOM ||
           bc->boundary_condition_flag_x_max_ == BOUNDARY_CONDITION_CUSTOM ||
           (two_d && (bc->boundary_condition_flag_y_min_ == BOUNDARY_CONDITION_CUSTOM ||
                      bc->boundary_condition_flag_y_max_ == BOUNDARY_CONDITION_CUSTOM));
}

/* Device forms of custom 2D hooks, keyed by the host hook they stand for. */
#define DEVICE_BC_MAX 64
typedef struct {
    set_boundary_condition_cartesian_2d_func host;
    hfv_device_bc_2d_func device;
    void *user;
} DeviceBc;
static DeviceBc device_bc[DEVICE_BC_MAX];
static int device_bc_count;
static pthread_mutex_t device_bc_lock = PTHREAD_MUTEX_INITIALIZER;

int hfv_register_device_bc_2d(set_boundary_condition_cartesian_2d_func host_hook, hfv_device_bc_2d_func device_hook,
                              void *user)
{
    if (!host_hook) return -1;
    int rc = 0;
    pthread_mutex_lock(&device_bc_lock);
    int k = 0;
    while (k < device_bc_count && device_bc[k].host != host_hook) k++;
    if (!device_hook) {
        if (k < device_bc_count) device_bc[k] = device_bc[--device_bc_count];
    } else if (k < device_bc_count || device_bc_count < DEVICE_BC_MAX) {
        if (k == device_bc_count) device_bc_count++;
        device_bc[k].host = host_hook;
        device_bc[k].device = device_hook;
        device_bc[k].user = user;
    } else {
        rc = -1;
    }
    pthread_mutex_unlock(&device_bc_lock);
    return rc;
}

static int find_device_bc(set_boundary_condition_cartesian_2d_func host_hook, DeviceBc *out)
{
    int found = 0;
    pthread_mutex_lock(&device_bc_lock);
    for (int k = 0; k < device_bc_count && !found; k++) {
        if (device_bc[k].host == host_hook) {
            *out = device_bc[k];
            found = 1;
        }
    }
    pthread_mutex_unlock(&device_bc_lock);
    return found;
}

/* Gives every custom side of a 2D run its registered device hook; 0 (and
 * none set) when some custom side has none, so the run keeps the host path. */
static int attach_device_bc(const BoundaryConditionParam *bc, hfv_mh2d *h)
{
    const int flags[4] = {bc->boundary_condition_flag_x_min_, bc->boundary_condition_flag_x_max_,
                          bc->boundary_condition_flag_y_min_, bc->boundary_condition_flag_y_max_};
    const set_boundary_condition_cartesian_2d_func hooks[4] = {
        bc->set_boundary_condition_cartesian_2d_x_min, bc->set_boundary_condition_cartesian_2d_x_max,
        bc->set_boundary_condition_cartesian_2d_y_min, bc->set_boundary_condition_cartesian_2d_y_max};
    DeviceBc found[4];
    for (int s = 0; s < 4; s++)
        if (flags[s] == BOUNDARY_CONDITION_CUSTOM && !find_device_bc(hooks[s], &found[s])) return 0;
    for (int s = 0; s < 4; s++) {
        if (flags[s] == BOUNDARY_CONDITION_CUSTOM && hfv_mh2d_set_device_bc(h, s, found[s].device, found[s].user) != HFV_OK) {
            for (int k = 0; k < s; k++)  /* all or none: back to the host path */
                if (flags[k] == BOUNDARY_CONDITION_CUSTOM) (void)hfv_mh2d_set_device_bc(h, k, NULL, NULL);
            return 0;
        }
    }
    return 1;
}

/* Snapshot bookkeeping of the reference loop (integrator_muscl_hancock.c:
 * 509-525, 1146-1155): the initial snapshot when asked for, then one after
 * each step that ends at t >= interval * (stored - initial + 1). */
typedef struct {
    int on;
    int offset;        /* 1 when the initial state was stored */
    long long resume;  /* steps taken when the last snapshot was written */
} Storing;

static double next_store_time(const StoringParam *sp, const Storing *s)
{
    return sp->storing_interval * (double)(sp->store_count_ - s->offset + 1);
}

static ErrorStatus storing_begin(Storing *s, const BoundaryConditionParam *bc, const System *system,
                                 const IntegratorParam *ip, SimulationStatus *status, StoringParam *sp)
{
    s->on = sp->is_storing ? 1 : 0;
    s->offset = (sp->is_storing && sp->store_initial) ? 1 : 0;
    s->resume = 0;
    if (s->on && sp->store_initial)
        return WRAP_TRACEBACK(store_snapshot(bc, system, ip, status, sp));
    return make_success_error_status();
}

/* Runs the 2D loop; max_steps < 0 and fixed_dt <= 0 give the reference loop.
 * first_order: the first-order Godunov scheme instead of MUSCL-Hancock. */
static ErrorStatus run_2d(BoundaryConditionParam *bc, System *system, IntegratorParam *ip, StoringParam *sp,
                          SimulationParam *prm, SimulationStatus *status, long long max_steps, double fixed_dt,
                          int first_order, int verbose)
{
    char msg[256];
    if (system->coord_sys_flag_ != COORD_SYS_CARTESIAN_2D) {
        snprintf(msg, sizeof msg, "Wrong coordinate system. Supported coordinate system: \"cartesian_2d\", got: \"%s\"",
                 system->coord_sys ? system->coord_sys : "");
        return WRAP_RAISE_ERROR(VALUE_ERROR, msg);
    }
    ErrorStatus e = WRAP_TRACEBACK(common_checks(system, bc, sp, 1));
    if (e.return_code != SUCCESS) return e;
    const int g = system->num_ghost_cells_side;
    const long tx = system->num_cells_x + 2 * g, ty = system->num_cells_y + 2 * g;
    if (!uniform(system->surface_area_x_, tx) || !uniform(system->surface_area_y_, ty) ||
        !uniform(system->volume_, tx * ty)) {
        return WRAP_RAISE_ERROR(NOT_IMPLEMENTED_ERROR, "The device integrator needs uniform Cartesian face areas and volumes.");
    }

    hfv_mh2d_config cfg;
    memset(&cfg, 0, sizeof cfg);
    cfg.num_cells_x = system->num_cells_x;
    cfg.num_cells_y = system->num_cells_y;
    cfg.num_ghost_cells_side = g;
    cfg.gamma = system->gamma;
    cfg.gravity = system->gravity;
    cfg.dx = system->dx_;
    cfg.dy = system->dy_;
    cfg.area_x = system->surface_area_x_[0];
    cfg.area_y = system->surface_area_y_[0];
    cfg.volume = system->volume_[0];
    cfg.slope_limiter = ip->slope_limiter_flag_;
    cfg.riemann_solver = ip->riemann_solver_flag_;
    cfg.tol = ip->tol;
    cfg.bc[0] = bc->boundary_condition_flag_x_min_;
    cfg.bc[1] = bc->boundary_condition_flag_x_max_;
    cfg.bc[2] = bc->boundary_condition_flag_y_min_;
    cfg.bc[3] = bc->boundary_condition_flag_y_max_;
    cfg.cfl = ip->cfl;
    cfg.cfl_initial_shrink_factor = ip->cfl_initial_shrink_factor;
    cfg.cfl_initial_shrink_num_steps = ip->cfl_initial_shrink_num_steps;
    cfg.tf = prm->tf;
    cfg.max_steps = max_steps;
    cfg.fixed_dt = fixed_dt;
    cfg.first_order = first_order;

    status->num_steps = 0;
    status->dt = 0.0;
    status->t = 0.0;
    sp->store_count_ = 0;
    Storing store;
    e = storing_begin(&store, bc, system, ip, status, sp);
    if (e.return_code != SUCCESS) return e;

    hfv_mh2d *h = NULL;
    int rc = hfv_mh2d_create(&cfg, &h);
    if (rc != HFV_OK) return device_error(rc);
    double *W[4] = {system->density_, system->velocity_x_, system->velocity_y_, system->pressure_};
    double *U[4] = {system->mass_, system->momentum_x_, system->momentum_y_, system->energy_};
    hfv_run_status st;
    memset(&st, 0, sizeof st);
    rc = hfv_mh2d_upload(h, (const double *const *)W, (const double *const *)U, NULL);
    if (rc == HFV_OK && store.on) rc = hfv_mh2d_set_pause(h, next_store_time(sp, &store), 0);
    /* custom hooks: on the device when every one has a registered device
     * form, else on the host around every step */
    const int custom = has_custom_bc(bc, 1) && !(rc == HFV_OK && attach_device_bc(bc, h));
    if (rc == HFV_OK) rc = hfv_mh2d_start(h, NULL);
    if (rc == HFV_OK && (store.on || custom)) {
        /* batches of queued launches; the device pauses after each step
         * that is due a snapshot, which is written from the host arrays.
         * With a custom boundary hook on the host every step goes one at a
         * time: the primitive state comes to the host, the reference's
         * boundary fill runs there (hooks included, boundary.c:513-777) and
         * the filled state goes back for the next step. */
        long long batch = 15, filled = 0;
        for (;;) {
            rc = hfv_mh2d_status(h, &st, NULL);
            if (rc != HFV_OK) break;
            if (custom && st.err_kind == HFV_ERR_NONE && st.num_steps > filled) {
                rc = hfv_mh2d_primitive(h, W, NULL);
                if (rc != HFV_OK) break;
                e = WRAP_TRACEBACK(set_boundary_condition(bc, system));
                if (e.return_code != SUCCESS) {
                    hfv_mh2d_destroy(h);
                    return e;
                }
                rc = hfv_mh2d_set_primitive(h, (const double *const *)W, NULL);
                if (rc != HFV_OK) break;
                filled = st.num_steps;
            }
            if (store.on && st.err_kind == HFV_ERR_NONE && st.num_steps > store.resume &&
                st.t >= next_store_time(sp, &store)) {
                rc = hfv_mh2d_download(h, custom ? NULL : W, U, NULL);
                if (rc != HFV_OK) break;
                fill_status(status, &st);
                e = WRAP_TRACEBACK(store_snapshot(bc, system, ip, status, sp));
                if (e.return_code != SUCCESS) {
                    hfv_mh2d_destroy(h);
                    return e;
                }
                store.resume = st.num_steps;
                rc = hfv_mh2d_set_pause(h, next_store_time(sp, &store), store.resume);
                if (rc != HFV_OK) break;
                continue;
            }
            if (st.finished) break;
            rc = custom ? hfv_mh2d_advance(h, 1, 0, NULL) : hfv_mh2d_advance(h, batch, 1, NULL);
            if (rc != HFV_OK) break;
            batch = (batch < 48) ? 48 : (batch < 480 ? 2 * batch : 480);
        }
    } else if (rc == HFV_OK) {
        /* batches of queued launches between status checks, growing to 480;
         * a launch after the end only forwards the state.  With a step limit
         * no more than the remaining max_steps - 1 launches are queued, and
         * a run that reaches tf or fails first stops at its next check. */
        long long batch = 15, left = (max_steps >= 0) ? (max_steps > 1 ? max_steps - 1 : 0) : -1;
        for (;;) {
            rc = hfv_mh2d_status(h, &st, NULL);
            if (rc != HFV_OK || st.finished || left == 0) break;
            const long long n = (left > 0 && left < batch) ? left : batch;
            rc = hfv_mh2d_advance(h, n, 1, NULL);
            if (rc != HFV_OK) break;
            if (left > 0) left -= n;
            batch = (batch < 48) ? 48 : (batch < 480 ? 2 * batch : 480);
        }
    }
    /* with custom hooks the host arrays already hold the filled primitive state */
    if (rc == HFV_OK) rc = hfv_mh2d_download(h, custom && st.err_kind == HFV_ERR_NONE ? NULL : W, U, NULL);
    hfv_mh2d_destroy(h);
    if (rc != HFV_OK) return device_error(rc);
    fill_status(status, &st);
    if (st.err_kind != HFV_ERR_NONE) return run_error(st.err_kind, verbose);
    return make_success_error_status();
}

/* scheme: 0 MUSCL-Hancock, 1 first-order Godunov, 2 random choice */
static ErrorStatus run_1d(BoundaryConditionParam *bc, System *system, IntegratorParam *ip, StoringParam *sp,
                          SimulationParam *prm, SimulationStatus *status, long long max_steps, double fixed_dt,
                          int scheme, int verbose)
{
    const int f = system->coord_sys_flag_;
    if (f != COORD_SYS_CARTESIAN_1D && f != COORD_SYS_CYLINDRICAL_1D && f != COORD_SYS_SPHERICAL_1D) {
        /* integrator_godunov_first_order.c:39-61 */
        char msg[256];
        snprintf(msg, sizeof msg,
                 "Wrong coordinate system. Supported coordinate system: \"cartesian_1d\", \"cylindrical_1d\" and "
                 "\"spherical_1d\", got: \"%s\"",
                 system->coord_sys ? system->coord_sys : "");
        return WRAP_RAISE_ERROR(VALUE_ERROR, msg);
    }
    ErrorStatus e = WRAP_TRACEBACK(common_checks(system, bc, sp, 0));
    if (e.return_code != SUCCESS) return e;
    hfv_mh1d_config cfg;
    memset(&cfg, 0, sizeof cfg);
    cfg.num_cells_x = system->num_cells_x;
    cfg.num_ghost_cells_side = system->num_ghost_cells_side;
    cfg.coord_sys_flag = f;
    cfg.gamma = system->gamma;
    cfg.dx = system->dx_;
    cfg.slope_limiter = ip->slope_limiter_flag_;
    cfg.riemann_solver = ip->riemann_solver_flag_;
    cfg.tol = ip->tol;
    cfg.bc[0] = bc->boundary_condition_flag_x_min_;
    cfg.bc[1] = bc->boundary_condition_flag_x_max_;
    cfg.cfl = ip->cfl;
    cfg.cfl_initial_shrink_factor = ip->cfl_initial_shrink_factor;
    cfg.cfl_initial_shrink_num_steps = ip->cfl_initial_shrink_num_steps;
    cfg.tf = prm->tf;
    cfg.max_steps = max_steps;
    cfg.fixed_dt = fixed_dt;
    cfg.first_order = scheme == 1;
    cfg.random_choice = scheme == 2;
    if (scheme == 2 && has_custom_bc(bc, 0)) {
        /* the hook would have to run between the sampling and the conversion
         * to conserved variables, inside the device step */
        return WRAP_RAISE_ERROR(NOT_IMPLEMENTED_ERROR,
                                "Custom boundary conditions with the random-choice integrator are not provided by this library.");
    }
    status->num_steps = 0;
    status->dt = 0.0;
    status->t = 0.0;
    sp->store_count_ = 0;
    Storing store;
    e = storing_begin(&store, bc, system, ip, status, sp);
    if (e.return_code != SUCCESS) return e;
    double *W[3] = {system->density_, system->velocity_x_, system->pressure_};
    double *U[3] = {system->mass_, system->momentum_x_, system->energy_};
    hfv_run_status st;
    const int custom = has_custom_bc(bc, 0);
    for (;;) {
        /* With storing, the run goes in segments that end at the steps due
         * a snapshot and resume from the state written back to the host.
         * With a custom boundary hook each segment is one step, after which
         * the reference's boundary fill runs on the host (hooks included,
         * boundary.c:411-511). */
        if (store.on) cfg.t_pause = next_store_time(sp, &store);
        if (custom) cfg.max_steps = (max_steps < 0 || cfg.steps_start + 1 < max_steps) ? cfg.steps_start + 1 : max_steps;
        const int rc =
            hfv_mh1d_run(&cfg, W, U, system->surface_area_x_, system->mid_points_x_, system->volume_, &st, NULL);
        if (rc != HFV_OK) return device_error(rc);
        bisect_warnings(st.bisect_fails, verbose);
        if (st.err_kind != HFV_ERR_NONE || st.num_steps == cfg.steps_start) break;
        if (custom) {
            e = WRAP_TRACEBACK(set_boundary_condition(bc, system));
            if (e.return_code != SUCCESS) return e;
        }
        const int due = store.on && st.num_steps > store.resume && st.t >= cfg.t_pause;
        if (due) {
            fill_status(status, &st);
            e = WRAP_TRACEBACK(store_snapshot(bc, system, ip, status, sp));
            if (e.return_code != SUCCESS) return e;
            store.resume = st.num_steps;
        }
        const int done = !(st.t < prm->tf) || (max_steps >= 0 && st.num_steps >= max_steps);
        if (done || (!due && !custom)) break;
        cfg.t_start = st.t;
        cfg.steps_start = st.num_steps;
    }
    fill_status(status, &st);
    if (st.err_kind != HFV_ERR_NONE) return run_error(st.err_kind, scheme == 2 ? 0 : verbose);
    return make_success_error_status();
}

ErrorStatus hfv_muscl_hancock_2d_run(BoundaryConditionParam *bc, System *system, IntegratorParam *ip,
                                     StoringParam *sp, Settings *settings, SimulationParam *prm,
                                     SimulationStatus *status, long long max_steps, double fixed_dt)
{
    return run_2d(bc, system, ip, sp, prm, status, max_steps, fixed_dt, 0, settings ? settings->verbose : 0);
}

ErrorStatus hfv_muscl_hancock_1d_run(BoundaryConditionParam *bc, System *system, IntegratorParam *ip,
                                     StoringParam *sp, Settings *settings, SimulationParam *prm,
                                     SimulationStatus *status, long long max_steps, double fixed_dt)
{
    return run_1d(bc, system, ip, sp, prm, status, max_steps, fixed_dt, 0, settings ? settings->verbose : 0);
}

ErrorStatus hfv_godunov_first_order_2d_run(BoundaryConditionParam *bc, System *system, IntegratorParam *ip,
                                           StoringParam *sp, Settings *settings, SimulationParam *prm,
                                           SimulationStatus *status, long long max_steps, double fixed_dt)
{
    return run_2d(bc, system, ip, sp, prm, status, max_steps, fixed_dt, 1, settings ? settings->verbose : 0);
}

ErrorStatus hfv_godunov_first_order_1d_run(BoundaryConditionParam *bc, System *system, IntegratorParam *ip,
                                           StoringParam *sp, Settings *settings, SimulationParam *prm,
                                           SimulationStatus *status, long long max_steps, double fixed_dt)
{
    return run_1d(bc, system, ip, sp, prm, status, max_steps, fixed_dt, 1, settings ? settings->verbose : 0);
}

ErrorStatus hfv_random_choice_1d_run(BoundaryConditionParam *bc, System *system, IntegratorParam *ip,
                                     StoringParam *sp, Settings *settings, SimulationParam *prm,
                                     SimulationStatus *status, long long max_steps, double fixed_dt)
{
    return run_1d(bc, system, ip, sp, prm, status, max_steps, fixed_dt, 2, settings ? settings->verbose : 0);
}

ErrorStatus random_choice_1d(BoundaryConditionParam *bc, System *system, IntegratorParam *ip, StoringParam *sp,
                             Settings *settings, SimulationParam *prm, SimulationStatus *status)
{
    return hfv_random_choice_1d_run(bc, system, ip, sp, settings, prm, status, -1, 0.0);
}

ErrorStatus muscl_hancock_2d(BoundaryConditionParam *bc, System *system, IntegratorParam *ip, StoringParam *sp,
                             Settings *settings, SimulationParam *prm, SimulationStatus *status)
{
    return hfv_muscl_hancock_2d_run(bc, system, ip, sp, settings, prm, status, -1, 0.0);
}

ErrorStatus muscl_hancock_1d(BoundaryConditionParam *bc, System *system, IntegratorParam *ip, StoringParam *sp,
                             Settings *settings, SimulationParam *prm, SimulationStatus *status)
{
    return hfv_muscl_hancock_1d_run(bc, system, ip, sp, settings, prm, status, -1, 0.0);
}

ErrorStatus godunov_first_order_2d(BoundaryConditionParam *bc, System *system, IntegratorParam *ip,
                                   StoringParam *sp, Settings *settings, SimulationParam *prm,
                                   SimulationStatus *status)
{
    return hfv_godunov_first_order_2d_run(bc, system, ip, sp, settings, prm, status, -1, 0.0);
}

ErrorStatus godunov_first_order_1d(BoundaryConditionParam *bc, System *system, IntegratorParam *ip,
                                   StoringParam *sp, Settings *settings, SimulationParam *prm,
                                   SimulationStatus *status)
{
    return hfv_godunov_first_order_1d_run(bc, system, ip, sp, settings, prm, status, -1, 0.0);
}
