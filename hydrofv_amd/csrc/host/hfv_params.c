/*
 * hfv_params.c -- parameter structs, string-to-flag resolution, validation and
 * the launch_simulation entry (the reference's integrator.c, boundary.c:51-409,
 * riemann_solver.c:23-45, slope_limiter.c:18-47, storing.c:27-100 and
 * hydro.c:29-265).  Option names, defaults and error messages are kept
 * because drivers and their users depend on them.
 */
#define _POSIX_C_SOURCE 199309L
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "hydrofv/hydro.h"

/* ---- boundary conditions (boundary.c:16-409) --------------------------- */
BoundaryConditionParam get_new_boundary_condition_param(void)
{
    BoundaryConditionParam p;
    memset(&p, 0, sizeof p);
    p.boundary_condition_flag_x_min_ = -1;
    p.boundary_condition_flag_x_max_ = -1;
    p.boundary_condition_flag_y_min_ = -1;
    p.boundary_condition_flag_y_max_ = -1;
    p.boundary_condition_flag_z_min_ = -1;
    p.boundary_condition_flag_z_max_ = -1;
    return p;
}

static int bc_flag(const char *name)
{
    if (strcmp(name, "none") == 0) return BOUNDARY_CONDITION_NONE;
    if (strcmp(name, "reflective") == 0) return BOUNDARY_CONDITION_REFLECTIVE;
    if (strcmp(name, "transmissive") == 0) return BOUNDARY_CONDITION_TRANSMISSIVE;
    if (strcmp(name, "periodic") == 0) return BOUNDARY_CONDITION_PERIODIC;
    if (strcmp(name, "custom") == 0) return BOUNDARY_CONDITION_CUSTOM;
    return -1;
}

static ErrorStatus resolve_side(const char *name, int *flag, const char *side)
{
    char msg[128];
    if (!name) {
        snprintf(msg, sizeof msg, "Boundary condition for %s is NULL.", side);
        return WRAP_RAISE_ERROR(POINTER_ERROR, msg);
    }
    *flag = bc_flag(name);
    if (*flag < 0) {
        snprintf(msg, sizeof msg, "Boundary condition for %s is not recognized.", side);
        return WRAP_RAISE_ERROR(VALUE_ERROR, msg);
    }
    return make_success_error_status();
}

ErrorStatus finalize_boundary_condition_param(const System *system, BoundaryConditionParam *p)
{
    ErrorStatus e;
    const int f = system->coord_sys_flag_;
    const int two_d = (f == COORD_SYS_CARTESIAN_2D);
    if (f == COORD_SYS_CARTESIAN_3D) {
        return WRAP_RAISE_ERROR(NOT_IMPLEMENTED_ERROR, "Cartesian 3D systems are not supported.");
    }
    if (f != COORD_SYS_CARTESIAN_1D && f != COORD_SYS_CYLINDRICAL_1D && f != COORD_SYS_SPHERICAL_1D && !two_d) {
        return WRAP_RAISE_ERROR(VALUE_ERROR, "Coordinate system flag not recognized.");
    }
    if (two_d) {
        e = resolve_side(p->boundary_condition_y_min, &p->boundary_condition_flag_y_min_, "y_min");
        if (e.return_code != SUCCESS) return e;
        e = resolve_side(p->boundary_condition_y_max, &p->boundary_condition_flag_y_max_, "y_max");
        if (e.return_code != SUCCESS) return e;
    }
    e = resolve_side(p->boundary_condition_x_min, &p->boundary_condition_flag_x_min_, "x_min");
    if (e.return_code != SUCCESS) return e;
    e = resolve_side(p->boundary_condition_x_max, &p->boundary_condition_flag_x_max_, "x_max");
    if (e.return_code != SUCCESS) return e;

    /* periodic must be paired (boundary.c:258-305) */
    if (two_d && (p->boundary_condition_flag_y_min_ == BOUNDARY_CONDITION_PERIODIC ||
                  p->boundary_condition_flag_y_max_ == BOUNDARY_CONDITION_PERIODIC) &&
        p->boundary_condition_flag_y_min_ != p->boundary_condition_flag_y_max_) {
        return WRAP_RAISE_ERROR(VALUE_ERROR, "Periodic boundary condition mismatch for y_min and y_max.");
    }
    if ((p->boundary_condition_flag_x_min_ == BOUNDARY_CONDITION_PERIODIC ||
         p->boundary_condition_flag_x_max_ == BOUNDARY_CONDITION_PERIODIC) &&
        p->boundary_condition_flag_x_min_ != p->boundary_condition_flag_x_max_) {
        return WRAP_RAISE_ERROR(VALUE_ERROR, "Periodic boundary condition mismatch for x_min and x_max.");
    }

    /* custom sides need their hook (boundary.c:307-371) */
    if (two_d) {
        if (p->boundary_condition_flag_x_min_ == BOUNDARY_CONDITION_CUSTOM && !p->set_boundary_condition_cartesian_2d_x_min)
            return WRAP_RAISE_ERROR(POINTER_ERROR, "Custom boundary condition pointer for x_min is NULL.");
        if (p->boundary_condition_flag_x_max_ == BOUNDARY_CONDITION_CUSTOM && !p->set_boundary_condition_cartesian_2d_x_max)
            return WRAP_RAISE_ERROR(POINTER_ERROR, "Custom boundary condition pointer for x_max is NULL.");
        if (p->boundary_condition_flag_y_min_ == BOUNDARY_CONDITION_CUSTOM && !p->set_boundary_condition_cartesian_2d_y_min)
            return WRAP_RAISE_ERROR(POINTER_ERROR, "Custom boundary condition pointer for y_min is NULL.");
        if (p->boundary_condition_flag_y_max_ == BOUNDARY_CONDITION_CUSTOM && !p->set_boundary_condition_cartesian_2d_y_max)
            return WRAP_RAISE_ERROR(POINTER_ERROR, "Custom boundary condition pointer for y_max is NULL.");
    } else {
        if (p->boundary_condition_flag_x_min_ == BOUNDARY_CONDITION_CUSTOM && !p->set_boundary_condition_1d_x_min)
            return WRAP_RAISE_ERROR(POINTER_ERROR, "Custom boundary condition pointer for x_min is NULL.");
        if (p->boundary_condition_flag_x_max_ == BOUNDARY_CONDITION_CUSTOM && !p->set_boundary_condition_1d_x_max)
            return WRAP_RAISE_ERROR(POINTER_ERROR, "Custom boundary condition pointer for x_max is NULL.");
    }
    return make_success_error_status();
}

/* ---- integrator parameters (integrator.c:16-106) ----------------------- */
IntegratorParam get_new_integrator_param(void)
{
    IntegratorParam p;
    p.integrator = NULL;
    p.riemann_solver = NULL;
    p.slope_limiter = NULL;
    p.cfl = 0.5;
    p.cfl_initial_shrink_factor = 0.2;
    p.cfl_initial_shrink_num_steps = 10;
    p.tol = 1e-6;
    p.integrator_flag_ = -1;
    p.riemann_solver_flag_ = -1;
    p.slope_limiter_flag_ = -1;
    return p;
}

ErrorStatus get_riemann_solver_flag(IntegratorParam *p)
{
    if (!p->riemann_solver) return WRAP_RAISE_ERROR(POINTER_ERROR, "Riemann solver is not set.");
    if (strcmp(p->riemann_solver, "riemann_solver_exact") == 0) p->riemann_solver_flag_ = RIEMANN_SOLVER_EXACT;
    else if (strcmp(p->riemann_solver, "riemann_solver_hllc") == 0) p->riemann_solver_flag_ = RIEMANN_SOLVER_HLLC;
    else return WRAP_RAISE_ERROR(VALUE_ERROR, "Riemann solver not recognized.");
    return make_success_error_status();
}

ErrorStatus get_slope_limiter_flag(IntegratorParam *p)
{
    if (!p || !p->slope_limiter) return WRAP_RAISE_ERROR(POINTER_ERROR, "Integrator parameter or slope limiter is NULL.");
    if (strcmp(p->slope_limiter, "minmod") == 0) p->slope_limiter_flag_ = SLOPE_LIMITER_MINMOD;
    else if (strcmp(p->slope_limiter, "van_leer") == 0) p->slope_limiter_flag_ = SLOPE_LIMITER_VAN_LEER;
    else if (strcmp(p->slope_limiter, "monotonized_central") == 0) p->slope_limiter_flag_ = SLOPE_LIMITER_MONOTONIZED_CENTRAL;
    else return WRAP_RAISE_ERROR(VALUE_ERROR, "Slope limiter not recognized.");
    return make_success_error_status();
}

static ErrorStatus get_integrator_flag(IntegratorParam *p)
{
    static const struct { const char *name; int flag; } names[] = {
        {"random_choice_1d", INTEGRATOR_RANDOM_CHOICE_1D},
        {"godunov_first_order_1d", INTEGRATOR_GODUNOV_FIRST_ORDER_1D},
        {"godunov_first_order_2d", INTEGRATOR_GODUNOV_FIRST_ORDER_2D},
        {"muscl_hancock_1d", INTEGRATOR_MUSCL_HANCOCK_1D},
        {"muscl_hancock_2d", INTEGRATOR_MUSCL_HANCOCK_2D},
    };
    if (!p->integrator) return WRAP_RAISE_ERROR(POINTER_ERROR, "Integrator is not set.");
    for (size_t k = 0; k < sizeof names / sizeof names[0]; k++) {
        if (strcmp(p->integrator, names[k].name) == 0) {
            p->integrator_flag_ = names[k].flag;
            return make_success_error_status();
        }
    }
    return WRAP_RAISE_ERROR(VALUE_ERROR, "Integrator not recognized.");
}

ErrorStatus finalize_integrator_param(IntegratorParam *p)
{
    ErrorStatus e = get_integrator_flag(p);
    if (e.return_code != SUCCESS) return e;
    e = get_riemann_solver_flag(p);
    if (e.return_code != SUCCESS) return e;
    e = get_slope_limiter_flag(p);
    if (e.return_code != SUCCESS) return e;
    if (p->cfl <= 0.0) return WRAP_RAISE_ERROR(VALUE_ERROR, "CFL number must be positive.");
    if (p->cfl_initial_shrink_factor <= 0.0) return WRAP_RAISE_ERROR(VALUE_ERROR, "CFL initial shrink factor must be positive.");
    if (p->cfl_initial_shrink_num_steps <= 0)
        return WRAP_RAISE_ERROR(VALUE_ERROR, "CFL initial shrink number of steps must be positive.");
    if (p->tol <= 0.0) return WRAP_RAISE_ERROR(VALUE_ERROR, "Tolerance must be positive.");
    return make_success_error_status();
}

/* ---- storing parameters (storing.c:27-100) ----------------------------- */
StoringParam get_new_storing_param(void)
{
    StoringParam p;
    memset(&p, 0, sizeof p);
    p.is_storing = false;
    p.store_initial = false;
    p.output_dir = NULL;
    p.storing_interval = -1.0;
    p.store_count_ = 0;
    return p;
}

/* finalize_storing_param: hfv_storing.c */

/* ---- utils.c:16-21 ----------------------------------------------------- */
double hydro_get_current_time(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + (double)ts.tv_nsec / 1.0e9;
}

/* ---- dispatch and launch (integrator.c:108-143, hydro.c:29-265) -------- */
ErrorStatus integrator_launch_simulation(BoundaryConditionParam *bc, System *system, IntegratorParam *ip,
                                         StoringParam *sp, Settings *st, SimulationParam *prm,
                                         SimulationStatus *status)
{
    switch (ip->integrator_flag_) {
    case INTEGRATOR_MUSCL_HANCOCK_1D:
        return WRAP_TRACEBACK(muscl_hancock_1d(bc, system, ip, sp, st, prm, status));
    case INTEGRATOR_MUSCL_HANCOCK_2D:
        return WRAP_TRACEBACK(muscl_hancock_2d(bc, system, ip, sp, st, prm, status));
    case INTEGRATOR_GODUNOV_FIRST_ORDER_1D:
        return WRAP_TRACEBACK(godunov_first_order_1d(bc, system, ip, sp, st, prm, status));
    case INTEGRATOR_GODUNOV_FIRST_ORDER_2D:
        return WRAP_TRACEBACK(godunov_first_order_2d(bc, system, ip, sp, st, prm, status));
    case INTEGRATOR_RANDOM_CHOICE_1D:
        return WRAP_TRACEBACK(random_choice_1d(bc, system, ip, sp, st, prm, status));
    default:
        return WRAP_RAISE_ERROR(VALUE_ERROR, "Integrator flag not recognized.");
    }
}

static ErrorStatus final_check(const System *system, const IntegratorParam *ip)
{
    char msg[256];
    switch (ip->integrator_flag_) {
    case INTEGRATOR_RANDOM_CHOICE_1D:
    case INTEGRATOR_GODUNOV_FIRST_ORDER_1D:
    case INTEGRATOR_GODUNOV_FIRST_ORDER_2D:
    case INTEGRATOR_GODUNOV_FIRST_ORDER_3D:
        if (system->num_ghost_cells_side < 1)
            return WRAP_RAISE_ERROR(VALUE_ERROR, "Number of ghost cells must be at least 1.");
        break;
    case INTEGRATOR_MUSCL_HANCOCK_1D:
    case INTEGRATOR_MUSCL_HANCOCK_2D:
        if (system->num_ghost_cells_side < 2)
            return WRAP_RAISE_ERROR(VALUE_ERROR, "Number of ghost cells must be at least 2.");
        break;
    default:
        return WRAP_RAISE_ERROR(VALUE_ERROR, "Integrator flag not recognized.");
    }
    const int f = system->coord_sys_flag_;
    switch (ip->integrator_flag_) {
    case INTEGRATOR_RANDOM_CHOICE_1D:
    case INTEGRATOR_GODUNOV_FIRST_ORDER_1D:
    case INTEGRATOR_MUSCL_HANCOCK_1D:
        if (f != COORD_SYS_CARTESIAN_1D && f != COORD_SYS_CYLINDRICAL_1D && f != COORD_SYS_SPHERICAL_1D) {
            snprintf(msg, sizeof msg,
                     "Wrong coordinate system. Supported coordinate system for integrator \"%s\": \"cartesian_1d\", "
                     "\"cylindrical_1d\" and \"spherical_1d\", got: \"%s\"",
                     ip->integrator, system->coord_sys);
            return WRAP_RAISE_ERROR(VALUE_ERROR, msg);
        }
        break;
    default:
        if (f != COORD_SYS_CARTESIAN_2D) {
            snprintf(msg, sizeof msg,
                     "Wrong coordinate system. Supported coordinate system for integrator \"%s\": \"cartesian_2d\", "
                     "got: \"%s\"",
                     ip->integrator, system->coord_sys);
            return WRAP_RAISE_ERROR(VALUE_ERROR, msg);
        }
        break;
    }
    if (ip->integrator_flag_ == INTEGRATOR_RANDOM_CHOICE_1D && ip->riemann_solver_flag_ != RIEMANN_SOLVER_EXACT) {
        snprintf(msg, sizeof msg,
                 "Supported Riemann solver for random choice method 1D: \"riemann_solver_exact\", got: \"%s\"",
                 ip->riemann_solver);
        return WRAP_RAISE_ERROR(VALUE_ERROR, msg);
    }
    return make_success_error_status();
}

ErrorStatus launch_simulation(BoundaryConditionParam *bc, System *system, IntegratorParam *ip, StoringParam *sp,
                              Settings *st, SimulationParam *prm, SimulationStatus *status)
{
    ErrorStatus e = WRAP_TRACEBACK(finalize_boundary_condition_param(system, bc));
    if (e.return_code != SUCCESS) return e;
    e = WRAP_TRACEBACK(finalize_integrator_param(ip));
    if (e.return_code != SUCCESS) return e;
    e = WRAP_TRACEBACK(finalize_storing_param(sp));
    if (e.return_code != SUCCESS) return e;
    e = WRAP_TRACEBACK(final_check(system, ip));
    if (e.return_code != SUCCESS) return e;
    return WRAP_TRACEBACK(integrator_launch_simulation(bc, system, ip, sp, st, prm, status));
}
