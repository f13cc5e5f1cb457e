/*
 * hydro.h -- hydroFV's C API as exported by libhydrofv_amd.so.
 *
 * Drop-in for the reference's include tree (src/hydro.h and the headers it
 * pulls in): the same type names, field order and layouts, flag values,
 * function names and signatures, so a driver written against the reference
 * (e.g. examples/kelvin_helmholtz_2d/kelvin_helmholtz_2d.c) compiles and links
 * against this library unchanged.  muscl_hancock_1d / muscl_hancock_2d,
 * godunov_first_order_1d / godunov_first_order_2d and random_choice_1d run
 * on the MI355X; the set-up helpers (system_init, conversions, boundary fill of
 * the initial state) operate on the host arrays of System as in the reference.
 *
 * Each declaration cites the reference interface it replaces.
 */
#ifndef HYDROFV_HYDRO_H
#define HYDROFV_HYDRO_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- common.h:16-69 --------------------------------------------------- */
#define IN_FILE static
typedef int64_t int64;

typedef struct ErrorStatus {
    int return_code;
    char *traceback;
    int traceback_code_;
} ErrorStatus;

typedef struct IntegratorParam {
    char *integrator;
    char *riemann_solver;
    char *slope_limiter;
    double cfl;
    double cfl_initial_shrink_factor;
    int cfl_initial_shrink_num_steps;
    double tol;
    int integrator_flag_;
    int riemann_solver_flag_;
    int slope_limiter_flag_;
} IntegratorParam;

typedef struct StoringParam {
    char *output_dir;
    bool is_storing;
    bool store_initial;
    double storing_interval;
    int storing_method_flag_;
    int store_count_;
} StoringParam;

typedef struct Settings {
    int verbose;
    bool no_progress_bar;
} Settings;

typedef struct SimulationParam {
    double tf;
} SimulationParam;

typedef struct SimulationStatus {
    int64 num_steps;
    double t;
    double dt;
} SimulationStatus;

/* ---- error.h:15-132 ---------------------------------------------------- */
#define UNKNOWN_ERROR -1
#define SUCCESS 0
#define FAILURE 1
#define VALUE_ERROR 2
#define POINTER_ERROR 3
#define MEMORY_ERROR 4
#define OS_ERROR 5
#define NOT_IMPLEMENTED_ERROR 6

#define TRACEBACK_NOT_INITIALIZED -1
#define TRACEBACK_SUCCESS 0
#define TRACEBACK_MALLOC_FAILED 1
#define TRACEBACK_TRUNCATED 2
#define TRACEBACK_SNPRINTF_FAILED 3

#define WRAP_RAISE_WARNING(error_msg) raise_warning(error_msg, __FILE__, __LINE__, __func__)
#define WRAP_RAISE_ERROR(error_code, error_msg) raise_error(error_code, error_msg, __FILE__, __LINE__, __func__)
#define WRAP_TRACEBACK(function_call) traceback(function_call, #function_call, __FILE__, __LINE__, __func__)

ErrorStatus make_success_error_status(void);                       /* error.h:67 */
void raise_warning(const char *warning_msg, const char *warning_file, const int warning_line,
                   const char *warning_func);                        /* error.h:77-82 */
ErrorStatus raise_error(const int error_code, const char *error_msg, const char *error_file,
                        const int error_line, const char *error_func); /* error.h:95-101 */
ErrorStatus traceback(ErrorStatus error_status, const char *function_call_source_code,
                      const char *error_file, const int error_line,
                      const char *error_func);                       /* error.h:114-120 */
void free_traceback(ErrorStatus *error_status);                      /* error.h:125 */
void print_and_free_traceback(ErrorStatus *error_status);            /* error.h:132 */

/* ---- system.h:16-202 --------------------------------------------------- */
#define COORD_SYS_CARTESIAN_1D 1
#define COORD_SYS_CARTESIAN_2D 2
#define COORD_SYS_CARTESIAN_3D 3
#define COORD_SYS_CYLINDRICAL_1D 4
#define COORD_SYS_SPHERICAL_1D 5

typedef struct System {
    const char *coord_sys;
    double gamma;
    double gravity;
    double x_min;
    double x_max;
    double y_min;
    double y_max;
    double z_min;
    double z_max;
    int num_cells_x;
    int num_cells_y;
    int num_cells_z;
    int num_ghost_cells_side;

    int coord_sys_flag_;
    double dx_;
    double dy_;
    double dz_;
    double *density_;
    double *velocity_x_;
    double *velocity_y_;
    double *velocity_z_;
    double *pressure_;
    double *mass_;
    double *momentum_x_;
    double *momentum_y_;
    double *momentum_z_;
    double *energy_;
    double *mid_points_x_;
    double *mid_points_y_;
    double *mid_points_z_;
    double *surface_area_x_;
    double *surface_area_y_;
    double *surface_area_z_;
    double *volume_;
} System;

System get_new_system_struct(void);                                  /* system.h:138 */
ErrorStatus system_init(System *system);                             /* system.h:147 */
void free_system_memory(System *system);                             /* system.h:154 */
void convert_conserved_to_primitive_1d(const int num_cells_x, const int num_ghost_cells_side, const double gamma,
                                       const double *volume, const double *mass, const double *momentum_x,
                                       const double *energy, double *density, double *velocity_x,
                                       double *pressure);            /* system.h:156-167 */
void convert_conserved_to_primitive_2d(const int num_cells_x, const int num_cells_y, const int num_ghost_cells_side,
                                       const double gamma, const double *volume, const double *mass,
                                       const double *momentum_x, const double *momentum_y, const double *energy,
                                       double *density, double *velocity_x, double *velocity_y,
                                       double *pressure);            /* system.h:169-183 */
ErrorStatus convert_conserved_to_primitive(System *system);          /* system.h:192 */
ErrorStatus convert_primitive_to_conserved(System *system);          /* system.h:202 */

/* ---- boundary.h:17-108 ------------------------------------------------- */
typedef void (*set_boundary_condition_1d_func)(double *density, double *velocity_x, double *pressure,
                                               const int num_ghost_cells_side, const int num_cells_x);
typedef void (*set_boundary_condition_cartesian_2d_func)(double *density, double *velocity_x, double *velocity_y,
                                                         double *pressure, const int num_ghost_cells_side,
                                                         const int num_cells_x, const int num_cells_y);

typedef struct BoundaryConditionParam {
    const char *boundary_condition_x_min;
    const char *boundary_condition_x_max;
    const char *boundary_condition_y_min;
    const char *boundary_condition_y_max;
    const char *boundary_condition_z_min;
    const char *boundary_condition_z_max;
    set_boundary_condition_1d_func set_boundary_condition_1d_x_min;
    set_boundary_condition_1d_func set_boundary_condition_1d_x_max;
    set_boundary_condition_cartesian_2d_func set_boundary_condition_cartesian_2d_x_min;
    set_boundary_condition_cartesian_2d_func set_boundary_condition_cartesian_2d_x_max;
    set_boundary_condition_cartesian_2d_func set_boundary_condition_cartesian_2d_y_min;
    set_boundary_condition_cartesian_2d_func set_boundary_condition_cartesian_2d_y_max;

    int boundary_condition_flag_x_min_;
    int boundary_condition_flag_x_max_;
    int boundary_condition_flag_y_min_;
    int boundary_condition_flag_y_max_;
    int boundary_condition_flag_z_min_;
    int boundary_condition_flag_z_max_;
} BoundaryConditionParam;

#define BOUNDARY_CONDITION_NONE 1
#define BOUNDARY_CONDITION_REFLECTIVE 2
#define BOUNDARY_CONDITION_TRANSMISSIVE 3
#define BOUNDARY_CONDITION_PERIODIC 4
#define BOUNDARY_CONDITION_CUSTOM 5

BoundaryConditionParam get_new_boundary_condition_param(void);      /* boundary.h:70 */
ErrorStatus finalize_boundary_condition_param(const System *system,
                                              BoundaryConditionParam *boundary_condition_param); /* :72-75 */
ErrorStatus set_boundary_condition(const BoundaryConditionParam *boundary_condition_param,
                                   System *system);                  /* boundary.h:85-88 */
ErrorStatus set_boundary_condition_1d(const BoundaryConditionParam *boundary_condition_param, double *density,
                                      double *velocity_x, double *pressure, const int num_ghost_cells_side,
                                      const int num_cells_x);        /* boundary.h:90-97 */
ErrorStatus set_boundary_condition_cartesian_2d(const BoundaryConditionParam *boundary_condition_param,
                                                double *density, double *velocity_x, double *velocity_y,
                                                double *pressure, const int num_ghost_cells_side,
                                                const int num_cells_x, const int num_cells_y); /* :99-108 */

/* ---- slope_limiter.h:14-47, riemann_solver.h:23-33 --------------------- */
#define SLOPE_LIMITER_MINMOD 1
#define SLOPE_LIMITER_VAN_LEER 2
#define SLOPE_LIMITER_MONOTONIZED_CENTRAL 3
#define RIEMANN_SOLVER_EXACT 1
#define RIEMANN_SOLVER_HLLC 2

ErrorStatus get_slope_limiter_flag(IntegratorParam *integrator_param);   /* slope_limiter.h:26-28 */
ErrorStatus get_riemann_solver_flag(IntegratorParam *integrator_param);  /* riemann_solver.h:33 */

/* The per-interface and per-cell functions, on the host with the
 * reference's arithmetic (the integrators run their device forms). */
ErrorStatus limit_slope(double *limited_slope, const IntegratorParam *integrator_param, const double slope_L,
                        const double slope_R);                       /* slope_limiter.h:42-47 */
/* riemann_solver.h:54-85.  solve_flux_2d also takes the exact solver, as the
 * integrators here do (the reference raises "Only HLLC Riemann solver is
 * available for cartesian 2D."). */
ErrorStatus solve_flux_1d(IntegratorParam *integrator_param, Settings *settings, double *flux_mass,
                          double *flux_momentum, double *flux_energy, const double gamma, const double rho_L,
                          const double u_L, const double p_L, const double rho_R, const double u_R,
                          const double p_R);
ErrorStatus solve_flux_2d(IntegratorParam *integrator_param, Settings *settings, double *flux_mass,
                          double *flux_momentum_x, double *flux_momentum_y, double *flux_energy, const double gamma,
                          const double rho_L, const double u_L, const double v_L, const double p_L,
                          const double rho_R, const double u_R, const double v_R, const double p_R);
double riemann_A_L_or_R(const double gamma, const double rho_X);      /* riemann_solver.h:95 */
double riemann_B_L_or_R(const double gamma, const double p_X);        /* riemann_solver.h:105 */
double guess_p(const double gamma, const double rho_L, const double u_L, const double p_L, const double a_L,
               const double rho_R, const double u_R, const double p_R, const double a_R,
               const double tol);                                    /* riemann_solver.h:123-134 */
void solve_vacuum_1d(double *sol_rho, double *sol_u, double *sol_p, const double gamma, const double rho_L,
                     const double u_L, const double p_L, const double a_L, const double rho_R, const double u_R,
                     const double p_R, const double a_R, const double speed); /* riemann_solver.h:153-167 */
void solve_vacuum_2d(double *sol_rho, double *sol_u, double *sol_v, double *sol_p, const double gamma,
                     const double rho_L, const double u_L, const double v_L, const double p_L, const double a_L,
                     const double rho_R, const double u_R, const double v_R, const double p_R, const double a_R,
                     const double speed);                            /* riemann_solver.h:169-186 */
ErrorStatus solve_flux_hllc_1d(double *flux_mass, double *flux_momentum, double *flux_energy, const double gamma,
                               const double rho_L, const double u_L, const double p_L, const double rho_R,
                               const double u_R, const double p_R,
                               const double tol);                    /* riemann_solver_hllc.h:32-44 */
ErrorStatus solve_flux_hllc_2d(double *flux_mass, double *flux_momentum_x, double *flux_momentum_y,
                               double *flux_energy, const double gamma, const double rho_L, const double u_L,
                               const double v_L, const double p_L, const double rho_R, const double u_R,
                               const double v_R, const double p_R,
                               const double tol);                    /* riemann_solver_hllc.h:46-61 */
ErrorStatus solve_exact_1d(double *sol_rho, double *sol_u, double *sol_p, const double gamma, const double rho_L,
                           const double u_L, const double p_L, const double rho_R, const double u_R,
                           const double p_R, const double tol, const double speed,
                           const int verbose);                       /* riemann_solver_exact.h:37-51 */
ErrorStatus solve_flux_exact_1d(double *flux_mass, double *flux_momentum, double *flux_energy, const double gamma,
                                const double rho_L, const double u_L, const double p_L, const double rho_R,
                                const double u_R, const double p_R, const double tol, const double speed,
                                const int verbose);                  /* riemann_solver_exact.h:73-87 */

/* ---- integrator.h:17-176 ----------------------------------------------- */
#define INTEGRATOR_RANDOM_CHOICE_1D 1
#define INTEGRATOR_GODUNOV_FIRST_ORDER_1D 2
#define INTEGRATOR_GODUNOV_FIRST_ORDER_2D 3
#define INTEGRATOR_GODUNOV_FIRST_ORDER_3D 4
#define INTEGRATOR_MUSCL_HANCOCK_1D 5
#define INTEGRATOR_MUSCL_HANCOCK_2D 6

IntegratorParam get_new_integrator_param(void);                      /* integrator.h:29 */
ErrorStatus finalize_integrator_param(IntegratorParam *integrator_param); /* integrator.h:38 */

#define HFV_INTEGRATOR_ARGS                                                                     \
    BoundaryConditionParam *boundary_condition_param, System *system, IntegratorParam *integrator_param, \
        StoringParam *storing_param, Settings *settings, SimulationParam *simulation_param,     \
        SimulationStatus *simulation_status

ErrorStatus integrator_launch_simulation(HFV_INTEGRATOR_ARGS);       /* integrator.h:53-61 */
/* MUSCL-Hancock on the GPU: integrator.h:145-153 (1D), 168-176 (2D). */
ErrorStatus muscl_hancock_1d(HFV_INTEGRATOR_ARGS);
ErrorStatus muscl_hancock_2d(HFV_INTEGRATOR_ARGS);
/* First-order Godunov on the GPU: integrator.h:99-107 (1D), 122-130 (2D). */
ErrorStatus godunov_first_order_1d(HFV_INTEGRATOR_ARGS);
ErrorStatus godunov_first_order_2d(HFV_INTEGRATOR_ARGS);
/* Random choice (Glimm) on the GPU: integrator.h:76-84. */
ErrorStatus random_choice_1d(HFV_INTEGRATOR_ARGS);

/* Extension (no reference counterpart): the same loops with a step limit
 * (max_steps >= 0) and/or a constant dt (fixed_dt > 0); muscl_hancock_{1d,2d}
 * godunov_first_order_{1d,2d} and random_choice_1d call these with
 * (-1, 0.0).  Used for fixed-length benchmark runs and for
 * fixed-dt forward solves. */
ErrorStatus hfv_muscl_hancock_1d_run(HFV_INTEGRATOR_ARGS, long long max_steps, double fixed_dt);
ErrorStatus hfv_muscl_hancock_2d_run(HFV_INTEGRATOR_ARGS, long long max_steps, double fixed_dt);
ErrorStatus hfv_godunov_first_order_1d_run(HFV_INTEGRATOR_ARGS, long long max_steps, double fixed_dt);
ErrorStatus hfv_godunov_first_order_2d_run(HFV_INTEGRATOR_ARGS, long long max_steps, double fixed_dt);
ErrorStatus hfv_random_choice_1d_run(HFV_INTEGRATOR_ARGS, long long max_steps, double fixed_dt);

/* Extension (no reference counterpart): a device form of a custom 2D hook
 * (boundary.h:25-33; the device hook type is the one of hfv_gpu.h).  While
 * `host_hook` is registered, muscl_hancock_2d and godunov_first_order_2d
 * fill the sides that name it with `device_hook` on the device state,
 * instead of copying the state to the host around every step.  A run with a
 * custom side whose hook is not registered keeps the host path.  Registering
 * a hook again replaces its entry; device_hook NULL removes it.  At most 64
 * entries; returns 0, or -1 for a NULL host_hook or a full table. */
#ifndef HFV_DEVICE_BC_2D_FUNC_DEFINED
#define HFV_DEVICE_BC_2D_FUNC_DEFINED
typedef void (*hfv_device_bc_2d_func)(double *density, double *velocity_x, double *velocity_y, double *pressure,
                                      int num_ghost_cells_side, int num_cells_x, int num_cells_y, long long pitch,
                                      void *stream, void *user);
#endif
int hfv_register_device_bc_2d(set_boundary_condition_cartesian_2d_func host_hook,
                              hfv_device_bc_2d_func device_hook, void *user);

/* ---- storing.h:21-49 --------------------------------------------------- */
StoringParam get_new_storing_param(void);                            /* storing.h:21 */
ErrorStatus finalize_storing_param(StoringParam *storing_param);     /* storing.h:30 */
/* storing.h:43-49: <output_dir>snapshot_<store_count_>.h5 in the reference's
 * HDF5 layout; the HDF5 C library is loaded at the first call. */
ErrorStatus store_snapshot(const BoundaryConditionParam *boundary_condition_param, const System *system,
                           const IntegratorParam *integrator_param, const SimulationStatus *simulation_status,
                           StoringParam *storing_param);

/* ---- utils.h:25-69 ----------------------------------------------------- */
double hydro_get_current_time(void);                                 /* utils.h:25 */
double get_sound_speed(double gamma, double rho, double p);          /* utils.h:32 */
double get_time_step_1d(const System *system, const double cfl);     /* utils.h:48-51 */
double get_time_step_2d(const System *system, const double cfl);     /* utils.h:66-69 */

/* ---- source_term.h:21-31 (host; the integrators apply them on the device) */
ErrorStatus add_geometry_source_term(const BoundaryConditionParam *boundary_condition_param, System *system,
                                     const double dt);               /* source_term.h:21-25 */
ErrorStatus add_gravity_source_term_2d(const BoundaryConditionParam *boundary_condition_param, System *system,
                                       const double dt);             /* source_term.h:27-31 */

/* ---- hydro.h:53-61 ----------------------------------------------------- */
ErrorStatus launch_simulation(HFV_INTEGRATOR_ARGS);

#ifdef __cplusplus
}
#endif

#endif
