// hfv_physics.cpp -- the reference's per-interface and per-cell C API as host
// functions of libhydrofv_amd.so: the slope limiter (slope_limiter.h), the
// Riemann solvers (riemann_solver.h, riemann_solver_hllc.h,
// riemann_solver_exact.h), the CFL time step and sound speed (utils.h) and
// the source terms (source_term.h), with the reference's signatures, flags,
// error codes and messages.  A driver that calls them directly links against
// this library as it links against the reference.
//
// The arithmetic is the device physics of hfv_math.hpp compiled for the
// host, with every power taken by the host C library's pow() as in the
// reference (HFV_HOST_LIBM_POW), and -ffp-contract=off: each function
// performs the reference's IEEE operation sequence and returns its bits.
#define HFV_HOST_LIBM_POW 1
#include "../device/hfv_math.hpp"

#include <stdio.h>

// The library exports only the C API: this file is built with
// -fvisibility=hidden, and the declarations of hydro.h are the exceptions.
#pragma GCC visibility push(default)
#include "hydrofv/hydro.h"
#pragma GCC visibility pop

using namespace hfv;

namespace {

constexpr int BISECTION_MAX_ITER = 5000;  // riemann_solver_exact.c: count > 5000 fails

// riemann_solver_exact.c:738-754: the warning every 500 bisection
// iterations (verbose > 0) and the failure message, for a solve whose
// bisection ran `count` iterations
void bisection_messages(int count, int verbose, ErrorStatus* err)
{
    char buf[256];
    if (verbose > 0) {
        for (int c = 500; c <= count && c <= BISECTION_MAX_ITER; c += 500) {
            snprintf(buf, sizeof buf,
                     "Bisection method did not converge after %d iterations for exact riemann solver.", c);
            WRAP_RAISE_WARNING(buf);
        }
    }
    if (err) {
        snprintf(buf, sizeof buf, "Bisection method did not converge after %d iterations for exact riemann solver.",
                 count);
        *err = WRAP_RAISE_ERROR(FAILURE, buf);
    }
}

// riemann_solver_exact.c:332-469 for the (rho, u, p) problem at x/t = speed
ErrorStatus exact_1d(double* sol_rho, double* sol_u, double* sol_p, double gamma, double rho_L, double u_L,
                     double p_L, double rho_R, double u_R, double p_R, double tol, double speed, int verbose)
{
    int count = 0;
    const Sample S = exact_sample<true>(gamma, tol, rho_L, u_L, 0.0, p_L, rho_R, u_R, 0.0, p_R, speed, &count);
    if (S.rc == ERR_EXACT_BRACKET) return WRAP_RAISE_ERROR(FAILURE, "Failed to bracket the root for bisection method.");
    if (S.rc == ERR_EXACT_BISECT) {
        ErrorStatus e;
        bisection_messages(count, verbose, &e);
        return e;
    }
    bisection_messages(count, verbose, nullptr);
    *sol_rho = S.w.r;
    *sol_u = S.w.u;
    *sol_p = S.w.p;
    return make_success_error_status();
}

ErrorStatus hllc_error()
{
    return WRAP_RAISE_ERROR(VALUE_ERROR,
                            "Invalid wave speed for the HLLC riemann solver. Possible fix: use a smaller CFL.");
}

}  // namespace

extern "C" {

// slope_limiter.c:93-115
ErrorStatus limit_slope(double* limited_slope, const IntegratorParam* integrator_param, const double slope_L,
                        const double slope_R)
{
    switch (integrator_param->slope_limiter_flag_) {
    case MINMOD: *limited_slope = limit<MINMOD>(slope_L, slope_R); break;
    case VAN_LEER: *limited_slope = limit<VAN_LEER>(slope_L, slope_R); break;
    case MC: *limited_slope = limit<MC>(slope_L, slope_R); break;
    default: return WRAP_RAISE_ERROR(VALUE_ERROR, "Limiter not recognized.");
    }
    return make_success_error_status();
}

// utils.c:23-26
double get_sound_speed(double gamma, double rho, double p) { return sound_speed(gamma, rho, p); }

// riemann_solver.c:144-152
double riemann_A_L_or_R(const double gamma, const double rho_X) { return coeff_A(gamma, rho_X); }
double riemann_B_L_or_R(const double gamma, const double p_X) { return coeff_B(gamma, p_X); }

// riemann_solver.c:154-224
double guess_p(const double gamma, const double rho_L, const double u_L, const double p_L, const double a_L,
               const double rho_R, const double u_R, const double p_R, const double a_R, const double tol)
{
    return hfv::guess_p(gamma, rho_L, u_L, p_L, a_L, rho_R, u_R, p_R, a_R, tol);
}

// riemann_solver.c:409-481
void solve_vacuum_1d(double* sol_rho, double* sol_u, double* sol_p, const double gamma, const double rho_L,
                     const double u_L, const double p_L, const double a_L, const double rho_R, const double u_R,
                     const double p_R, const double a_R, const double speed)
{
    const Prim w = vacuum(gamma, rho_L, u_L, 0.0, p_L, a_L, rho_R, u_R, 0.0, p_R, a_R, speed);
    *sol_rho = w.r;
    *sol_u = w.u;
    *sol_p = w.p;
}

// riemann_solver.c:667-733
void solve_vacuum_2d(double* sol_rho, double* sol_u, double* sol_v, double* sol_p, const double gamma,
                     const double rho_L, const double u_L, const double v_L, const double p_L, const double a_L,
                     const double rho_R, const double u_R, const double v_R, const double p_R, const double a_R,
                     const double speed)
{
    const Prim w = vacuum(gamma, rho_L, u_L, v_L, p_L, a_L, rho_R, u_R, v_R, p_R, a_R, speed);
    *sol_rho = w.r;
    *sol_u = w.u;
    *sol_v = w.v;
    *sol_p = w.p;
}

// riemann_solver_hllc.c:42-200.  On the error the fluxes are left as they
// were, as in the reference.
ErrorStatus solve_flux_hllc_1d(double* flux_mass, double* flux_momentum, double* flux_energy, const double gamma,
                               const double rho_L, const double u_L, const double p_L, const double rho_R,
                               const double u_R, const double p_R, const double tol)
{
    const double wl[3] = {rho_L, u_L, p_L}, wr[3] = {rho_R, u_R, p_R};
    double f[3];
    if (flux_hllc<3>(gamma, tol, wl, wr, f) != ERR_NONE) return hllc_error();
    *flux_mass = f[0];
    *flux_momentum = f[1];
    *flux_energy = f[2];
    return make_success_error_status();
}

// riemann_solver_hllc.c:202-376
ErrorStatus solve_flux_hllc_2d(double* flux_mass, double* flux_momentum_x, double* flux_momentum_y,
                               double* flux_energy, const double gamma, const double rho_L, const double u_L,
                               const double v_L, const double p_L, const double rho_R, const double u_R,
                               const double v_R, const double p_R, const double tol)
{
    const double wl[4] = {rho_L, u_L, v_L, p_L}, wr[4] = {rho_R, u_R, v_R, p_R};
    double f[4];
    if (flux_hllc<4>(gamma, tol, wl, wr, f) != ERR_NONE) return hllc_error();
    *flux_mass = f[0];
    *flux_momentum_x = f[1];
    *flux_momentum_y = f[2];
    *flux_energy = f[3];
    return make_success_error_status();
}

// riemann_solver_exact.c:332-469.  On an error the solution is left as it
// was (the reference leaves it unset).
ErrorStatus solve_exact_1d(double* sol_rho, double* sol_u, double* sol_p, const double gamma, const double rho_L,
                           const double u_L, const double p_L, const double rho_R, const double u_R, const double p_R,
                           const double tol, const double speed, const int verbose)
{
    return WRAP_TRACEBACK(exact_1d(sol_rho, sol_u, sol_p, gamma, rho_L, u_L, p_L, rho_R, u_R, p_R, tol, speed,
                                   verbose));
}

// riemann_solver_exact.c:471-510.  On an error the fluxes are those of a
// NaN state (the reference computes them from an unset solution).
ErrorStatus solve_flux_exact_1d(double* flux_mass, double* flux_momentum, double* flux_energy, const double gamma,
                                const double rho_L, const double u_L, const double p_L, const double rho_R,
                                const double u_R, const double p_R, const double tol, const double speed,
                                const int verbose)
{
    double r = NAN, u = NAN, p = NAN;
    ErrorStatus e =
        WRAP_TRACEBACK(exact_1d(&r, &u, &p, gamma, rho_L, u_L, p_L, rho_R, u_R, p_R, tol, speed, verbose));
    Sample S;
    S.w = Prim{r, u, 0.0, p};
    S.rc = ERR_NONE;
    double f[3];
    flux_from_sample<3>(gamma, S, f);
    *flux_mass = f[0];
    *flux_momentum = f[1];
    *flux_energy = f[2];
    return e;
}

// riemann_solver.c:47-97
ErrorStatus solve_flux_1d(IntegratorParam* integrator_param, Settings* settings, double* flux_mass,
                          double* flux_momentum, double* flux_energy, const double gamma, const double rho_L,
                          const double u_L, const double p_L, const double rho_R, const double u_R, const double p_R)
{
    switch (integrator_param->riemann_solver_flag_) {
    case EXACT:
        return WRAP_TRACEBACK(solve_flux_exact_1d(flux_mass, flux_momentum, flux_energy, gamma, rho_L, u_L, p_L,
                                                  rho_R, u_R, p_R, integrator_param->tol, 0.0, settings->verbose));
    case HLLC:
        return WRAP_TRACEBACK(solve_flux_hllc_1d(flux_mass, flux_momentum, flux_energy, gamma, rho_L, u_L, p_L,
                                                 rho_R, u_R, p_R, integrator_param->tol));
    default: return WRAP_RAISE_ERROR(VALUE_ERROR, "Riemann solver flag not recognized.");
    }
}

// riemann_solver.c:99-142.  The exact solver is accepted in 2D, as by this
// library's integrators (the reference raises "Only HLLC Riemann solver is
// available for cartesian 2D."): the normal components are the 1D exact
// flux, the transverse momentum is the mass flux times the upwind v.
ErrorStatus solve_flux_2d(IntegratorParam* integrator_param, Settings* settings, double* flux_mass,
                          double* flux_momentum_x, double* flux_momentum_y, double* flux_energy, const double gamma,
                          const double rho_L, const double u_L, const double v_L, const double p_L,
                          const double rho_R, const double u_R, const double v_R, const double p_R)
{
    switch (integrator_param->riemann_solver_flag_) {
    case EXACT: {
        int count = 0;
        const double tol = integrator_param->tol;
        const Sample S = exact_sample<true>(gamma, tol, rho_L, u_L, v_L, p_L, rho_R, u_R, v_R, p_R, 0.0, &count);
        if (S.rc == ERR_EXACT_BRACKET)
            return WRAP_RAISE_ERROR(FAILURE, "Failed to bracket the root for bisection method.");
        if (S.rc == ERR_EXACT_BISECT) {
            ErrorStatus e;
            bisection_messages(count, settings->verbose, &e);
            return e;
        }
        bisection_messages(count, settings->verbose, nullptr);
        double f[4];
        flux_from_sample<4>(gamma, S, f);
        *flux_mass = f[0];
        *flux_momentum_x = f[1];
        *flux_momentum_y = f[2];
        *flux_energy = f[3];
        return make_success_error_status();
    }
    case HLLC:
        return WRAP_TRACEBACK(solve_flux_hllc_2d(flux_mass, flux_momentum_x, flux_momentum_y, flux_energy, gamma,
                                                 rho_L, u_L, v_L, p_L, rho_R, u_R, v_R, p_R, integrator_param->tol));
    default: return WRAP_RAISE_ERROR(VALUE_ERROR, "Riemann solver flag not recognized.");
    }
}

// utils.c:28-53
double get_time_step_1d(const System* system, const double cfl)
{
    const int g = system->num_ghost_cells_side, n = system->num_cells_x;
    double smax = -INFINITY;
    for (int i = g; i < n + g; i++) {
        const double s = sound_speed(system->gamma, system->density_[i], system->pressure_[i]) +
                         fabs(system->velocity_x_[i]);
        if (s > smax) smax = s;
    }
    return cfl * system->dx_ / smax;
}

// utils.c:55-95.  The maxima do not depend on the order of the cells (a
// maximum of non-negative values, NaN never taken), so the rows are read
// in memory order where the reference walks columns.
double get_time_step_2d(const System* system, const double cfl)
{
    const int g = system->num_ghost_cells_side, nx = system->num_cells_x, ny = system->num_cells_y;
    const long long tx = nx + 2 * g;
    double sx = -INFINITY, sy = -INFINITY;
    for (int j = g; j < ny + g; j++) {
        for (int i = g; i < nx + g; i++) {
            const long long k = j * tx + i;
            const double a = sound_speed(system->gamma, system->density_[k], system->pressure_[k]);
            const double vx = a + fabs(system->velocity_x_[k]);
            const double vy = a + fabs(system->velocity_y_[k]);
            if (vx > sx) sx = vx;
            if (vy > sy) sy = vy;
        }
    }
    return cfl / (sx / system->dx_ + sy / system->dy_);
}

// source_term.c:45-183: RK4 per interior cell, then c2p and the boundary fill
ErrorStatus add_geometry_source_term(const BoundaryConditionParam* boundary_condition_param, System* system,
                                     const double dt)
{
    double alpha;
    switch (system->coord_sys_flag_) {
    case COORD_SYS_CARTESIAN_1D:
    case COORD_SYS_CARTESIAN_2D:
    case COORD_SYS_CARTESIAN_3D:
        return WRAP_RAISE_ERROR(VALUE_ERROR, "Wrong coordinate system, expected cylindrical 1D or spherical 1D.");
    case COORD_SYS_CYLINDRICAL_1D: alpha = 1.0; break;
    case COORD_SYS_SPHERICAL_1D: alpha = 2.0; break;
    default: return WRAP_RAISE_ERROR(VALUE_ERROR, "Unknown coordinate system flag.");
    }
    const int g = system->num_ghost_cells_side, n = system->num_cells_x;
    for (int i = g; i < n + g; i++) {
        double u[3] = {system->mass_[i], system->momentum_x_[i], system->energy_[i]};
        geometry_source_rk4(alpha, system->gamma, system->volume_[i], system->mid_points_x_[i], system->density_[i],
                            dt, u);
        system->mass_[i] = u[0];
        system->momentum_x_[i] = u[1];
        system->energy_[i] = u[2];
    }
    ErrorStatus e = WRAP_TRACEBACK(convert_conserved_to_primitive(system));
    if (e.return_code != SUCCESS) return e;
    return WRAP_TRACEBACK(set_boundary_condition(boundary_condition_param, system));
}

// source_term.c:185-259: nothing when gravity is 0; else the RK4 gravity
// term per interior cell, then c2p and the boundary fill
ErrorStatus add_gravity_source_term_2d(const BoundaryConditionParam* boundary_condition_param, System* system,
                                       const double dt)
{
    if (system->gravity == 0.0) return make_success_error_status();
    const int g = system->num_ghost_cells_side, nx = system->num_cells_x, ny = system->num_cells_y;
    const long long tx = nx + 2 * g;
    for (int j = g; j < ny + g; j++) {
        for (int i = g; i < nx + g; i++) {
            const long long k = j * tx + i;
            gravity_source(dt, system->gravity, system->mass_[k], &system->momentum_y_[k], &system->energy_[k]);
        }
    }
    ErrorStatus e = WRAP_TRACEBACK(convert_conserved_to_primitive(system));
    if (e.return_code != SUCCESS) return e;
    return WRAP_TRACEBACK(set_boundary_condition(boundary_condition_param, system));
}

}  // extern "C"
