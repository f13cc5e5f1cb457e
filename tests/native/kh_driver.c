/*
 * kh_driver.c -- a C program that uses libhydrofv_amd.so through
 * include/hydrofv/hydro.h alone, the way the reference's example drivers use
 * the reference library (examples/kelvin_helmholtz_2d/kelvin_helmholtz_2d.c:
 * system_init, the initial primitive state written into the System's
 * arrays, set_boundary_condition, convert_primitive_to_conserved,
 * launch_simulation).  Test infrastructure: tests/test_gpu_c_driver.py runs
 * it and checks the run against the oracle.
 *
 *   kh_driver N TF OUT
 *
 * Shear layer on [0,1]^2, N x N cells, 2 ghost cells, gamma 5/3, periodic,
 * MUSCL-Hancock + HLLC + MC at CFL 0.8 to t = TF.  OUT receives, as raw
 * little-endian values: int32 nx, ny, g; int64 num_steps; double t, dt;
 * double dt0 (get_time_step_2d of the initial state at the shrunk CFL);
 * then W0[4], U0[4] (after the conversion) and W[4], U[4] (after the run),
 * each (ny + 2g) x (nx + 2g) doubles, x fastest.  The per-interface API is
 * called once as well (limit_slope, solve_flux_2d, get_sound_speed), so
 * that linking this program shows those symbols are there.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "hydro.h"

static int write_fields(FILE *f, double *const *fields, size_t n)
{
    for (int k = 0; k < 4; k++)
        if (fwrite(fields[k], sizeof(double), n, f) != n) return -1;
    return 0;
}

static int fail(ErrorStatus *st)
{
    print_and_free_traceback(st);
    return 1;
}

int main(int argc, char **argv)
{
    if (argc != 4) {
        fprintf(stderr, "usage: %s N TF OUT\n", argv[0]);
        return 2;
    }
    const int n = atoi(argv[1]);
    const double tf = atof(argv[2]);
    const char *out_path = argv[3];

    System system = get_new_system_struct();
    system.coord_sys = "cartesian_2d";
    system.gamma = 5.0 / 3.0;
    system.x_min = 0.0;
    system.x_max = 1.0;
    system.y_min = 0.0;
    system.y_max = 1.0;
    system.num_cells_x = n;
    system.num_cells_y = n;
    system.num_ghost_cells_side = 2;
    ErrorStatus st = system_init(&system);
    if (st.return_code != SUCCESS) return fail(&st);

    const int g = system.num_ghost_cells_side;
    const int tx = n + 2 * g, ty = n + 2 * g;
    for (int j = g; j < g + n; j++) {
        for (int i = g; i < g + n; i++) {
            const int k = j * tx + i;
            const double x = system.mid_points_x_[i], y = system.mid_points_y_[j];
            const int inner = fabs(y - 0.5) < 0.25;
            /* a fixed, cell-dependent perturbation of amplitude 1e-3 */
            const uint32_t h = (uint32_t)(i * 73856093u) ^ (uint32_t)(j * 19349663u);
            const double r = (double)(h % 2001u) / 1000.0 - 1.0;
            system.density_[k] = inner ? 2.0 : 1.0;
            system.velocity_x_[k] = (inner ? 0.5 : -0.5) + 1e-3 * r;
            system.velocity_y_[k] = 0.01 * sin(4.0 * M_PI * x) - 1e-3 * r;
            system.pressure_[k] = 2.5;
        }
    }

    BoundaryConditionParam bc = get_new_boundary_condition_param();
    bc.boundary_condition_x_min = "periodic";
    bc.boundary_condition_x_max = "periodic";
    bc.boundary_condition_y_min = "periodic";
    bc.boundary_condition_y_max = "periodic";
    st = finalize_boundary_condition_param(&system, &bc);
    if (st.return_code != SUCCESS) return fail(&st);
    st = set_boundary_condition(&bc, &system);
    if (st.return_code != SUCCESS) return fail(&st);
    st = convert_primitive_to_conserved(&system);
    if (st.return_code != SUCCESS) return fail(&st);

    const size_t cells = (size_t)tx * ty;
    double *keep = malloc(8 * cells * sizeof(double));
    if (!keep) return 1;
    double *const W[4] = {system.density_, system.velocity_x_, system.velocity_y_, system.pressure_};
    double *const U[4] = {system.mass_, system.momentum_x_, system.momentum_y_, system.energy_};
    for (int k = 0; k < 4; k++) {
        for (size_t c = 0; c < cells; c++) {
            keep[k * cells + c] = W[k][c];
            keep[(4 + k) * cells + c] = U[k][c];
        }
    }

    IntegratorParam integrator_param = get_new_integrator_param();
    integrator_param.integrator = "muscl_hancock_2d";
    integrator_param.riemann_solver = "riemann_solver_hllc";
    integrator_param.slope_limiter = "monotonized_central";
    integrator_param.cfl = 0.8;
    st = finalize_integrator_param(&integrator_param);
    if (st.return_code != SUCCESS) return fail(&st);

    /* the per-interface API on the initial state */
    const double dt0 = get_time_step_2d(&system, integrator_param.cfl * integrator_param.cfl_initial_shrink_factor);
    double slope = 0.0, fm = 0.0, fmx = 0.0, fmy = 0.0, fe = 0.0;
    Settings settings = {.verbose = 0, .no_progress_bar = true};
    st = limit_slope(&slope, &integrator_param, 0.3, 0.1);
    if (st.return_code != SUCCESS) return fail(&st);
    st = solve_flux_2d(&integrator_param, &settings, &fm, &fmx, &fmy, &fe, system.gamma, 2.0, 0.5, 0.0, 2.5, 1.0,
                       -0.5, 0.0, 2.5);
    if (st.return_code != SUCCESS) return fail(&st);
    printf("limit_slope(0.3, 0.1) = %.17g; HLLC mass flux of the shear states = %.17g; sound speed = %.17g\n",
           slope, fm, get_sound_speed(system.gamma, 2.0, 2.5));

    StoringParam storing_param = get_new_storing_param();
    storing_param.is_storing = false;
    SimulationParam simulation_param = {.tf = tf};
    SimulationStatus status;
    st = launch_simulation(&bc, &system, &integrator_param, &storing_param, &settings, &simulation_param, &status);
    if (st.return_code != SUCCESS) return fail(&st);
    printf("steps %lld, t %.17g, dt %.17g\n", (long long)status.num_steps, status.t, status.dt);

    FILE *f = fopen(out_path, "wb");
    if (!f) return 1;
    const int32_t hdr[3] = {n, n, g};
    const int64_t steps = status.num_steps;
    const double times[3] = {status.t, status.dt, dt0};
    double *const W0[4] = {keep, keep + cells, keep + 2 * cells, keep + 3 * cells};
    double *const U0[4] = {keep + 4 * cells, keep + 5 * cells, keep + 6 * cells, keep + 7 * cells};
    int bad = fwrite(hdr, sizeof hdr, 1, f) != 1 || fwrite(&steps, sizeof steps, 1, f) != 1 ||
              fwrite(times, sizeof times, 1, f) != 1 || write_fields(f, W0, cells) || write_fields(f, U0, cells) ||
              write_fields(f, W, cells) || write_fields(f, U, cells);
    bad |= fclose(f) != 0;
    free(keep);
    free_system_memory(&system);
    return bad ? 1 : 0;
}
