/*
 * hfv_system.c -- the System container on the host: defaults, allocation,
 * geometry, primitive <-> conserved conversion and the ghost-cell fill of
 * the initial state (the interfaces of the reference's system.h and
 * boundary.h).  These run once while a driver sets up its initial
 * condition; the time loop itself runs on the device (hfv_integrator.c).
 *
 * The geometry and conversion formulas and their operand order are those of
 * the reference (system.c:291-529, 735-917) because the device integrator's
 * results are compared bit for bit against runs that start from them.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hydrofv/hydro.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

System get_new_system_struct(void)
{
    System s;
    memset(&s, 0, sizeof s);
    s.coord_sys = NULL;
    s.gamma = -1.0;
    s.gravity = 0.0;
    s.x_min = s.x_max = s.y_min = s.y_max = s.z_min = s.z_max = -1.0;
    s.num_cells_x = s.num_cells_y = s.num_cells_z = -1;
    s.num_ghost_cells_side = -1;
    s.coord_sys_flag_ = -1;
    s.dx_ = s.dy_ = s.dz_ = -1.0;
    return s;
}

static ErrorStatus coord_flag(System *s)
{
    static const struct { const char *name; int flag; } names[] = {
        {"cartesian_1d", COORD_SYS_CARTESIAN_1D}, {"cartesian_2d", COORD_SYS_CARTESIAN_2D},
        {"cartesian_3d", COORD_SYS_CARTESIAN_3D}, {"cylindrical_1d", COORD_SYS_CYLINDRICAL_1D},
        {"spherical_1d", COORD_SYS_SPHERICAL_1D},
    };
    if (!s->coord_sys) return WRAP_RAISE_ERROR(POINTER_ERROR, "Coordinate system pointer is NULL.");
    for (size_t k = 0; k < sizeof names / sizeof names[0]; k++) {
        if (strcmp(s->coord_sys, names[k].name) == 0) {
            s->coord_sys_flag_ = names[k].flag;
            return make_success_error_status();
        }
    }
    return WRAP_RAISE_ERROR(VALUE_ERROR, "Coordinate system not recognized.");
}

static ErrorStatus check_input(const System *s)
{
    char msg[256];
    if (s->gamma < 1.0) {
        snprintf(msg, sizeof msg, "Gamma must be greater than 1. Got: %.3g", s->gamma);
        return WRAP_RAISE_ERROR(VALUE_ERROR, msg);
    }
    const void *arrays[] = {s->density_, s->velocity_x_, s->velocity_y_, s->velocity_z_, s->pressure_, s->mass_,
                            s->momentum_x_, s->momentum_y_, s->momentum_z_, s->energy_, s->mid_points_x_,
                            s->mid_points_y_, s->mid_points_z_, s->surface_area_x_, s->surface_area_y_,
                            s->surface_area_z_, s->volume_};
    for (size_t k = 0; k < sizeof arrays / sizeof arrays[0]; k++) {
        if (arrays[k]) return WRAP_RAISE_ERROR(POINTER_ERROR, "System array pointers must be NULL before system_init.");
    }
    if (s->num_ghost_cells_side < 1) {
        snprintf(msg, sizeof msg, "Number of ghost cells per side must be at least 1. Got: %d", s->num_ghost_cells_side);
        return WRAP_RAISE_ERROR(VALUE_ERROR, msg);
    }
    if (s->coord_sys_flag_ == COORD_SYS_CARTESIAN_3D) {
        return WRAP_RAISE_ERROR(NOT_IMPLEMENTED_ERROR, "Cartesian 3D systems are not supported.");
    }
    if (s->coord_sys_flag_ == COORD_SYS_CARTESIAN_2D) {
        if (s->num_cells_y < 1) {
            snprintf(msg, sizeof msg, "Number of cells in y-direction must be positive. Got: %d", s->num_cells_y);
            return WRAP_RAISE_ERROR(VALUE_ERROR, msg);
        }
        if (s->y_min >= s->y_max) {
            snprintf(msg, sizeof msg, "y_min must be less than y_max. Got: y_min = %.3g, y_max = %.3g", s->y_min, s->y_max);
            return WRAP_RAISE_ERROR(VALUE_ERROR, msg);
        }
    }
    if (s->num_cells_x < 1) {
        snprintf(msg, sizeof msg, "Number of cells in x-direction must be positive. Got: %d", s->num_cells_x);
        return WRAP_RAISE_ERROR(VALUE_ERROR, msg);
    }
    if (s->x_min >= s->x_max) {
        snprintf(msg, sizeof msg, "x_min must be less than x_max. Got: x_min = %.3g, x_max = %.3g", s->x_min, s->x_max);
        return WRAP_RAISE_ERROR(VALUE_ERROR, msg);
    }
    return make_success_error_status();
}

void free_system_memory(System *s)
{
    double **arrays[] = {&s->density_, &s->velocity_x_, &s->velocity_y_, &s->velocity_z_, &s->pressure_, &s->mass_,
                         &s->momentum_x_, &s->momentum_y_, &s->momentum_z_, &s->energy_, &s->mid_points_x_,
                         &s->mid_points_y_, &s->mid_points_z_, &s->surface_area_x_, &s->surface_area_y_,
                         &s->surface_area_z_, &s->volume_};
    for (size_t k = 0; k < sizeof arrays / sizeof arrays[0]; k++) {
        free(*arrays[k]);
        *arrays[k] = NULL;
    }
}

/* geometry with the ghost cells inside [x_min, x_max] (system.c:291-529) */
static void init_geometry(System *s)
{
    const int g = s->num_ghost_cells_side;
    const int tx = s->num_cells_x + 2 * g;
    const double hx = (s->x_max - s->x_min) / (s->num_cells_x + 2 * g);
    s->dx_ = hx;
    for (int i = 0; i < tx; i++) s->mid_points_x_[i] = s->x_min + (i + 0.5) * hx;
    switch (s->coord_sys_flag_) {
    case COORD_SYS_CARTESIAN_2D: {
        const int ty = s->num_cells_y + 2 * g;
        const double hy = (s->y_max - s->y_min) / (s->num_cells_y + 2 * g);
        s->dy_ = hy;
        for (int j = 0; j < ty; j++) s->mid_points_y_[j] = s->y_min + (j + 0.5) * hy;
        for (int i = 0; i <= tx; i++) s->surface_area_x_[i] = hy;
        for (int j = 0; j <= ty; j++) s->surface_area_y_[j] = hx;
        const double v = hx * hy;
        for (long k = 0; k < (long)tx * ty; k++) s->volume_[k] = v;
        break;
    }
    case COORD_SYS_CARTESIAN_1D:
        for (int i = 0; i <= tx; i++) s->surface_area_x_[i] = 1.0;
        for (int i = 0; i < tx; i++) s->volume_[i] = hx;
        break;
    case COORD_SYS_CYLINDRICAL_1D:
        for (int i = 0; i <= tx; i++) s->surface_area_x_[i] = 2.0 * M_PI * (s->x_min + i * hx);
        for (int i = 0; i < tx; i++) {
            const double r0 = s->x_min + i * hx, r1 = s->x_min + (i + 1) * hx;
            s->volume_[i] = M_PI * (r1 * r1 - r0 * r0);
        }
        break;
    case COORD_SYS_SPHERICAL_1D:
        for (int i = 0; i <= tx; i++) {
            const double r = s->x_min + i * hx;
            s->surface_area_x_[i] = 4.0 * M_PI * r * r;
        }
        for (int i = 0; i < tx; i++) {
            const double r0 = s->x_min + i * hx, r1 = s->x_min + (i + 1) * hx;
            s->volume_[i] = (4.0 / 3.0) * M_PI * (r1 * r1 * r1 - r0 * r0 * r0);
        }
        break;
    default:
        break;
    }
}

ErrorStatus system_init(System *s)
{
    ErrorStatus e = WRAP_TRACEBACK(coord_flag(s));
    if (e.return_code != SUCCESS) return e;
    e = WRAP_TRACEBACK(check_input(s));
    if (e.return_code != SUCCESS) return e;

    const int g = s->num_ghost_cells_side;
    const int two_d = (s->coord_sys_flag_ == COORD_SYS_CARTESIAN_2D);
    const size_t tx = (size_t)s->num_cells_x + 2 * (size_t)g;
    const size_t ty = two_d ? (size_t)s->num_cells_y + 2 * (size_t)g : 1;
    const size_t n = tx * ty;

    s->density_ = calloc(n, sizeof(double));
    s->velocity_x_ = calloc(n, sizeof(double));
    s->pressure_ = calloc(n, sizeof(double));
    s->mass_ = calloc(n, sizeof(double));
    s->momentum_x_ = calloc(n, sizeof(double));
    s->energy_ = calloc(n, sizeof(double));
    s->volume_ = malloc(n * sizeof(double));
    s->mid_points_x_ = malloc(tx * sizeof(double));
    s->surface_area_x_ = malloc((tx + 1) * sizeof(double));
    int ok = s->density_ && s->velocity_x_ && s->pressure_ && s->mass_ && s->momentum_x_ && s->energy_ &&
             s->volume_ && s->mid_points_x_ && s->surface_area_x_;
    if (two_d) {
        s->velocity_y_ = calloc(n, sizeof(double));
        s->momentum_y_ = calloc(n, sizeof(double));
        s->mid_points_y_ = malloc(ty * sizeof(double));
        s->surface_area_y_ = malloc((ty + 1) * sizeof(double));
        ok = ok && s->velocity_y_ && s->momentum_y_ && s->mid_points_y_ && s->surface_area_y_;
    }
    if (!ok) {
        free_system_memory(s);
        return WRAP_RAISE_ERROR(MEMORY_ERROR, "Memory allocation failed.");
    }
    init_geometry(s);
    return make_success_error_status();
}

/* system.c:735-761 */
void convert_conserved_to_primitive_1d(const int num_cells_x, const int num_ghost_cells_side, const double gamma,
                                       const double *volume, const double *mass, const double *momentum_x,
                                       const double *energy, double *density, double *velocity_x,
                                       double *pressure)
{
    const int tx = num_cells_x + 2 * num_ghost_cells_side;
    for (int i = 0; i < tx; i++) {
        const double m = mass[i], vol = volume[i];
        density[i] = m / vol;
        velocity_x[i] = momentum_x[i] / m;
        pressure[i] = (gamma - 1.0) * (energy[i] - 0.5 * m * velocity_x[i] * velocity_x[i]) / vol;
    }
}

/* system.c:763-802 */
void convert_conserved_to_primitive_2d(const int num_cells_x, const int num_cells_y, const int num_ghost_cells_side,
                                       const double gamma, const double *volume, const double *mass,
                                       const double *momentum_x, const double *momentum_y, const double *energy,
                                       double *density, double *velocity_x, double *velocity_y, double *pressure)
{
    const long n = (long)(num_cells_x + 2 * num_ghost_cells_side) * (num_cells_y + 2 * num_ghost_cells_side);
    for (long k = 0; k < n; k++) {
        const double m = mass[k], vol = volume[k];
        density[k] = m / vol;
        velocity_x[k] = momentum_x[k] / m;
        velocity_y[k] = momentum_y[k] / m;
        pressure[k] = (gamma - 1.0) * (energy[k] - 0.5 * m * (velocity_x[k] * velocity_x[k] + velocity_y[k] * velocity_y[k])) / vol;
    }
}

ErrorStatus convert_conserved_to_primitive(System *s)
{
    switch (s->coord_sys_flag_) {
    case COORD_SYS_CARTESIAN_1D:
    case COORD_SYS_CYLINDRICAL_1D:
    case COORD_SYS_SPHERICAL_1D:
        convert_conserved_to_primitive_1d(s->num_cells_x, s->num_ghost_cells_side, s->gamma, s->volume_, s->mass_,
                                          s->momentum_x_, s->energy_, s->density_, s->velocity_x_, s->pressure_);
        return make_success_error_status();
    case COORD_SYS_CARTESIAN_2D:
        convert_conserved_to_primitive_2d(s->num_cells_x, s->num_cells_y, s->num_ghost_cells_side, s->gamma, s->volume_,
                                          s->mass_, s->momentum_x_, s->momentum_y_, s->energy_, s->density_,
                                          s->velocity_x_, s->velocity_y_, s->pressure_);
        return make_success_error_status();
    case COORD_SYS_CARTESIAN_3D:
        return WRAP_RAISE_ERROR(NOT_IMPLEMENTED_ERROR, "");
    default:
        return WRAP_RAISE_ERROR(VALUE_ERROR, "Coordinate system flag not recognized.");
    }
}

/* system.c:849-917 */
ErrorStatus convert_primitive_to_conserved(System *s)
{
    const double gm = s->gamma;
    const int g = s->num_ghost_cells_side;
    switch (s->coord_sys_flag_) {
    case COORD_SYS_CARTESIAN_1D:
    case COORD_SYS_CYLINDRICAL_1D:
    case COORD_SYS_SPHERICAL_1D: {
        const int tx = s->num_cells_x + 2 * g;
        for (int i = 0; i < tx; i++) {
            const double r = s->density_[i], u = s->velocity_x_[i], p = s->pressure_[i], vol = s->volume_[i];
            s->mass_[i] = r * vol;
            s->momentum_x_[i] = s->mass_[i] * u;
            s->energy_[i] = s->volume_[i] * (0.5 * r * u * u + p / (gm - 1.0));
        }
        return make_success_error_status();
    }
    case COORD_SYS_CARTESIAN_2D: {
        const long n = (long)(s->num_cells_x + 2 * g) * (s->num_cells_y + 2 * g);
        for (long k = 0; k < n; k++) {
            const double r = s->density_[k], u = s->velocity_x_[k], v = s->velocity_y_[k], p = s->pressure_[k];
            const double vol = s->volume_[k];
            s->mass_[k] = r * vol;
            s->momentum_x_[k] = s->mass_[k] * u;
            s->momentum_y_[k] = s->mass_[k] * v;
            s->energy_[k] = vol * (0.5 * r * (u * u + v * v) + p / (gm - 1.0));
        }
        return make_success_error_status();
    }
    case COORD_SYS_CARTESIAN_3D:
        return WRAP_RAISE_ERROR(NOT_IMPLEMENTED_ERROR, "");
    default:
        return WRAP_RAISE_ERROR(VALUE_ERROR, "Coordinate system flag not recognized.");
    }
}

/* ---- ghost-cell fill on primitives (boundary.c:411-777) ---------------- */
/* One line of cells (a row or a column) at base + k*stride; 'nf' fields,
 * field 'normal' is negated by a reflective side.  Sides are processed low
 * then high, layer by layer, as in the reference. */
static ErrorStatus fill_line(int lo, int hi, double *const *w, int nf, int normal, long base, long stride, int n,
                             int g)
{
#define AT(f, k) w[f][base + (long)(k) * stride]
    if (lo == BOUNDARY_CONDITION_REFLECTIVE || lo == BOUNDARY_CONDITION_TRANSMISSIVE) {
        for (int i = 0; i < g; i++)
            for (int f = 0; f < nf; f++) {
                const double v = AT(f, g + i);
                AT(f, g - 1 - i) = (lo == BOUNDARY_CONDITION_REFLECTIVE && f == normal) ? -v : v;
            }
    } else if (lo != BOUNDARY_CONDITION_NONE && lo != BOUNDARY_CONDITION_PERIODIC && lo != BOUNDARY_CONDITION_CUSTOM) {
        return WRAP_RAISE_ERROR(VALUE_ERROR, "Boundary condition flag not recognized for the lower side.");
    }
    if (hi == BOUNDARY_CONDITION_REFLECTIVE || hi == BOUNDARY_CONDITION_TRANSMISSIVE) {
        for (int i = 0; i < g; i++)
            for (int f = 0; f < nf; f++) {
                const double v = AT(f, g + n - 1 - i);
                AT(f, g + n + i) = (hi == BOUNDARY_CONDITION_REFLECTIVE && f == normal) ? -v : v;
            }
    } else if (hi == BOUNDARY_CONDITION_PERIODIC) {
        for (int i = 0; i < g; i++)
            for (int f = 0; f < nf; f++) {
                AT(f, g + n + i) = AT(f, g + i);
                AT(f, g - 1 - i) = AT(f, g + n - 1 - i);
            }
    } else if (hi != BOUNDARY_CONDITION_NONE && hi != BOUNDARY_CONDITION_CUSTOM) {
        return WRAP_RAISE_ERROR(VALUE_ERROR, "Boundary condition flag not recognized for the upper side.");
    }
#undef AT
    return make_success_error_status();
}

ErrorStatus set_boundary_condition_1d(const BoundaryConditionParam *bc, double *density, double *velocity_x,
                                      double *pressure, const int num_ghost_cells_side, const int num_cells_x)
{
    double *w[3] = {density, velocity_x, pressure};
    if (bc->boundary_condition_flag_x_min_ == BOUNDARY_CONDITION_CUSTOM) {
        bc->set_boundary_condition_1d_x_min(density, velocity_x, pressure, num_ghost_cells_side, num_cells_x);
    }
    /* the built-in lower side goes before the upper side's custom hook */
    ErrorStatus e = fill_line(bc->boundary_condition_flag_x_min_, BOUNDARY_CONDITION_NONE, w, 3, 1, 0, 1,
                              num_cells_x, num_ghost_cells_side);
    if (e.return_code != SUCCESS) return e;
    e = fill_line(BOUNDARY_CONDITION_NONE, bc->boundary_condition_flag_x_max_, w, 3, 1, 0, 1, num_cells_x,
                  num_ghost_cells_side);
    if (e.return_code != SUCCESS) return e;
    if (bc->boundary_condition_flag_x_max_ == BOUNDARY_CONDITION_CUSTOM) {
        bc->set_boundary_condition_1d_x_max(density, velocity_x, pressure, num_ghost_cells_side, num_cells_x);
    }
    return make_success_error_status();
}

ErrorStatus set_boundary_condition_cartesian_2d(const BoundaryConditionParam *bc, double *density, double *velocity_x,
                                                double *velocity_y, double *pressure, const int num_ghost_cells_side,
                                                const int num_cells_x, const int num_cells_y)
{
    double *w[4] = {density, velocity_x, velocity_y, pressure};
    const int g = num_ghost_cells_side;
    const long tx = num_cells_x + 2 * g;
    const int flags[4] = {bc->boundary_condition_flag_x_min_, bc->boundary_condition_flag_x_max_,
                          bc->boundary_condition_flag_y_min_, bc->boundary_condition_flag_y_max_};
    const set_boundary_condition_cartesian_2d_func hooks[4] = {
        bc->set_boundary_condition_cartesian_2d_x_min, bc->set_boundary_condition_cartesian_2d_x_max,
        bc->set_boundary_condition_cartesian_2d_y_min, bc->set_boundary_condition_cartesian_2d_y_max};
    /* sides in the reference's order: x_min, x_max, y_min, y_max */
    for (int side = 0; side < 4; side++) {
        if (flags[side] == BOUNDARY_CONDITION_CUSTOM) {
            hooks[side](density, velocity_x, velocity_y, pressure, g, num_cells_x, num_cells_y);
            continue;
        }
        const int lo = (side % 2 == 0) ? flags[side] : BOUNDARY_CONDITION_NONE;
        const int hi = (side % 2 == 1) ? flags[side] : BOUNDARY_CONDITION_NONE;
        ErrorStatus e = make_success_error_status();
        if (side < 2) {
            for (int j = g; j < g + num_cells_y && e.return_code == SUCCESS; j++)
                e = fill_line(lo, hi, w, 4, 1, j * tx, 1, num_cells_x, g);
        } else {
            for (int i = g; i < g + num_cells_x && e.return_code == SUCCESS; i++)
                e = fill_line(lo, hi, w, 4, 2, i, tx, num_cells_y, g);
        }
        if (e.return_code != SUCCESS) return e;
    }
    return make_success_error_status();
}

ErrorStatus set_boundary_condition(const BoundaryConditionParam *bc, System *s)
{
    switch (s->coord_sys_flag_) {
    case COORD_SYS_CARTESIAN_1D:
    case COORD_SYS_CYLINDRICAL_1D:
    case COORD_SYS_SPHERICAL_1D:
        return WRAP_TRACEBACK(set_boundary_condition_1d(bc, s->density_, s->velocity_x_, s->pressure_,
                                                        s->num_ghost_cells_side, s->num_cells_x));
    case COORD_SYS_CARTESIAN_2D:
        return WRAP_TRACEBACK(set_boundary_condition_cartesian_2d(bc, s->density_, s->velocity_x_, s->velocity_y_,
                                                                  s->pressure_, s->num_ghost_cells_side,
                                                                  s->num_cells_x, s->num_cells_y));
    default:
        return WRAP_RAISE_ERROR(VALUE_ERROR, "Coordinate system flag not recognized.");
    }
}
