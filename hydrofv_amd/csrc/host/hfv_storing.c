/*
 * hfv_storing.c -- snapshots in the reference's HDF5 format (storing.c:40-843):
 * finalize_storing_param and store_snapshot.
 *
 * A snapshot is one file <output_dir>snapshot_<n>.h5 with three groups:
 *   /simulation_status  num_steps (int64), simulation_time, dt (double)
 *   /fields             density, velocity_x[, velocity_y], pressure: the whole
 *                       arrays, ghost cells included, flattened (x fastest)
 *   /parameters         x_min, x_max[, y_min, y_max] (double),
 *                       num_cells_x[, num_cells_y], num_ghost_cells_side (int),
 *                       dx[, dy], gamma (double), coordinate_system and the
 *                       boundary-condition names (variable-length strings)
 * so the reference's plotting scripts read this library's output.
 *
 * The HDF5 C library is loaded on the first snapshot (dlopen), so the
 * library has no link-time dependency on it: HYDROFV_HDF5_LIBRARY names the
 * shared object, else libhdf5.so.103 from the loader's search path or
 * /opt/conda/lib.  Only the stable C API of HDF5 1.10 is used; its handle
 * types and constants below are that API's published values.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <errno.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>
#include <time.h>

#include "hydrofv/hydro.h"

/* ---- the HDF5 1.10 C API, bound at run time ---------------------------- */
typedef int64_t h5_hid;
typedef int h5_herr;
typedef unsigned long long h5_hsize;
#define H5_ACC_TRUNC 0x0002u
#define H5_P_DEFAULT ((h5_hid)0)
#define H5_S_ALL ((h5_hid)0)
#define H5_S_SCALAR 0
#define H5_T_VARIABLE ((size_t)-1)

static struct {
    int loaded;
    h5_herr (*open)(void);
    h5_hid (*fcreate)(const char *, unsigned, h5_hid, h5_hid);
    h5_herr (*fclose)(h5_hid);
    h5_hid (*gcreate)(h5_hid, const char *, h5_hid, h5_hid, h5_hid);
    h5_herr (*gclose)(h5_hid);
    h5_hid (*screate)(int);
    h5_hid (*screate_simple)(int, const h5_hsize *, const h5_hsize *);
    h5_herr (*sclose)(h5_hid);
    h5_hid (*tcopy)(h5_hid);
    h5_herr (*tset_size)(h5_hid, size_t);
    h5_herr (*tclose)(h5_hid);
    h5_hid (*dcreate)(h5_hid, const char *, h5_hid, h5_hid, h5_hid, h5_hid, h5_hid);
    h5_herr (*dwrite)(h5_hid, h5_hid, h5_hid, h5_hid, h5_hid, const void *);
    h5_herr (*dclose)(h5_hid);
    h5_hid *native_double, *native_llong, *native_int, *c_s1;
} H5;

static int h5_load(void)
{
    if (H5.loaded) return 1;
    void *lib = NULL;
    const char *env = getenv("HYDROFV_HDF5_LIBRARY");
    if (env && env[0]) lib = dlopen(env, RTLD_NOW | RTLD_LOCAL);
    if (!lib) lib = dlopen("libhdf5.so.103", RTLD_NOW | RTLD_LOCAL);
    if (!lib) lib = dlopen("/opt/conda/lib/libhdf5.so.103", RTLD_NOW | RTLD_LOCAL);
    if (!lib) return 0;
#define H5_SYM(field, name)                         \
    do {                                            \
        *(void **)(&H5.field) = dlsym(lib, name);   \
        if (!H5.field) return 0;                    \
    } while (0)
    H5_SYM(open, "H5open");
    H5_SYM(fcreate, "H5Fcreate");
    H5_SYM(fclose, "H5Fclose");
    H5_SYM(gcreate, "H5Gcreate2");
    H5_SYM(gclose, "H5Gclose");
    H5_SYM(screate, "H5Screate");
    H5_SYM(screate_simple, "H5Screate_simple");
    H5_SYM(sclose, "H5Sclose");
    H5_SYM(tcopy, "H5Tcopy");
    H5_SYM(tset_size, "H5Tset_size");
    H5_SYM(tclose, "H5Tclose");
    H5_SYM(dcreate, "H5Dcreate2");
    H5_SYM(dwrite, "H5Dwrite");
    H5_SYM(dclose, "H5Dclose");
    H5_SYM(native_double, "H5T_NATIVE_DOUBLE_g");
    H5_SYM(native_llong, "H5T_NATIVE_LLONG_g");
    H5_SYM(native_int, "H5T_NATIVE_INT_g");
    H5_SYM(c_s1, "H5T_C_S1_g");
#undef H5_SYM
    /* the type ids are valid once the library is initialised */
    if (H5.open() < 0) return 0;
    H5.loaded = 1;
    return 1;
}

/* ---- storing.c:40-228 --------------------------------------------------- */
ErrorStatus finalize_storing_param(StoringParam *p)
{
    char msg[1024];
    if (!p->is_storing) return make_success_error_status();
    if (p->storing_interval <= 0.0) return WRAP_RAISE_ERROR(VALUE_ERROR, "Storing interval must be positive.");
    if (!p->output_dir) {
        /* snapshots_YYYYMMDD_HHMMSS/ (owned by the parameters from here on) */
        const size_t len = strlen("snapshots_YYYYMMDD_HHMMSS/") + 1;
        char *dir = malloc(len);
        if (!dir) return WRAP_RAISE_ERROR(MEMORY_ERROR, "Failed to allocate memory for directory path.");
        const time_t now = time(NULL);
        strftime(dir, len, "snapshots_%Y%m%d_%H%M%S/", localtime(&now));
        p->output_dir = dir;
    } else {
        const size_t n = strlen(p->output_dir);
        if (n == 0 || p->output_dir[n - 1] != '/') {
            snprintf(msg, sizeof msg,
                     "Directory path for storing snapshots must end with a trailing slash (\"/\"). Got: \"%s\".",
                     p->output_dir);
            return WRAP_RAISE_ERROR(VALUE_ERROR, msg);
        }
    }
    if (mkdir(p->output_dir, 0777) == -1) {
        struct stat st;
        if (stat(p->output_dir, &st) != 0) {
            snprintf(msg, sizeof msg, "Failed to access path for storing snapshots: \"%s\".", p->output_dir);
            return WRAP_RAISE_ERROR(OS_ERROR, msg);
        }
        if (S_ISDIR(st.st_mode)) {
            snprintf(msg, sizeof msg,
                     "Directory for storing snapshots already exists. The files will be overwritten. Directory: \"%s\".",
                     p->output_dir);
            WRAP_RAISE_WARNING(msg);
        } else {
            snprintf(msg, sizeof msg, "Failed to create directory for storing snapshots: \"%s\".", p->output_dir);
            return WRAP_RAISE_ERROR(OS_ERROR, msg);
        }
    }
    return make_success_error_status();
}

/* ---- storing.c:230-758: one snapshot file ------------------------------ */
#define MAX_DSETS 32

typedef struct {
    h5_hid file, groups[3], scalar, field, vstr;
    h5_hid dsets[MAX_DSETS];
    int ndsets;
    int failed;
} Snapshot;

static void snap_close(Snapshot *s)
{
    for (int k = 0; k < s->ndsets; k++)
        if (s->dsets[k] >= 0) H5.dclose(s->dsets[k]);
    if (s->vstr >= 0) H5.tclose(s->vstr);
    if (s->scalar >= 0) H5.sclose(s->scalar);
    if (s->field >= 0) H5.sclose(s->field);
    for (int k = 0; k < 3; k++)
        if (s->groups[k] >= 0) H5.gclose(s->groups[k]);
    if (s->file >= 0) H5.fclose(s->file);
}

/* creates dataset `name` in group g with the given type and space, writes *buf */
static void put(Snapshot *s, int g, const char *name, h5_hid type, h5_hid space, const void *buf)
{
    if (s->failed || s->ndsets == MAX_DSETS) {
        s->failed = 1;
        return;
    }
    const h5_hid d = H5.dcreate(s->groups[g], name, type, space, H5_P_DEFAULT, H5_P_DEFAULT, H5_P_DEFAULT);
    s->dsets[s->ndsets++] = d;
    if (d < 0 || H5.dwrite(d, type, H5_S_ALL, H5_S_ALL, H5_P_DEFAULT, buf) < 0) s->failed = 1;
}

static ErrorStatus write_snapshot(const BoundaryConditionParam *bc, const System *sy,
                                  const SimulationStatus *status, const char *path, int two_d)
{
    Snapshot s;
    memset(&s, 0, sizeof s);
    s.file = s.groups[0] = s.groups[1] = s.groups[2] = s.scalar = s.field = s.vstr = -1;
    s.file = H5.fcreate(path, H5_ACC_TRUNC, H5_P_DEFAULT, H5_P_DEFAULT);
    if (s.file < 0) return WRAP_RAISE_ERROR(OS_ERROR, "Failed to create HDF5 file.");
    s.groups[0] = H5.gcreate(s.file, "/simulation_status", H5_P_DEFAULT, H5_P_DEFAULT, H5_P_DEFAULT);
    s.groups[1] = H5.gcreate(s.file, "/fields", H5_P_DEFAULT, H5_P_DEFAULT, H5_P_DEFAULT);
    s.groups[2] = H5.gcreate(s.file, "/parameters", H5_P_DEFAULT, H5_P_DEFAULT, H5_P_DEFAULT);
    if (s.groups[0] < 0 || s.groups[1] < 0 || s.groups[2] < 0) {
        snap_close(&s);
        return WRAP_RAISE_ERROR(OS_ERROR, "Failed to create HDF5 group.");
    }
    const h5_hsize tx = (h5_hsize)sy->num_cells_x + 2 * (h5_hsize)sy->num_ghost_cells_side;
    const h5_hsize ty = two_d ? (h5_hsize)sy->num_cells_y + 2 * (h5_hsize)sy->num_ghost_cells_side : 1;
    const h5_hsize total = tx * ty;
    s.scalar = H5.screate(H5_S_SCALAR);
    s.field = H5.screate_simple(1, &total, NULL);
    if (s.scalar < 0 || s.field < 0) {
        snap_close(&s);
        return WRAP_RAISE_ERROR(OS_ERROR, "Failed to create HDF5 dataspace.");
    }
    s.vstr = H5.tcopy(*H5.c_s1);
    if (s.vstr < 0 || H5.tset_size(s.vstr, H5_T_VARIABLE) < 0) s.failed = 1;

    const h5_hid dbl = *H5.native_double, i64 = *H5.native_llong, i32 = *H5.native_int;
    enum { STATUS = 0, FIELDS = 1, PARAMS = 2 };
    put(&s, STATUS, "num_steps", i64, s.scalar, &status->num_steps);
    put(&s, STATUS, "simulation_time", dbl, s.scalar, &status->t);
    put(&s, STATUS, "dt", dbl, s.scalar, &status->dt);
    if (s.failed) {
        snap_close(&s);
        return WRAP_RAISE_ERROR(OS_ERROR, "Failed to create HDF5 simulation status datasets.");
    }
    put(&s, FIELDS, "density", dbl, s.field, sy->density_);
    put(&s, FIELDS, "velocity_x", dbl, s.field, sy->velocity_x_);
    if (two_d) put(&s, FIELDS, "velocity_y", dbl, s.field, sy->velocity_y_);
    put(&s, FIELDS, "pressure", dbl, s.field, sy->pressure_);
    if (s.failed) {
        snap_close(&s);
        return WRAP_RAISE_ERROR(OS_ERROR, "Failed to create HDF5 fields datasets.");
    }
    put(&s, PARAMS, "x_min", dbl, s.scalar, &sy->x_min);
    put(&s, PARAMS, "x_max", dbl, s.scalar, &sy->x_max);
    if (two_d) {
        put(&s, PARAMS, "y_min", dbl, s.scalar, &sy->y_min);
        put(&s, PARAMS, "y_max", dbl, s.scalar, &sy->y_max);
    }
    put(&s, PARAMS, "num_cells_x", i32, s.scalar, &sy->num_cells_x);
    if (two_d) put(&s, PARAMS, "num_cells_y", i32, s.scalar, &sy->num_cells_y);
    put(&s, PARAMS, "num_ghost_cells_side", i32, s.scalar, &sy->num_ghost_cells_side);
    put(&s, PARAMS, "dx", dbl, s.scalar, &sy->dx_);
    if (two_d) put(&s, PARAMS, "dy", dbl, s.scalar, &sy->dy_);
    put(&s, PARAMS, "gamma", dbl, s.scalar, &sy->gamma);
    put(&s, PARAMS, "coordinate_system", s.vstr, s.scalar, &sy->coord_sys);
    put(&s, PARAMS, "boundary_condition_x_min", s.vstr, s.scalar, &bc->boundary_condition_x_min);
    put(&s, PARAMS, "boundary_condition_x_max", s.vstr, s.scalar, &bc->boundary_condition_x_max);
    if (two_d) {
        put(&s, PARAMS, "boundary_condition_y_min", s.vstr, s.scalar, &bc->boundary_condition_y_min);
        put(&s, PARAMS, "boundary_condition_y_max", s.vstr, s.scalar, &bc->boundary_condition_y_max);
    }
    const int failed = s.failed;
    snap_close(&s);
    if (failed) return WRAP_RAISE_ERROR(OS_ERROR, "Failed to create HDF5 parameters datasets.");
    return make_success_error_status();
}

/* ---- storing.c:761-843 -------------------------------------------------- */
ErrorStatus store_snapshot(const BoundaryConditionParam *bc, const System *system, const IntegratorParam *ip,
                           const SimulationStatus *status, StoringParam *sp)
{
    (void)ip;
    if (!sp->output_dir) return WRAP_RAISE_ERROR(POINTER_ERROR, "Directory path for storing is NULL.");
    int two_d;
    switch (system->coord_sys_flag_) {
    case COORD_SYS_CARTESIAN_1D:
    case COORD_SYS_CYLINDRICAL_1D:
    case COORD_SYS_SPHERICAL_1D:
        two_d = 0;
        break;
    case COORD_SYS_CARTESIAN_2D:
        two_d = 1;
        break;
    case COORD_SYS_CARTESIAN_3D:
        return WRAP_RAISE_ERROR(NOT_IMPLEMENTED_ERROR, "Storing snapshot for Cartesian 3D system is not implemented.");
    default:
        return WRAP_RAISE_ERROR(VALUE_ERROR, "Coordinate system flag not recognized.");
    }
    const int len = snprintf(NULL, 0, "%ssnapshot_%d.h5", sp->output_dir, sp->store_count_) + 1;
    char *path = malloc((size_t)len);
    if (!path) return WRAP_RAISE_ERROR(MEMORY_ERROR, "Failed to allocate memory for file path string.");
    snprintf(path, (size_t)len, "%ssnapshot_%d.h5", sp->output_dir, sp->store_count_);
    ErrorStatus e;
    if (!h5_load()) {
        e = WRAP_RAISE_ERROR(OS_ERROR,
                             "Failed to load the HDF5 library (libhdf5.so.103); set HYDROFV_HDF5_LIBRARY to its path.");
    } else {
        e = WRAP_TRACEBACK(write_snapshot(bc, system, status, path, two_d));
        if (e.return_code == SUCCESS) sp->store_count_++;
    }
    free(path);
    return e;
}
