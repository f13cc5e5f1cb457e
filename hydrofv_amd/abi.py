"""ctypes mirror of hydroFV's C API (the drop-in boundary).

The struct layouts and function signatures are those of the reference headers
(src/common.h:22-69, src/system.h:93-130, src/boundary.h:17-56,
src/integrator.h:145-176, src/hydro.h:53-61).  ``libhydrofv_amd.so`` exports
the same symbols with the same layouts, so one binding drives either library;
``bind(lib)`` attaches the prototypes to any ``ctypes.CDLL`` that exports them.
"""
from __future__ import annotations

import ctypes as C

c_double_p = C.POINTER(C.c_double)

# error.h:15-22
UNKNOWN_ERROR = -1
SUCCESS = 0
FAILURE = 1
VALUE_ERROR = 2
POINTER_ERROR = 3
MEMORY_ERROR = 4
OS_ERROR = 5
NOT_IMPLEMENTED_ERROR = 6

# system.h:16-20
COORD_SYS_CARTESIAN_1D = 1
COORD_SYS_CARTESIAN_2D = 2
COORD_SYS_CARTESIAN_3D = 3
COORD_SYS_CYLINDRICAL_1D = 4
COORD_SYS_SPHERICAL_1D = 5

# boundary.h:59-63
BOUNDARY_CONDITION_NONE = 1
BOUNDARY_CONDITION_REFLECTIVE = 2
BOUNDARY_CONDITION_TRANSMISSIVE = 3
BOUNDARY_CONDITION_PERIODIC = 4
BOUNDARY_CONDITION_CUSTOM = 5

# slope_limiter.h:14-16, riemann_solver.h:23-24
SLOPE_LIMITER_MINMOD = 1
SLOPE_LIMITER_VAN_LEER = 2
SLOPE_LIMITER_MONOTONIZED_CENTRAL = 3
RIEMANN_SOLVER_EXACT = 1
RIEMANN_SOLVER_HLLC = 2


class ErrorStatus(C.Structure):  # common.h:22-27
    _fields_ = [
        ("return_code", C.c_int),
        ("traceback", C.c_void_p),
        ("traceback_code_", C.c_int),
    ]


class IntegratorParam(C.Structure):  # common.h:29-41
    _fields_ = [
        ("integrator", C.c_char_p),
        ("riemann_solver", C.c_char_p),
        ("slope_limiter", C.c_char_p),
        ("cfl", C.c_double),
        ("cfl_initial_shrink_factor", C.c_double),
        ("cfl_initial_shrink_num_steps", C.c_int),
        ("tol", C.c_double),
        ("integrator_flag_", C.c_int),
        ("riemann_solver_flag_", C.c_int),
        ("slope_limiter_flag_", C.c_int),
    ]


class StoringParam(C.Structure):  # common.h:43-51
    _fields_ = [
        ("output_dir", C.c_char_p),
        ("is_storing", C.c_bool),
        ("store_initial", C.c_bool),
        ("storing_interval", C.c_double),
        ("storing_method_flag_", C.c_int),
        ("store_count_", C.c_int),
    ]


class Settings(C.Structure):  # common.h:53-57
    _fields_ = [("verbose", C.c_int), ("no_progress_bar", C.c_bool)]


class SimulationParam(C.Structure):  # common.h:59-62
    _fields_ = [("tf", C.c_double)]


class SimulationStatus(C.Structure):  # common.h:64-69
    _fields_ = [("num_steps", C.c_int64), ("t", C.c_double), ("dt", C.c_double)]


class System(C.Structure):  # system.h:93-130
    _fields_ = [
        ("coord_sys", C.c_char_p),
        ("gamma", C.c_double),
        ("gravity", C.c_double),
        ("x_min", C.c_double),
        ("x_max", C.c_double),
        ("y_min", C.c_double),
        ("y_max", C.c_double),
        ("z_min", C.c_double),
        ("z_max", C.c_double),
        ("num_cells_x", C.c_int),
        ("num_cells_y", C.c_int),
        ("num_cells_z", C.c_int),
        ("num_ghost_cells_side", C.c_int),
        ("coord_sys_flag_", C.c_int),
        ("dx_", C.c_double),
        ("dy_", C.c_double),
        ("dz_", C.c_double),
        ("density_", c_double_p),
        ("velocity_x_", c_double_p),
        ("velocity_y_", c_double_p),
        ("velocity_z_", c_double_p),
        ("pressure_", c_double_p),
        ("mass_", c_double_p),
        ("momentum_x_", c_double_p),
        ("momentum_y_", c_double_p),
        ("momentum_z_", c_double_p),
        ("energy_", c_double_p),
        ("mid_points_x_", c_double_p),
        ("mid_points_y_", c_double_p),
        ("mid_points_z_", c_double_p),
        ("surface_area_x_", c_double_p),
        ("surface_area_y_", c_double_p),
        ("surface_area_z_", c_double_p),
        ("volume_", c_double_p),
    ]


# boundary.h:17-33
BC1DFunc = C.CFUNCTYPE(None, c_double_p, c_double_p, c_double_p, C.c_int, C.c_int)
BC2DFunc = C.CFUNCTYPE(None, c_double_p, c_double_p, c_double_p, c_double_p, C.c_int, C.c_int, C.c_int)


class BoundaryConditionParam(C.Structure):  # boundary.h:35-56
    _fields_ = [
        ("boundary_condition_x_min", C.c_char_p),
        ("boundary_condition_x_max", C.c_char_p),
        ("boundary_condition_y_min", C.c_char_p),
        ("boundary_condition_y_max", C.c_char_p),
        ("boundary_condition_z_min", C.c_char_p),
        ("boundary_condition_z_max", C.c_char_p),
        ("set_boundary_condition_1d_x_min", BC1DFunc),
        ("set_boundary_condition_1d_x_max", BC1DFunc),
        ("set_boundary_condition_cartesian_2d_x_min", BC2DFunc),
        ("set_boundary_condition_cartesian_2d_x_max", BC2DFunc),
        ("set_boundary_condition_cartesian_2d_y_min", BC2DFunc),
        ("set_boundary_condition_cartesian_2d_y_max", BC2DFunc),
        ("boundary_condition_flag_x_min_", C.c_int),
        ("boundary_condition_flag_x_max_", C.c_int),
        ("boundary_condition_flag_y_min_", C.c_int),
        ("boundary_condition_flag_y_max_", C.c_int),
        ("boundary_condition_flag_z_min_", C.c_int),
        ("boundary_condition_flag_z_max_", C.c_int),
    ]


_P = C.POINTER
_DP = C.POINTER(C.c_double)
_SEVEN = [
    _P(BoundaryConditionParam),
    _P(System),
    _P(IntegratorParam),
    _P(StoringParam),
    _P(Settings),
    _P(SimulationParam),
    _P(SimulationStatus),
]

# name -> (restype, argtypes) for the part of the reference API that both
# libraries export.
PROTOTYPES = {
    "make_success_error_status": (ErrorStatus, []),
    "free_traceback": (None, [_P(ErrorStatus)]),
    "print_and_free_traceback": (None, [_P(ErrorStatus)]),
    "get_new_system_struct": (System, []),
    "system_init": (ErrorStatus, [_P(System)]),
    "free_system_memory": (None, [_P(System)]),
    "convert_primitive_to_conserved": (ErrorStatus, [_P(System)]),
    "convert_conserved_to_primitive": (ErrorStatus, [_P(System)]),
    "get_new_boundary_condition_param": (BoundaryConditionParam, []),
    "finalize_boundary_condition_param": (ErrorStatus, [_P(System), _P(BoundaryConditionParam)]),
    "set_boundary_condition": (ErrorStatus, [_P(BoundaryConditionParam), _P(System)]),
    "get_new_integrator_param": (IntegratorParam, []),
    "finalize_integrator_param": (ErrorStatus, [_P(IntegratorParam)]),
    "get_new_storing_param": (StoringParam, []),
    "finalize_storing_param": (ErrorStatus, [_P(StoringParam)]),
    "launch_simulation": (ErrorStatus, _SEVEN),
    "integrator_launch_simulation": (ErrorStatus, _SEVEN),
    "muscl_hancock_1d": (ErrorStatus, _SEVEN),
    "muscl_hancock_2d": (ErrorStatus, _SEVEN),
    "godunov_first_order_1d": (ErrorStatus, _SEVEN),
    "godunov_first_order_2d": (ErrorStatus, _SEVEN),
    "random_choice_1d": (ErrorStatus, _SEVEN),
    "store_snapshot": (ErrorStatus, [_P(BoundaryConditionParam), _P(System), _P(IntegratorParam),
                                     _P(SimulationStatus), _P(StoringParam)]),
    # per-interface and per-cell functions (slope_limiter.h, riemann_solver*.h, utils.h, source_term.h)
    "limit_slope": (ErrorStatus, [_DP, _P(IntegratorParam), C.c_double, C.c_double]),
    "solve_flux_1d": (ErrorStatus, [_P(IntegratorParam), _P(Settings)] + [_DP] * 3 + [C.c_double] * 7),
    "solve_flux_2d": (ErrorStatus, [_P(IntegratorParam), _P(Settings)] + [_DP] * 4 + [C.c_double] * 9),
    "riemann_A_L_or_R": (C.c_double, [C.c_double] * 2),
    "riemann_B_L_or_R": (C.c_double, [C.c_double] * 2),
    "guess_p": (C.c_double, [C.c_double] * 10),
    "solve_vacuum_1d": (None, [_DP] * 3 + [C.c_double] * 10),
    "solve_vacuum_2d": (None, [_DP] * 4 + [C.c_double] * 12),
    "solve_flux_hllc_1d": (ErrorStatus, [_DP] * 3 + [C.c_double] * 8),
    "solve_flux_hllc_2d": (ErrorStatus, [_DP] * 4 + [C.c_double] * 10),
    "solve_exact_1d": (ErrorStatus, [_DP] * 3 + [C.c_double] * 9 + [C.c_int]),
    "solve_flux_exact_1d": (ErrorStatus, [_DP] * 3 + [C.c_double] * 9 + [C.c_int]),
    "get_sound_speed": (C.c_double, [C.c_double] * 3),
    "get_time_step_1d": (C.c_double, [_P(System), C.c_double]),
    "get_time_step_2d": (C.c_double, [_P(System), C.c_double]),
    "add_geometry_source_term": (ErrorStatus, [_P(BoundaryConditionParam), _P(System), C.c_double]),
    "add_gravity_source_term_2d": (ErrorStatus, [_P(BoundaryConditionParam), _P(System), C.c_double]),
}


def bind(lib: C.CDLL, names=None) -> C.CDLL:
    """Attach restype/argtypes for every name in PROTOTYPES (or ``names``)."""
    for name in names or PROTOTYPES:
        restype, argtypes = PROTOTYPES[name]
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    return lib


def traceback_text(lib: C.CDLL, status: ErrorStatus) -> str:
    """Return the traceback string of a failed call and free it."""
    if not status.traceback:
        return ""
    text = C.string_at(status.traceback).decode(errors="replace")
    lib.free_traceback(C.byref(status))
    return text
