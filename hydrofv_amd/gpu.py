"""ctypes mirror of include/hydrofv/hfv_gpu.h (the device-level entry points)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi

dp = C.POINTER(C.c_double)

HFV_OK = 0
HFV_EINVAL, HFV_ENOMEM, HFV_EDEVICE, HFV_ETOOBIG = -1, -2, -3, -4
HFV_NOALLOC_SCRATCH_BYTES = 64 << 20
ERR_NAMES = {0: "none", 1: "dt is zero", 2: "HLLC invalid wave speed", 3: "exact: bracket failure",
             4: "exact: bisection did not converge"}


class MH2DConfig(C.Structure):
    _fields_ = [
        ("num_cells_x", C.c_int), ("num_cells_y", C.c_int), ("num_ghost_cells_side", C.c_int),
        ("gamma", C.c_double), ("gravity", C.c_double),
        ("dx", C.c_double), ("dy", C.c_double),
        ("area_x", C.c_double), ("area_y", C.c_double), ("volume", C.c_double),
        ("slope_limiter", C.c_int), ("riemann_solver", C.c_int),
        ("tol", C.c_double),
        ("bc", C.c_int * 4),
        ("cfl", C.c_double), ("cfl_initial_shrink_factor", C.c_double),
        ("cfl_initial_shrink_num_steps", C.c_int),
        ("tf", C.c_double),
        ("max_steps", C.c_longlong),
        ("fixed_dt", C.c_double),
        ("strip_rows", C.c_int),
        ("hllc_form", C.c_int),
        ("first_order", C.c_int),
        ("full_tiles", C.c_int),
    ]


HLLC_SELECTS, HLLC_BRANCHES = 0, 1  # hfv_mh2d_config.hllc_form


class MH1DConfig(C.Structure):
    _fields_ = [
        ("num_cells_x", C.c_int), ("num_ghost_cells_side", C.c_int),
        ("coord_sys_flag", C.c_int),
        ("gamma", C.c_double), ("dx", C.c_double),
        ("slope_limiter", C.c_int), ("riemann_solver", C.c_int),
        ("tol", C.c_double),
        ("bc", C.c_int * 2),
        ("cfl", C.c_double), ("cfl_initial_shrink_factor", C.c_double),
        ("cfl_initial_shrink_num_steps", C.c_int),
        ("tf", C.c_double),
        ("max_steps", C.c_longlong),
        ("fixed_dt", C.c_double),
        ("first_order", C.c_int),
        ("random_choice", C.c_int),
        ("t_start", C.c_double),
        ("steps_start", C.c_longlong),
        ("t_pause", C.c_double),
    ]


class RunStatus(C.Structure):
    _fields_ = [
        ("t", C.c_double), ("dt", C.c_double), ("t_prev", C.c_double),
        ("num_steps", C.c_longlong),
        ("err_kind", C.c_int), ("cur", C.c_int),
        ("err_cell", C.c_longlong),
        ("smax", C.c_double * 2),
        ("finished", C.c_int),
        ("bisect_fails", C.c_longlong),
    ]


_P = C.POINTER
_SEVEN = abi.PROTOTYPES["muscl_hancock_2d"][1]

PROTOTYPES = {
    "hfv_mh2d_create": (C.c_int, [_P(MH2DConfig), _P(C.c_void_p)]),
    "hfv_mh2d_destroy": (None, [C.c_void_p]),
    "hfv_mh2d_upload": (C.c_int, [C.c_void_p, _P(dp), _P(dp), C.c_void_p]),
    "hfv_mh2d_start": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hfv_mh2d_advance": (C.c_int, [C.c_void_p, C.c_longlong, C.c_int, C.c_void_p]),
    "hfv_mh2d_status": (C.c_int, [C.c_void_p, _P(RunStatus), C.c_void_p]),
    "hfv_mh2d_download": (C.c_int, [C.c_void_p, _P(dp), _P(dp), C.c_void_p]),
    "hfv_mh2d_fields": (C.c_int, [C.c_void_p, _P(C.c_void_p), _P(C.c_void_p), _P(C.c_void_p), _P(C.c_longlong)]),
    "hfv_mh2d_run_noalloc": (C.c_int, [_P(MH2DConfig), _P(C.c_void_p), _P(C.c_void_p), C.c_longlong, C.c_longlong,
                                       _P(RunStatus), C.c_void_p]),
    "hfv_mh2d_set_device_bc": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "hfv_register_device_bc_2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "hfv_mh1d_run": (C.c_int, [_P(MH1DConfig), _P(dp), _P(dp), dp, dp, dp, _P(RunStatus), C.c_void_p]),
    "hfv_probe_flux": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_longlong, dp, dp, dp,
                                 _P(C.c_int)]),
    "hfv_probe_limit": (C.c_int, [C.c_int, C.c_longlong, dp, dp, dp]),
    "hfv_probe_pow": (C.c_int, [C.c_int, C.c_longlong, dp, dp, dp]),
    "hfv_probe_sqrt": (C.c_int, [C.c_int, C.c_longlong, dp, dp]),
    "hfv_probe_div": (C.c_int, [C.c_int, C.c_longlong, dp, dp, dp]),
    "hfv_probe_exact_finish": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_int, C.c_longlong, dp, dp, dp,
                                         _P(C.c_int)]),
    "hfv_device_error_string": (C.c_char_p, []),
    "hfv_muscl_hancock_1d_run": (abi.ErrorStatus, list(_SEVEN) + [C.c_longlong, C.c_double]),
    "hfv_muscl_hancock_2d_run": (abi.ErrorStatus, list(_SEVEN) + [C.c_longlong, C.c_double]),
    "hfv_godunov_first_order_1d_run": (abi.ErrorStatus, list(_SEVEN) + [C.c_longlong, C.c_double]),
    "hfv_godunov_first_order_2d_run": (abi.ErrorStatus, list(_SEVEN) + [C.c_longlong, C.c_double]),
    "hfv_random_choice_1d_run": (abi.ErrorStatus, list(_SEVEN) + [C.c_longlong, C.c_double]),
}


def bind(lib):
    for name, (restype, argtypes) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    return lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous, "arrays must be contiguous float64"
    return a.ctypes.data_as(dp)


def _ptr_array(arrs):
    out = (dp * len(arrs))()
    for k, a in enumerate(arrs):
        out[k] = _ptr(a)
    return out


def check(lib, rc: int, what: str):
    if rc != HFV_OK:
        raise RuntimeError(f"{what} failed with code {rc}: {lib.hfv_device_error_string().decode()}")


def probe_flux(lib, nf: int, solver: int, gamma: float, tol: float, wl: np.ndarray, wr: np.ndarray):
    """Evaluate the device flux function on n interface states (host arrays)."""
    wl = np.ascontiguousarray(wl, dtype=np.float64)
    wr = np.ascontiguousarray(wr, dtype=np.float64)
    n = wl.shape[0]
    f = np.zeros_like(wl)
    err = np.zeros(n, dtype=np.int32)
    check(lib, lib.hfv_probe_flux(nf, solver, gamma, tol, n, _ptr(wl), _ptr(wr), _ptr(f),
                                  err.ctypes.data_as(C.POINTER(C.c_int))), "hfv_probe_flux")
    return f, err


def probe_limit(lib, limiter: int, a: np.ndarray, b: np.ndarray):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = np.zeros_like(a)
    check(lib, lib.hfv_probe_limit(limiter, a.size, _ptr(a), _ptr(b), _ptr(out)), "hfv_probe_limit")
    return out


def probe_pow(lib, impl: int, x: np.ndarray, y: np.ndarray):
    """x^y on the device: impl 0 the library pow, 1 the kernels' pow_pos."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros_like(x)
    check(lib, lib.hfv_probe_pow(impl, x.size, _ptr(x), _ptr(y), _ptr(out)), "hfv_probe_pow")
    return out



def probe_sqrt(lib, impl: int, x: np.ndarray):
    """sqrt(x) on the device: impl 0 the library sqrt, 1 the kernels' grouped
    root, 2 its unscaled core (arguments in [2^-767, +inf) only)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    check(lib, lib.hfv_probe_sqrt(impl, x.size, _ptr(x), _ptr(out)), "hfv_probe_sqrt")
    return out


def probe_div(lib, impl: int, a: np.ndarray, b: np.ndarray):
    """a / b on the device: impl 0 the division, 1 the kernels' grouped form,
    2 its unscaled core (operands of magnitude in [2^-300, 2^300] only)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = np.zeros_like(a)
    check(lib, lib.hfv_probe_div(impl, a.size, _ptr(a), _ptr(b), _ptr(out)), "hfv_probe_div")
    return out


def probe_exact_finish(lib, impl: int, gamma: float, tol: float, cap: int, wl: np.ndarray, wr: np.ndarray):
    """The exact solver's pressure iteration: Newton for cap iterations, then
    the finish of a face not converged, serially (impl 0) or by a wavefront
    per face (impl 1); impl 2: Newton for cap iterations, the rest of the
    loop by four lanes per face (the 2D step's set-aside faces), then the
    finish.  Returns (p*, rc); rc -1: converged, -3: vacuum, else the
    failure kind."""
    wl = np.ascontiguousarray(wl, dtype=np.float64)
    wr = np.ascontiguousarray(wr, dtype=np.float64)
    n = wl.shape[0]
    p = np.zeros(n)
    rc = np.zeros(n, dtype=np.int32)
    check(lib, lib.hfv_probe_exact_finish(impl, gamma, tol, cap, n, _ptr(wl), _ptr(wr), _ptr(p),
                                          rc.ctypes.data_as(C.POINTER(C.c_int))), "hfv_probe_exact_finish")
    return p, rc


class MH2D:
    """Device-resident 2D run (hfv_mh2d_*): upload once, queue steps, read back."""

    def __init__(self, lib, cfg: MH2DConfig):
        self.lib = lib
        self.cfg = cfg
        h = C.c_void_p()
        check(lib, lib.hfv_mh2d_create(C.byref(cfg), C.byref(h)), "hfv_mh2d_create")
        self.h = h
        self.nx, self.ny, self.g = cfg.num_cells_x, cfg.num_cells_y, cfg.num_ghost_cells_side
        self.shape = (self.ny + 2 * self.g, self.nx + 2 * self.g)

    def upload(self, W, U, stream=None):
        W = [np.ascontiguousarray(w, dtype=np.float64) for w in W]
        U = [np.ascontiguousarray(u, dtype=np.float64) for u in U]
        for a in W + U:
            assert a.shape == self.shape
        check(self.lib, self.lib.hfv_mh2d_upload(self.h, _ptr_array(W), _ptr_array(U), stream), "upload")
        if stream is None:
            self.status(None)  # synchronises the null stream before the host arrays go away

    def start(self, stream=None):
        check(self.lib, self.lib.hfv_mh2d_start(self.h, stream), "start")

    def advance(self, launches: int, use_graph: bool = True, stream=None):
        check(self.lib, self.lib.hfv_mh2d_advance(self.h, launches, int(use_graph), stream), "advance")

    def status(self, stream=None) -> RunStatus:
        st = RunStatus()
        check(self.lib, self.lib.hfv_mh2d_status(self.h, C.byref(st), stream), "status")
        return st

    def download(self, stream=None):
        W = [np.empty(self.shape) for _ in range(4)]
        U = [np.empty(self.shape) for _ in range(4)]
        check(self.lib, self.lib.hfv_mh2d_download(self.h, _ptr_array(W), _ptr_array(U), stream), "download")
        return W, U

    def close(self):
        if self.h:
            self.lib.hfv_mh2d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
