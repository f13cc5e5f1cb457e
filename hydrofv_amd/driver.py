"""Host-side driver for any library that exports hydroFV's C API.

This mirrors how the reference's example programs drive the library
(e.g. examples/kelvin_helmholtz_2d/kelvin_helmholtz_2d.c:132-201): build a
``System`` with ``system_init``, write the initial primitive state into its
arrays, fill the ghost cells, convert to conserved variables, then call
``launch_simulation``.  The same code runs against ``libhydrofv_amd.so`` (the
MI355X implementation) and against any other build of the same API.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import abi


class HydroError(RuntimeError):
    """A call returned a non-SUCCESS ErrorStatus."""

    def __init__(self, code: int, text: str):
        super().__init__(f"error code {code}:\n{text}")
        self.code = code
        self.text = text


def check(lib, status: abi.ErrorStatus) -> None:
    if status.return_code != abi.SUCCESS:
        raise HydroError(status.return_code, abi.traceback_text(lib, status))


@dataclass
class RunResult:
    num_steps: int
    t: float
    dt: float


class Simulation:
    """One ``System`` owned by ``lib`` (allocated by its ``system_init``)."""

    def __init__(self, lib, coord_sys: str, gamma: float, nx: int, ny: int = 0, g: int = 2,
                 x: tuple = (0.0, 1.0), y: tuple = (0.0, 1.0), gravity: float = 0.0):
        self.lib = abi.bind(lib)
        self._keep = []  # keeps C strings and callbacks alive
        sysm = lib.get_new_system_struct()
        sysm.coord_sys = self._cstr(coord_sys)
        sysm.gamma = gamma
        sysm.gravity = gravity
        sysm.x_min, sysm.x_max = x
        sysm.num_cells_x = nx
        sysm.num_ghost_cells_side = g
        if coord_sys == "cartesian_2d":
            sysm.y_min, sysm.y_max = y
            sysm.num_cells_y = ny
        self.system = sysm
        check(lib, lib.system_init(C.byref(self.system)))
        self.dim = 2 if coord_sys == "cartesian_2d" else 1
        self.g = g
        self.nx = nx
        self.ny = ny if self.dim == 2 else 1
        self.tx = nx + 2 * g
        self.ty = self.ny + 2 * g if self.dim == 2 else 1
        self.bc = lib.get_new_boundary_condition_param()
        self._device_bc = []  # host hooks registered with a device form
        self._closed = False

    # -- helpers ---------------------------------------------------------
    def _cstr(self, s):
        if s is None:
            return None
        b = C.c_char_p(s.encode())
        self._keep.append(b)
        return b

    def _view(self, ptr, n):
        return np.ctypeslib.as_array(ptr, shape=(n,))

    def field(self, name: str) -> np.ndarray:
        """Writable numpy view of a cell array, shaped (ty, tx) in 2D."""
        ptr = getattr(self.system, name + "_")
        arr = self._view(ptr, self.tx * self.ty)
        return arr.reshape(self.ty, self.tx) if self.dim == 2 else arr

    @property
    def mid_x(self):
        return self._view(self.system.mid_points_x_, self.tx)

    @property
    def mid_y(self):
        return self._view(self.system.mid_points_y_, self.ty)

    def primitive(self):
        names = ["density", "velocity_x", "velocity_y", "pressure"] if self.dim == 2 else [
            "density", "velocity_x", "pressure"]
        return [self.field(n).copy() for n in names]

    def conserved(self):
        names = ["mass", "momentum_x", "momentum_y", "energy"] if self.dim == 2 else [
            "mass", "momentum_x", "energy"]
        return [self.field(n).copy() for n in names]

    def set_primitive(self, fields, interior_only=False):
        names = ["density", "velocity_x", "velocity_y", "pressure"] if self.dim == 2 else [
            "density", "velocity_x", "pressure"]
        g = self.g
        for n, a in zip(names, fields):
            dst = self.field(n)
            if interior_only:
                if self.dim == 2:
                    dst[g:g + self.ny, g:g + self.nx] = a
                else:
                    dst[g:g + self.nx] = a
            else:
                dst[...] = a

    # -- API calls -------------------------------------------------------
    def set_boundary(self, x_min, x_max, y_min=None, y_max=None, custom=None, device=None):
        """Boundary conditions by name; `custom` maps a hook field of
        BoundaryConditionParam to a Python function (boundary.h:17-33).
        `device` maps some of those fields to (device hook address, user
        pointer) pairs, registered with hfv_register_device_bc_2d (this
        library only) so that 2D runs fill those sides on the device."""
        bc = self.bc
        bc.boundary_condition_x_min = self._cstr(x_min)
        bc.boundary_condition_x_max = self._cstr(x_max)
        bc.boundary_condition_y_min = self._cstr(y_min)
        bc.boundary_condition_y_max = self._cstr(y_max)
        for name, fn in (custom or {}).items():
            cb = (abi.BC2DFunc if self.dim == 2 else abi.BC1DFunc)(fn)
            self._keep.append(cb)
            setattr(bc, name, cb)
            if device and name in device:
                self._register_device_bc(cb, *device[name])
        check(self.lib, self.lib.finalize_boundary_condition_param(C.byref(self.system), C.byref(bc)))

    def _register_device_bc(self, host_cb, device_fn, user):
        reg = self.lib.hfv_register_device_bc_2d
        reg.restype = C.c_int
        reg.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        host = C.cast(host_cb, C.c_void_p).value
        if reg(host, device_fn, user) != 0:
            raise HydroError(-1, "hfv_register_device_bc_2d failed (table full)")
        self._device_bc.append(host)

    def apply_boundary(self):
        check(self.lib, self.lib.set_boundary_condition(C.byref(self.bc), C.byref(self.system)))

    def primitive_to_conserved(self):
        check(self.lib, self.lib.convert_primitive_to_conserved(C.byref(self.system)))

    def conserved_to_primitive(self):
        check(self.lib, self.lib.convert_conserved_to_primitive(C.byref(self.system)))

    def params(self, integrator, riemann_solver, slope_limiter, cfl, tol=1e-6,
               shrink_factor=None, shrink_steps=None):
        ip = self.lib.get_new_integrator_param()
        ip.integrator = self._cstr(integrator)
        ip.riemann_solver = self._cstr(riemann_solver)
        ip.slope_limiter = self._cstr(slope_limiter)
        ip.cfl = cfl
        ip.tol = tol
        if shrink_factor is not None:
            ip.cfl_initial_shrink_factor = shrink_factor
        if shrink_steps is not None:
            ip.cfl_initial_shrink_num_steps = shrink_steps
        return ip

    def launch(self, integrator, riemann_solver, slope_limiter, cfl, tf, tol=1e-6,
               shrink_factor=None, shrink_steps=None, entry="launch_simulation",
               max_steps=None, fixed_dt=None, store_dir=None, store_interval=None,
               store_initial=False, verbose=0) -> RunResult:
        """Run the integrator.  ``max_steps`` / ``fixed_dt`` select the
        library's step-limited / constant-dt loop (hfv_muscl_hancock_*_run,
        hfv_godunov_first_order_*_run, hfv_random_choice_1d_run).  ``store_dir`` (ending in "/") turns
        on snapshots every ``store_interval`` of simulation time.  ``verbose``
        is Settings.verbose (the exact solver's bisection warnings)."""
        ip = self.params(integrator, riemann_solver, slope_limiter, cfl, tol, shrink_factor, shrink_steps)
        sp = self.lib.get_new_storing_param()
        sp.is_storing = store_dir is not None
        if store_dir is not None:
            self._store_dir = store_dir.encode()
            sp.output_dir = self._store_dir
            sp.storing_interval = float(store_interval)
            sp.store_initial = bool(store_initial)
        st = abi.Settings(verbose=verbose, no_progress_bar=True)
        prm = abi.SimulationParam(tf=tf)
        status = abi.SimulationStatus()
        if max_steps is not None or fixed_dt is not None:
            scheme = next((k for k in ("godunov_first_order", "random_choice") if integrator.startswith(k)),
                          "muscl_hancock")
            entry = f"hfv_{scheme}_{self.dim}d_run"
        if entry != "launch_simulation":
            # direct integrator call: finalize the parameters the way
            # launch_simulation would (hydro.c:41-57)
            check(self.lib, self.lib.finalize_integrator_param(C.byref(ip)))
            check(self.lib, self.lib.finalize_storing_param(C.byref(sp)))
        fn = getattr(self.lib, entry)
        args = [C.byref(self.bc), C.byref(self.system), C.byref(ip), C.byref(sp), C.byref(st),
                C.byref(prm), C.byref(status)]
        if entry.startswith("hfv_"):
            args += [-1 if max_steps is None else int(max_steps), 0.0 if fixed_dt is None else float(fixed_dt)]
        self.last_params = ip
        self.last_status = status  # also after a failed run
        check(self.lib, fn(*args))
        return RunResult(status.num_steps, status.t, status.dt)

    def close(self):
        if not self._closed:
            for host in self._device_bc:
                self.lib.hfv_register_device_bc_2d(host, None, None)
            self._device_bc = []
            self.lib.free_system_memory(C.byref(self.system))
            self._closed = True

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
