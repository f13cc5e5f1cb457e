"""Set-ups of the benchmark problems through the C API, in the order the
reference's example drivers use (system_init, initial primitive state, ghost
fill, primitive -> conserved; e.g. examples/kelvin_helmholtz_2d/
kelvin_helmholtz_2d.c:53-128).  Each returns a driver.Simulation bound to the
given library, ready for launch()."""
from __future__ import annotations

from . import driver, workloads


def kelvin_helmholtz(lib, n: int, g: int = 2, noise: float = 1e-3) -> driver.Simulation:
    s = driver.Simulation(lib, "cartesian_2d", 5.0 / 3.0, n, n, g)
    s.set_boundary("periodic", "periodic", "periodic", "periodic")
    s.set_primitive(workloads.kelvin_helmholtz(n, g, noise=noise), interior_only=True)
    s.apply_boundary()
    s.primitive_to_conserved()
    return s


def sedov(lib, n: int, g: int = 2, lo: float = -1.0, hi: float = 1.0) -> driver.Simulation:
    s = driver.Simulation(lib, "cartesian_2d", 1.4, n, n, g, x=(lo, hi), y=(lo, hi))
    s.set_boundary("transmissive", "transmissive", "transmissive", "transmissive")
    s.set_primitive(workloads.sedov(n, g, lo=lo, hi=hi), interior_only=True)
    s.apply_boundary()
    s.primitive_to_conserved()
    return s


def sod(lib, n: int = 1000, g: int = 2, coord: str = "cartesian_1d",
        bc=("transmissive", "transmissive")) -> driver.Simulation:
    s = driver.Simulation(lib, coord, 1.4, n, g=g)
    s.set_primitive(workloads.sod(n, g))
    s.primitive_to_conserved()
    s.set_boundary(*bc)
    return s


def four_quadrant(lib, n: int, states, g: int = 2) -> driver.Simulation:
    s = driver.Simulation(lib, "cartesian_2d", 1.4, n, n, g)
    s.set_boundary("transmissive", "transmissive", "transmissive", "transmissive")
    s.set_primitive(workloads.four_quadrant(n, states, g), interior_only=True)
    s.apply_boundary()
    s.primitive_to_conserved()
    return s


def device_config(sim: driver.Simulation, limiter: int, solver: int, cfl: float, tf: float = float("inf"),
                  max_steps: int = -1, fixed_dt: float = 0.0, tol: float = 1e-6, strip_rows: int = 0,
                  hllc_form: int = 0, full_tiles: int = 0):
    """hfv_mh2d_config for a 2D Simulation (flags as in the reference headers)."""
    from . import gpu
    from .abi import BOUNDARY_CONDITION_NONE
    sy = sim.system
    cfg = gpu.MH2DConfig(
        num_cells_x=sim.nx, num_cells_y=sim.ny, num_ghost_cells_side=sim.g, gamma=sy.gamma,
        gravity=sy.gravity, dx=sy.dx_, dy=sy.dy_, area_x=sy.surface_area_x_[0], area_y=sy.surface_area_y_[0],
        volume=sy.volume_[0], slope_limiter=limiter, riemann_solver=solver, tol=tol, cfl=cfl,
        cfl_initial_shrink_factor=0.2, cfl_initial_shrink_num_steps=10, tf=tf, max_steps=max_steps,
        fixed_dt=fixed_dt, strip_rows=strip_rows, hllc_form=hllc_form, full_tiles=full_tiles)
    flags = [sim.bc.boundary_condition_flag_x_min_, sim.bc.boundary_condition_flag_x_max_,
             sim.bc.boundary_condition_flag_y_min_, sim.bc.boundary_condition_flag_y_max_]
    for k, f in enumerate(flags):
        cfg.bc[k] = f if f > 0 else BOUNDARY_CONDITION_NONE
    return cfg
