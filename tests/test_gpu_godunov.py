"""Device parity of the first-order Godunov integrators
(godunov_first_order_{1d,2d}, integrator_godunov_first_order.c) against the
CPU oracle, through the library's C API (launch_simulation and the
step-limited hfv_godunov_first_order_*_run).  The faces take the cells' own
states, so the arithmetic is the MUSCL-Hancock path's without reconstruction:
bitwise except where pow() runs (two-rarefaction guess, exact solver,
vacuum samplers), and within 1e-12 relative everywhere (1e-10 for the exact
solver's shock tubes)."""
import numpy as np
import pytest

import _cases as K
from hydrofv_amd import driver
from test_gpu_parity_2d import compare

pytestmark = pytest.mark.gpu

SHAPES = [(13, 7, 1), (9, 11, 2), (130, 70, 1)]


def run_pair_2d(gpulib, setup, steps, cfl=0.8, solver="riemann_solver_hllc", tol=1e-6, exact_time=True):
    sim = setup(gpulib)
    grid = K.oracle_grid(sim, "minmod", solver, tol=tol)
    res = sim.launch("godunov_first_order_2d", solver, "minmod", cfl, tf=np.inf, tol=tol, max_steps=steps)
    r = grid.run(cfl, max_steps=steps, first_order=True)
    assert r.error == 0
    assert res.num_steps == r.num_steps == steps
    if exact_time:
        assert (res.t, res.dt) == (r.t, r.dt)
    else:
        assert abs(res.t - r.t) <= 1e-12 * r.t
    W, U = K.state(sim)
    return sim, W, U, grid.W, grid.U


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"{s[0]}x{s[1]}g{s[2]}")
@pytest.mark.parametrize("bcs", list(K.BC_SETS), ids=str)
def test_godunov_2d_grids(gpulib, shape, bcs):
    """One and two ghost layers; 130 columns make three strips, the last one
    partly filled, and 70 rows several row strips."""
    nx, ny, g = shape
    setup = lambda lib: K.setup_random_2d(lib, nx, ny, g, K.BC_SETS[bcs], 500 + nx + ny + g)
    sim, W, U, Wo, Uo = run_pair_2d(gpulib, setup, 11)
    compare(f"godunov {shape} {bcs}", W, U, Wo, Uo)
    sim.close()


def test_godunov_2d_kh_to_tf(gpulib):
    """Kelvin-Helmholtz 128^2 with the reference's run-to-tf loop."""
    sim = K.setup_kh(gpulib, 128)
    grid = K.oracle_grid(sim, "minmod", "riemann_solver_hllc")
    res = sim.launch("godunov_first_order_2d", "riemann_solver_hllc", "minmod", 0.8, 0.05)
    r = grid.run(0.8, tf=0.05, first_order=True)
    assert r.error == 0 and (res.num_steps, res.t, res.dt) == (r.num_steps, r.t, r.dt)
    W, U = K.state(sim)
    compare("godunov KH 128 to tf", W, U, grid.W, grid.U)


def test_godunov_2d_gravity(gpulib):
    setup = lambda lib: K.setup_random_2d(lib, 13, 9, 1, K.BC_SETS["reflective"], 12, gravity=0.7)
    sim, W, U, Wo, Uo = run_pair_2d(gpulib, setup, 11)
    compare("godunov gravity", W, U, Wo, Uo)


def test_godunov_2d_exact_solver(gpulib):
    """The exact solver in 2D (an extension: the reference's 2D flux
    dispatch accepts HLLC only), first order, on a small Sedov blast."""
    sim, W, U, Wo, Uo = run_pair_2d(gpulib, lambda lib: K.setup_sedov(lib, 32), 15, cfl=0.5,
                                    solver="riemann_solver_exact", tol=1e-8, exact_time=False)
    from test_gpu_parity_2d import interior_and_bc_mask
    compare("godunov sedov 32 exact", W, U, Wo, Uo, tol=1e-10, mask=interior_and_bc_mask(sim))


GODUNOV_1D = [
    ("sod400_hllc", lambda L: K.setup_sod(L, 400), "riemann_solver_hllc", 0.9, 0.2, 1e-12),
    ("sod200_exact", lambda L: K.setup_sod(L, 200), "riemann_solver_exact", 0.9, 0.2, 1e-10),
    ("cyl64_hllc", lambda L: K.setup_sod(L, 64, coord="cylindrical_1d", bc=("reflective", "transmissive")),
     "riemann_solver_hllc", 0.9, 0.2, 1e-12),
    ("sph64_exact", lambda L: K.setup_sod(L, 64, coord="spherical_1d", bc=("reflective", "transmissive")),
     "riemann_solver_exact", 0.9, 0.2, 1e-10),
    ("sod50g1_periodic", lambda L: K.setup_sod(L, 50, g=1, bc=("periodic", "periodic")),
     "riemann_solver_hllc", 0.8, 0.05, 1e-12),
    ("sod50g1_none", lambda L: K.setup_sod(L, 50, g=1, bc=("none", "reflective")),
     "riemann_solver_hllc", 0.8, 0.05, 1e-12),
]


@pytest.mark.parametrize("key,setup,solver,cfl,tf,tol", GODUNOV_1D, ids=[c[0] for c in GODUNOV_1D])
def test_godunov_1d(gpulib, key, setup, solver, cfl, tf, tol):
    sim = setup(gpulib)
    grid = K.oracle_grid(sim, "minmod", solver)
    res = sim.launch("godunov_first_order_1d", solver, "minmod", cfl, tf)
    r = grid.run(cfl, tf=tf, first_order=True)
    assert r.error == 0
    assert res.num_steps == r.num_steps and res.t == r.t
    assert abs(res.dt - r.dt) <= tol * r.dt
    W, U = K.state(sim)
    compare(f"godunov 1d {key}", W[:, None], U[:, None], grid.W[:, None], grid.U[:, None], tol=tol)


def test_godunov_1d_reports_a_failed_flux(gpulib):
    """Unlike muscl_hancock_1d, godunov_first_order_1d returns a failed
    flux's error (integrator_godunov_first_order.c:172-199): a NaN pressure
    in one cell makes the HLLC wave-speed tests fail."""
    sim = K.setup_sod(gpulib, 40)
    W = np.stack(sim.primitive())
    W[2][10] = np.nan
    sim.set_primitive(list(W))
    sim.primitive_to_conserved()
    with pytest.raises(driver.HydroError) as e:
        sim.launch("godunov_first_order_1d", "riemann_solver_hllc", "minmod", 0.9, 0.2)
    assert e.value.code == 2
    assert "Invalid wave speed for the HLLC riemann solver" in e.value.text


def test_godunov_2d_reports_a_failed_flux(gpulib):
    sim = driver.Simulation(gpulib, "cartesian_2d", 1.4, 8, 8, 1)
    sim.set_boundary("periodic", "periodic", "periodic", "periodic")
    W = K.random_state_2d(8, 8, 9)
    W[3][4, 4] = np.nan
    sim.set_primitive(W, interior_only=True)
    sim.apply_boundary()
    sim.primitive_to_conserved()
    with pytest.raises(driver.HydroError) as e:
        sim.launch("godunov_first_order_2d", "riemann_solver_hllc", "minmod", 0.5, tf=1.0)
    assert e.value.code == 2
    assert "Invalid wave speed for the HLLC riemann solver" in e.value.text
