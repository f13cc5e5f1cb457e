"""Device parity of the 1D MUSCL-Hancock path (muscl_hancock_1d) against the
CPU oracle: the Sod shock tube of the benchmark set, the curvilinear
geometric source term, the exact solver and every boundary condition."""
import numpy as np
import pytest

import _cases as K
from test_gpu_parity_2d import compare

pytestmark = pytest.mark.gpu


def run_1d(gpulib, setup, limiter, solver, cfl, tf=np.inf, steps=None, tol=1e-6, exact_time=True):
    sim = setup(gpulib)
    grid = K.oracle_grid(sim, limiter, solver, tol=tol)
    if steps is None:
        res = sim.launch("muscl_hancock_1d", solver, limiter, cfl, tf, tol=tol)
        r = grid.run(cfl, tf=tf)
    else:
        res = sim.launch("muscl_hancock_1d", solver, limiter, cfl, tf, tol=tol, max_steps=steps)
        r = grid.run(cfl, tf=tf, max_steps=steps)
    assert r.error == 0
    if exact_time:
        assert (res.num_steps, res.t, res.dt) == (r.num_steps, r.t, r.dt)
    else:
        assert res.num_steps == r.num_steps and res.t == r.t
        assert abs(res.dt - r.dt) <= 1e-10 * r.dt
    W, U = K.state(sim)
    return W, U, grid.W, grid.U, res


def test_sod_1000_hllc_minmod_bitwise(gpulib):
    """Sod, N=1000, HLLC + minmod, CFL 0.9, to t=0.2 (497 steps); the device
    result also matches the exact Riemann solution: L1 errors of rho, u and p
    (the oracle's exact sampler at tol 1e-12, cell centres) at the level of a
    second-order scheme on 1000 cells (about 9e-4, 9e-4 and 5e-4)."""
    import _oracle as O
    W, U, Wo, Uo, res = run_1d(gpulib, lambda lib: K.setup_sod(lib, 1000), "minmod", "riemann_solver_hllc",
                               0.9, tf=0.2)
    assert res.num_steps == 497
    compare("sod", W[:, None], U[:, None], Wo[:, None], Uo[:, None])
    g, n = 2, 1000
    xi = ((np.arange(n + 2 * g) + 0.5) / (n + 2 * g))[g:g + n]
    sol = O.oracle_exact_sample(1.4, 1e-12, [1.0, 0.0, 1.0], [0.125, 0.0, 0.1], (xi - 0.5) / 0.2)
    h = 1.0 / (n + 2 * g)
    l1 = [float(np.sum(np.abs(W[k][g:g + n] - sol[:, k])) * h) for k in range(3)]
    print("Sod L1 vs exact (rho, u, p):", l1)
    assert all(e < 1.2e-3 for e in l1), l1


@pytest.mark.parametrize("limiter", K.LIMITERS)
def test_sod_exact_solver(gpulib, limiter):
    """Exact solver in 1D: pow() may differ from glibc in the last place."""
    W, U, Wo, Uo, res = run_1d(gpulib, lambda lib: K.setup_sod(lib, 200), limiter, "riemann_solver_exact",
                               0.9, tf=0.2, exact_time=False)
    compare("sod exact", W[:, None], U[:, None], Wo[:, None], Uo[:, None], bitwise=False, tol=1e-10)


@pytest.mark.parametrize("coord", ["cylindrical_1d", "spherical_1d"])
def test_curvilinear_geometric_source(gpulib, coord):
    setup = lambda lib: K.setup_sod(lib, 64, coord=coord, bc=("reflective", "transmissive"))
    W, U, Wo, Uo, res = run_1d(gpulib, setup, "monotonized_central", "riemann_solver_hllc", 0.9, tf=0.2)
    compare(coord, W[:, None], U[:, None], Wo[:, None], Uo[:, None])


@pytest.mark.parametrize("bc", [("periodic", "periodic"), ("reflective", "reflective"), ("none", "transmissive")])
@pytest.mark.parametrize("limiter", K.LIMITERS)
def test_boundaries_and_limiters(gpulib, bc, limiter):
    setup = lambda lib: K.setup_sod(lib, 50, g=3, bc=bc)
    W, U, Wo, Uo, res = run_1d(gpulib, setup, limiter, "riemann_solver_hllc", 0.8, steps=40)
    compare(f"{bc} {limiter}", W[:, None], U[:, None], Wo[:, None], Uo[:, None])
