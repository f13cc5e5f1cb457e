"""The oracle's first-order Godunov loops (godunov_first_order_{1d,2d},
integrator_godunov_first_order.c) against the committed fixtures of the
reference build (tests/golden/runs_godunov.npz), bitwise; and, independently
of the reference, against the exact Riemann solution (first-order
convergence), exact conservation and the 1D path."""
import os

import numpy as np
import pytest

import _cases as K
import _oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SHAPES = [(13, 7, 1), (9, 11, 2)]


def load():
    return np.load(os.path.join(GOLD, "runs_godunov.npz"), allow_pickle=False)


def eq(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b), equal_nan=True)


def grid_from(d, key, coord, nx, ny, g, bcs, solver="riemann_solver_hllc", gravity=0.0):
    grid = O.Grid(coord, nx, ny, g, float(d[f"{key}/gamma"]) if f"{key}/gamma" in d.files else 1.4,
                  bc=bcs, solver=solver, gravity=gravity)
    grid.W[...] = d[f"{key}/W0"]
    grid.U[...] = d[f"{key}/U0"]
    return grid


def check(grid, d, key, r):
    assert r.error == 0
    assert (r.num_steps, r.t, r.dt) == (int(d[f"{key}/n"]), float(d[f"{key}/t"]), float(d[f"{key}/dt"]))
    assert eq(grid.U, d[f"{key}/U1"])
    assert eq(grid.W, d[f"{key}/W1"])


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"{s[0]}x{s[1]}g{s[2]}")
@pytest.mark.parametrize("bname", list(K.BC_SETS))
def test_2d_godunov_golden(shape, bname):
    d = load()
    nx, ny, g = shape
    key = f"2d_{nx}x{ny}g{g}_{bname}"
    grid = grid_from(d, key, "cartesian_2d", nx, ny, g, K.BC_SETS[bname])
    check(grid, d, key, grid.run(0.8, max_steps=11, first_order=True))


def test_2d_godunov_kh_and_gravity_golden():
    d = load()
    grid = grid_from(d, "2d_kh64_40", "cartesian_2d", 64, 64, 2, ("periodic",) * 4)
    check(grid, d, "2d_kh64_40", grid.run(0.8, max_steps=40, first_order=True))
    grid = grid_from(d, "2d_gravity_13x9g1", "cartesian_2d", 13, 9, 1, K.BC_SETS["reflective"], gravity=0.7)
    check(grid, d, "2d_gravity_13x9g1", grid.run(0.8, max_steps=11, first_order=True))


GODUNOV_1D = [
    ("1d_sod400_hllc", "cartesian_1d", 400, 2, ("transmissive",) * 2, "riemann_solver_hllc", 0.9, 0.2),
    ("1d_sod200_exact", "cartesian_1d", 200, 2, ("transmissive",) * 2, "riemann_solver_exact", 0.9, 0.2),
    ("1d_cyl64_hllc", "cylindrical_1d", 64, 2, ("reflective", "transmissive"), "riemann_solver_hllc", 0.9, 0.2),
    ("1d_sph64_exact", "spherical_1d", 64, 2, ("reflective", "transmissive"), "riemann_solver_exact", 0.9, 0.2),
    ("1d_sod50g1_periodic", "cartesian_1d", 50, 1, ("periodic",) * 2, "riemann_solver_hllc", 0.8, 0.05),
]


@pytest.mark.parametrize("key,coord,n,g,bcs,solver,cfl,tf", GODUNOV_1D, ids=[c[0] for c in GODUNOV_1D])
def test_1d_godunov_golden(key, coord, n, g, bcs, solver, cfl, tf):
    d = load()
    grid = O.Grid(coord, n, 1, g, 1.4, bc=bcs, solver=solver)
    grid.W[...] = d[f"{key}/W0"]
    grid.U[...] = d[f"{key}/U0"]
    check(grid, d, key, grid.run(cfl, tf=tf, first_order=True))


def sod_l1(n, solver):
    """L1 errors (rho, u, p) of first-order Godunov on Sod at t = 0.2 against
    the exact solution (the oracle's exact sampler at tol 1e-12)."""
    g = 1
    grid = O.Grid("cartesian_1d", n, 1, g, 1.4, bc=("transmissive",) * 2, solver=solver)
    x = (np.arange(n + 2 * g) + 0.5) / (n + 2 * g)
    left = x < 0.5
    grid.W[0] = np.where(left, 1.0, 0.125)
    grid.W[1] = 0.0
    grid.W[2] = np.where(left, 1.0, 0.1)
    grid.p2c()
    assert grid.run(0.9, tf=0.2, first_order=True).error == 0
    xi = x[g:g + n]
    sol = O.oracle_exact_sample(1.4, 1e-12, [1.0, 0.0, 1.0], [0.125, 0.0, 0.1], (xi - 0.5) / 0.2)
    h = 1.0 / (n + 2 * g)
    return np.array([np.sum(np.abs(grid.W[k][g:g + n] - sol[:, k])) * h for k in range(3)])


@pytest.mark.parametrize("solver", ["riemann_solver_hllc", "riemann_solver_exact"])
def test_sod_converges_at_first_order(solver):
    """Doubling the resolution cuts the L1 error of a shock tube by a factor
    between 2^0.5 and 2^1 (first order at the contact and rarefaction, with
    the shock's contribution falling as 1/N)."""
    e1, e2 = sod_l1(200, solver), sod_l1(400, solver)
    ratio = e1 / e2
    assert np.all(ratio > 2 ** 0.5) and np.all(ratio < 2.2), ratio
    assert e2[0] < 1.2e-2, e2


def test_godunov_2d_conserves_with_periodic_boundaries():
    nx = ny = 20
    grid = O.Grid("cartesian_2d", nx, ny, 1, 1.4, bc=("periodic",) * 4)
    W = K.random_state_2d(nx, ny, 5)
    for k in range(4):
        grid.W[k][1:1 + ny, 1:1 + nx] = W[k]
    grid.apply_bc()
    grid.p2c()
    before = grid.U[:, 1:1 + ny, 1:1 + nx].sum(axis=(1, 2))
    assert grid.run(0.8, max_steps=25, first_order=True).error == 0
    after = grid.U[:, 1:1 + ny, 1:1 + nx].sum(axis=(1, 2))
    assert np.allclose(after, before, rtol=0, atol=1e-12 * np.abs(before).max())


def test_godunov_x_sweep_matches_1d_path():
    """Without y-variation (v = 0) the 2D first-order step reproduces the 1D
    one row by row (they differ only in the rounding of the extensive
    variables)."""
    n = 60
    g2 = O.Grid("cartesian_2d", n, 3, 1, 1.4, bc=("transmissive", "transmissive", "periodic", "periodic"))
    g1 = O.Grid("cartesian_1d", n, 1, 1, 1.4, bc=("transmissive", "transmissive"))
    x = (np.arange(n) + 0.5) / n
    rho = np.where(x < 0.5, 1.0, 0.125)
    p = np.where(x < 0.5, 1.0, 0.1)
    for k, a in enumerate([rho, np.zeros(n), np.zeros(n), p]):
        g2.W[k][1:4, 1:1 + n] = a
    for k, a in enumerate([rho, np.zeros(n), p]):
        g1.W[k][1:1 + n] = a
    g2.apply_bc(); g2.p2c(); g1.apply_bc(); g1.p2c()
    dt = 0.4 * g1.p.dx
    assert g2.run(0.5, max_steps=50, fixed_dt=dt, first_order=True).error == 0
    assert g1.run(0.5, max_steps=50, fixed_dt=dt, first_order=True).error == 0
    for k2, k1 in ((0, 0), (1, 1), (3, 2)):
        assert np.allclose(g2.W[k2][2, 1:1 + n], g1.W[k1][1:1 + n], rtol=1e-11, atol=1e-13), k2
