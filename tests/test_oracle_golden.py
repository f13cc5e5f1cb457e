"""The CPU oracle against the committed golden fixtures (tests/golden/), which
were produced by the reference compiled from its own sources
(tests/golden/make_golden.py).  Bitwise equality is required throughout:
the oracle keeps the reference's IEEE operation sequence.

Also pinned here, independently of the reference build:
  * the exact solver against the reference's exact Sod solution at t=0.2
    (examples/sod_shock_1d/sol_cartesian_1d_0.2.npz, copied as data);
  * closed forms of the limiters;
  * exact conservation of the 2D update with periodic boundaries;
  * mirror symmetry of a symmetric problem.
"""
import os

import numpy as np
import pytest

import _cases as K
import _oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def eq(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b), equal_nan=True)


# -- per-interface functions ---------------------------------------------
def test_limiters_golden():
    d = load("scalar_vectors.npz")
    for name in K.LIMITERS:
        assert eq(O.limit_batch(name, d["lim_a"], d["lim_b"]), d[f"lim_{name}"]), name


def test_limiter_closed_forms():
    a = np.array([1.0, 2.0, -3.0, 0.5, 1.0, -1.0, 0.0])
    b = np.array([2.0, 1.0, -1.0, -0.5, 1.0, -1.0, 4.0])
    assert eq(O.limit_batch("minmod", a, b), [1.0, 1.0, -1.0, 0.0, 1.0, -1.0, 0.0])
    vl = O.limit_batch("van_leer", a, b)
    assert np.allclose(vl[:3], [2 * 1 * 2 / 3, 2 * 2 * 1 / 3, 2 * (-3) * (-1) / (-4)], rtol=1e-9)
    assert eq(O.limit_batch("monotonized_central", a, b), [1.5, 1.5, -2.0, 0.0, 1.0, -1.0, 0.0])


@pytest.mark.parametrize("kind,dim,solver,gamma,tol", [
    ("hllc2d", 2, "riemann_solver_hllc", 1.4, 1e-6),
    ("hllc1d", 1, "riemann_solver_hllc", 1.4, 1e-6),
    ("exact1d", 1, "riemann_solver_exact", 1.4, 1e-6),
    ("exact1d_g53_tol8", 1, "riemann_solver_exact", 5.0 / 3.0, 1e-8),
])
def test_fluxes_golden(kind, dim, solver, gamma, tol):
    d = load("scalar_vectors.npz")
    wl, wr = d["wl"], d["wr"]
    if dim == 1:
        wl, wr = wl[:, [0, 1, 3]], wr[:, [0, 1, 3]]
    f, err = O.flux_batch(dim, solver, gamma, tol, wl, wr)
    assert np.array_equal(err, d[f"{kind}_err"])
    ok = err == 0
    assert eq(f[ok], d[f"{kind}_f"][ok])
    assert (d[f"{kind}_err"] != 0).sum() >= 1  # the NaN-pressure case fails


def test_exact_flux_2d_composition():
    """The 2D exact flux (no reference counterpart): mass and normal momentum
    equal the 1D exact flux bitwise, the energy equals it when v = 0, and the
    transverse momentum flux is the mass flux times the upwinded v."""
    wl, wr = K.riemann_states(2000, 5)
    f2, e2 = O.flux_batch(2, "riemann_solver_exact", 1.4, 1e-6, wl, wr)
    f1, e1 = O.flux_batch(1, "riemann_solver_exact", 1.4, 1e-6, wl[:, [0, 1, 3]], wr[:, [0, 1, 3]])
    assert np.array_equal(e1, e2)
    ok = e1 == 0
    assert eq(f2[ok][:, :2], f1[ok][:, :2])
    z_l, z_r = wl.copy(), wr.copy()
    z_l[:, 2] = z_r[:, 2] = 0.0
    f0, _ = O.flux_batch(2, "riemann_solver_exact", 1.4, 1e-6, z_l, z_r)
    assert eq(f0[ok][:, 3], f1[ok][:, 2])
    # transverse flux: either v_L or v_R times the mass flux
    m = ok & ~np.isnan(f2[:, 0]) & (f2[:, 0] != 0)
    ratio = f2[m, 2] / f2[m, 0]
    assert np.all((ratio == wl[m, 2]) | (ratio == wr[m, 2]) | np.isclose(ratio, wl[m, 2]) | np.isclose(ratio, wr[m, 2]))


def test_exact_solver_reproduces_reference_sod_solution():
    """examples/sod_shock_1d/sol_cartesian_1d_0.2.npz is the exact solution at
    t=0.2 sampled at x = 0, 0.001, ..., 0.999 with tol 1e-6: bitwise."""
    d = load("sod_exact_t0.2.npz")
    x = d["mid_points"]
    sol = O.oracle_exact_sample(1.4, 1e-6, [1.0, 0.0, 1.0], [0.125, 0.0, 0.1], (x - 0.5) / 0.2)
    assert eq(sol[:, 0], d["density"])
    assert eq(sol[:, 1], d["velocity"])
    assert eq(sol[:, 2], d["pressure"])


# -- full runs -----------------------------------------------------------
def _grid_from(d, key, coord, nx, ny, g, bcs, limiter, solver="riemann_solver_hllc", gravity=0.0, x=(0, 1)):
    grid = O.Grid(coord, nx, ny, g, float(d[f"{key}/gamma"]) if f"{key}/gamma" in d.files else 1.4, x=x,
                  bc=bcs, limiter=limiter, solver=solver, gravity=gravity)
    grid.W[...] = d[f"{key}/W0"]
    grid.U[...] = d[f"{key}/U0"]
    return grid


@pytest.mark.parametrize("shape", K.SHAPES, ids=lambda s: f"{s[0]}x{s[1]}g{s[2]}")
@pytest.mark.parametrize("bname", list(K.BC_SETS))
@pytest.mark.parametrize("limiter", K.LIMITERS)
def test_2d_runs_golden(shape, bname, limiter):
    d = load("runs_2d.npz")
    nx, ny, g = shape
    key = f"{nx}x{ny}g{g}_{bname}_{limiter}"
    grid = _grid_from(d, key, "cartesian_2d", nx, ny, g, K.BC_SETS[bname], limiter)
    r = grid.run(0.8, max_steps=11)
    assert r.error == 0
    assert (r.num_steps, r.t, r.dt) == (int(d[f"{key}/n"]), float(d[f"{key}/t"]), float(d[f"{key}/dt"]))
    assert eq(grid.U, d[f"{key}/U1"])
    assert eq(grid.W, d[f"{key}/W1"])


def test_kh64_golden():
    d = load("runs_2d.npz")
    key = "kh64_mc_50"
    grid = _grid_from(d, key, "cartesian_2d", 64, 64, 2, ("periodic",) * 4, "monotonized_central")
    r = grid.run(0.8, max_steps=50)
    assert (r.num_steps, r.t) == (50, float(d[f"{key}/t"]))
    assert eq(grid.U, d[f"{key}/U1"]) and eq(grid.W, d[f"{key}/W1"])


def test_gravity_golden():
    d = load("runs_2d.npz")
    key = "gravity_13x9_vl"
    grid = _grid_from(d, key, "cartesian_2d", 13, 9, 2, K.BC_SETS["reflective"], "van_leer", gravity=0.7)
    r = grid.run(0.8, max_steps=11)
    assert r.t == float(d[f"{key}/t"])
    assert eq(grid.U, d[f"{key}/U1"]) and eq(grid.W, d[f"{key}/W1"])


@pytest.mark.parametrize("key,coord,n,g,bcs,limiter,solver,cfl,tf", [
    ("sod1000_hllc_minmod", "cartesian_1d", 1000, 2, ("transmissive",) * 2, "minmod", "riemann_solver_hllc", 0.9, 0.2),
    ("sod200_exact_vl", "cartesian_1d", 200, 2, ("transmissive",) * 2, "van_leer", "riemann_solver_exact", 0.9, 0.2),
    ("cyl64_hllc_mc", "cylindrical_1d", 64, 2, ("reflective", "transmissive"), "monotonized_central",
     "riemann_solver_hllc", 0.9, 0.2),
    ("sph64_hllc_mc", "spherical_1d", 64, 2, ("reflective", "transmissive"), "monotonized_central",
     "riemann_solver_hllc", 0.9, 0.2),
    ("sod50g3_periodic_minmod", "cartesian_1d", 50, 3, ("periodic",) * 2, "minmod", "riemann_solver_hllc", 0.8, 0.05),
])
def test_1d_runs_golden(key, coord, n, g, bcs, limiter, solver, cfl, tf):
    d = load("runs_1d.npz")
    grid = O.Grid(coord, n, 1, g, 1.4, bc=bcs, limiter=limiter, solver=solver)
    grid.W[...] = d[f"{key}/W0"]
    grid.U[...] = d[f"{key}/U0"]
    r = grid.run(cfl, tf=tf)
    assert r.error == 0
    assert (r.num_steps, r.t, r.dt) == (int(d[f"{key}/n"]), float(d[f"{key}/t"]), float(d[f"{key}/dt"]))
    assert eq(grid.U, d[f"{key}/U1"]) and eq(grid.W, d[f"{key}/W1"])


def test_sod_1000_steps_and_accuracy():
    """Sod N=1000 to t=0.2: 497 steps, and the L1 errors against the exact
    solution stay at the level of a second-order scheme on 1000 cells."""
    d = load("runs_1d.npz")
    assert int(d["sod1000_hllc_minmod/n"]) == 497
    W = d["sod1000_hllc_minmod/W1"]
    g, n = 2, 1000
    x = (np.arange(n + 2 * g) + 0.5) / (n + 2 * g)
    xi = x[g:g + n]
    sol = O.oracle_exact_sample(1.4, 1e-12, [1.0, 0.0, 1.0], [0.125, 0.0, 0.1], (xi - 0.5) / 0.2)
    h = 1.0 / (n + 2 * g)
    for k in range(3):
        l1 = np.sum(np.abs(W[k][g:g + n] - sol[:, k])) * h
        assert l1 < 1.2e-3, (k, l1)


def test_periodic_update_conserves_mass_and_momentum():
    """With periodic boundaries every interface flux enters one cell and
    leaves another, so the interior totals change only by rounding."""
    sim_lib = None
    nx = ny = 24
    grid = O.Grid("cartesian_2d", nx, ny, 2, 1.4, bc=("periodic",) * 4, limiter="monotonized_central")
    W = K.random_state_2d(nx, ny, 3)
    for k in range(4):
        grid.W[k][2:2 + ny, 2:2 + nx] = W[k]
    grid.apply_bc()
    grid.p2c()
    before = grid.U[:, 2:2 + ny, 2:2 + nx].sum(axis=(1, 2))
    r = grid.run(0.8, max_steps=20)
    assert r.error == 0
    after = grid.U[:, 2:2 + ny, 2:2 + nx].sum(axis=(1, 2))
    assert np.allclose(after, before, rtol=0, atol=1e-12 * np.abs(before).max())


def test_y_only_problem_keeps_rows_and_mirror_symmetry():
    """A shock tube along y (u = 0, state symmetric about y = 0.5) keeps every
    row identical bitwise and stays mirror-symmetric in y to round-off.  (A
    general 2D problem does not keep its mirror symmetry: the y-sweep density
    predictor multiplies the x-velocity slope, integrator_muscl_hancock.c:979.)"""
    nx, ny = 6, 40
    grid = O.Grid("cartesian_2d", nx, ny, 2, 1.4, bc=("periodic", "periodic", "reflective", "reflective"),
                  limiter="monotonized_central")
    yc = (np.arange(ny) + 0.5) / ny
    inner = np.abs(yc - 0.5) < 0.2
    rho = np.where(inner, 1.0, 0.125)[:, None] * np.ones((1, nx))
    p = np.where(inner, 1.0, 0.1)[:, None] * np.ones((1, nx))
    z = np.zeros((ny, nx))
    for k, a in enumerate([rho, z, z, p]):
        grid.W[k][2:2 + ny, 2:2 + nx] = a
    grid.apply_bc()
    grid.p2c()
    r = grid.run(0.8, max_steps=40)
    assert r.error == 0
    rho = grid.W[0][2:2 + ny, 2:2 + nx]
    assert all(np.array_equal(rho[:, 0], rho[:, i]) for i in range(nx))
    assert np.allclose(rho, rho[::-1, :], rtol=1e-12, atol=0)
    assert np.abs(grid.W[1][2:2 + ny, 2:2 + nx]).max() == 0.0


def test_x_sweep_matches_1d_path():
    """A 2D problem without y-variation (v = 0) run at fixed dt reproduces the
    1D MUSCL-Hancock path row by row (the two differ only in the rounding of
    dt*A/V and of the extensive variables)."""
    n = 80
    g2 = O.Grid("cartesian_2d", n, 4, 2, 1.4, bc=("transmissive", "transmissive", "periodic", "periodic"),
                limiter="van_leer")
    g1 = O.Grid("cartesian_1d", n, 1, 2, 1.4, bc=("transmissive", "transmissive"), limiter="van_leer")
    x = (np.arange(n) + 0.5) / n
    rho = np.where(x < 0.5, 1.0, 0.125)
    p = np.where(x < 0.5, 1.0, 0.1)
    for k, a in enumerate([rho, np.zeros(n), np.zeros(n), p]):
        g2.W[k][2:6, 2:2 + n] = a
    for k, a in enumerate([rho, np.zeros(n), p]):
        g1.W[k][2:2 + n] = a
    g2.apply_bc(); g2.p2c(); g1.apply_bc(); g1.p2c()
    dt = 0.4 * g1.p.dx
    assert g2.run(0.5, max_steps=60, fixed_dt=dt).error == 0
    assert g1.run(0.5, max_steps=60, fixed_dt=dt).error == 0
    for k2, k1 in ((0, 0), (1, 1), (3, 2)):
        a = g2.W[k2][3, 2:2 + n]
        b = g1.W[k1][2:2 + n]
        assert np.allclose(a, b, rtol=1e-11, atol=1e-13), k2
