"""Device parity of the 2D MUSCL-Hancock path against the CPU oracle.

Each case builds the initial state through the library's host API, runs the
device integrator through the C ABI (muscl_hancock_2d / hfv_muscl_hancock_2d_run)
and the oracle on a copy of the same state, and compares every cell
(interior, ghost cells filled by the boundary condition, ghost conserved
cells and corners alike).  The acceptance tolerance of the path is 1e-12
relative.  The device reproduces the reference's IEEE operation sequence, so
results are bitwise equal except where pow() is evaluated (the
two-rarefaction pressure guess, the exact solver, the vacuum samplers): the
device libm's pow and glibc's can differ in the last place.  Bitwise
equality is asserted where pow cannot occur (probe states outside those
branches) and reported otherwise.
"""
import numpy as np
import pytest

import _cases as K
from hydrofv_amd import driver, gpu

pytestmark = pytest.mark.gpu

TOL = 1e-12  # max |a-b| / max(|b|, tiny) per field


def test_probe_limiters_bitwise(gpulib):
    from hydrofv_amd import gpu
    import _oracle as O
    wl, wr = K.riemann_states(4000, 3)
    a = np.concatenate([wl[:, 1] - wr[:, 1], wl[:, 3] - wr[:, 3], [0.0, -0.0, 1.0, -1.0, 2.0, 1e-300, -5e-11, 3.0]])
    b = np.concatenate([wr[:, 2], wr[:, 0] - wl[:, 0], [1.0, 1.0, 1.0, -1.0, -2.0, 1e-300, -5e-11, 3.0]])
    for k, name in enumerate(K.LIMITERS, start=1):
        dev = gpu.probe_limit(gpulib, k, a, b)
        ref = O.limit_batch(name, a, b)
        bad = np.flatnonzero((dev != ref) & ~(np.isnan(dev) & np.isnan(ref)))
        assert bad.size == 0, (name, [(a[i].hex(), b[i].hex(), dev[i].hex(), ref[i].hex()) for i in bad[:5]])


@pytest.mark.parametrize("nf,solver", [(4, "riemann_solver_hllc"), (3, "riemann_solver_hllc"),
                                       (3, "riemann_solver_exact"), (4, "riemann_solver_exact")])
def test_probe_fluxes(gpulib, nf, solver):
    """Device flux functions against the oracle on branch-covering states:
    HLLC bitwise; the exact solver within 1e-12 (pow rounding)."""
    from hydrofv_amd import gpu
    import _oracle as O
    wl, wr = K.riemann_states(4000, 17)
    if nf == 3:
        wl, wr = wl[:, [0, 1, 3]], wr[:, [0, 1, 3]]
    sid = 2 if solver == "riemann_solver_hllc" else 1
    fd, ed = gpu.probe_flux(gpulib, nf, sid, 1.4, 1e-6, wl, wr)
    fo, eo = O.flux_batch(2 if nf == 4 else 1, solver, 1.4, 1e-6, wl, wr)
    # device failure kinds -> the reference's return codes (VALUE_ERROR, FAILURE)
    ed_code = np.select([ed == 0, ed == 2, (ed == 3) | (ed == 4)], [0, 2, 1], default=-1)
    assert np.array_equal(ed_code, eo), [(i, ed[i], eo[i]) for i in np.flatnonzero(ed_code != eo)[:5]]
    ok = eo == 0
    fin = np.isfinite(fo) & ok[:, None]
    assert np.all(np.abs(fd[fin] - fo[fin]) <= 1e-12 * np.maximum(np.abs(fo[fin]), 1.0))
    assert np.array_equal(np.isnan(fd[ok]), np.isnan(fo[ok]))
    if solver == "riemann_solver_hllc":
        # bitwise wherever neither the two-rarefaction guess nor a vacuum sampler (pow) runs
        L, R = (wl, wr) if nf == 4 else (wl[:, [0, 1, 2, 2]], wr[:, [0, 1, 2, 2]])
        rl, ul, pl = wl[:, 0], wl[:, 1], wl[:, -1]
        rr, ur, pr = wr[:, 0], wr[:, 1], wr[:, -1]
        with np.errstate(invalid="ignore", divide="ignore"):
            al = np.sqrt(1.4 * pl / rl)
            ar = np.sqrt(1.4 * pr / rr)
            ppv = 0.5 * (pl + pr) - 0.125 * (ur - ul) * (rl + rr) * (al + ar)
            vac = (rl <= 0) | (rr <= 0) | ((al + ar) * 2.0 / 2.4 <= (ur - ul))
            trrs = ppv < np.minimum(pl, pr)
        pow_free = ok & ~vac & ~trrs & np.isfinite(ppv)
        eq = np.all((fd == fo) | (np.isnan(fd) & np.isnan(fo)), axis=1)
        bad = np.flatnonzero(pow_free & ~eq)
        assert bad.size == 0, [(wl[i].tolist(), wr[i].tolist(), fd[i].tolist(), fo[i].tolist()) for i in bad[:3]]
        assert pow_free.sum() > 1000


def _ratio_two_states(seed, n, plo_range):
    """Face states whose pressure ratio p_max / p_min lies within a few ulps
    of 2 (both orientations), with u_L = u_R so that the PVRS value lies
    between the pressures and the ratio test alone picks the guess."""
    rng = np.random.default_rng(seed)
    plo = np.exp(rng.uniform(np.log(plo_range[0]), np.log(plo_range[1]), n))
    k = rng.integers(-4, 5, n)
    phi = ((2.0 * plo).view(np.int64) + k).view(np.float64)
    swap = rng.random(n) < 0.5
    pl, pr = np.where(swap, phi, plo), np.where(swap, plo, phi)
    rl, rr = rng.uniform(0.5, 2.0, n), rng.uniform(0.5, 2.0, n)
    u = rng.uniform(-0.3, 0.3, n)
    vl, vr = rng.uniform(-0.2, 0.2, n), rng.uniform(-0.2, 0.2, n)
    return np.stack([rl, u, vl, pl], axis=1), np.stack([rr, u, vr, pr], axis=1)


@pytest.mark.parametrize("plo_range", [(0.5, 2.0), (2.0 ** -970, 2.0 ** -960), (2.0 ** -975, 2.0 ** -965),
                                       (2.0 ** 1010, 2.0 ** 1019), (2.0 ** 1015, 2.0 ** 1019.5)])
def test_probe_hllc_pressure_ratio_two(gpulib, plo_range):
    """The 2D kernels' HLLC guess takes p_max / p_min <= 2 as
    p_max - 2 p_min <= 2^-52 p_min (every pair of a wavefront with
    2^-970 <= p_min and p_max <= 2^1020; otherwise it divides): pressure
    ratios within four ulps of 2, in the middle of that range, near its
    bounds and past them, give flux_hllc's and the oracle's bits."""
    from hydrofv_amd import gpu
    import _oracle as O
    wl, wr = _ratio_two_states(3, 4096, plo_range)
    fd, ed = gpu.probe_flux(gpulib, 4, 2, 1.4, 1e-6, wl, wr)
    fo, eo = O.flux_batch(2, "riemann_solver_hllc", 1.4, 1e-6, wl, wr)
    ed_code = np.select([ed == 0, ed == 2], [0, 2], default=-1)
    assert np.array_equal(ed_code, eo), [(i, ed[i], eo[i]) for i in np.flatnonzero(ed_code != eo)[:5]]
    eq = np.all((fd.view(np.uint64) == fo.view(np.uint64)) | (np.isnan(fd) & np.isnan(fo)), axis=1)
    bad = np.flatnonzero(~eq)
    assert bad.size == 0, [(wl[i].tolist(), wr[i].tolist(), fd[i].tolist(), fo[i].tolist()) for i in bad[:3]]
    plo, phi = np.minimum(wl[:, 3], wr[:, 3]), np.maximum(wl[:, 3], wr[:, 3])
    within = phi / plo <= 2.0
    assert 500 < within.sum() < 3600, within.sum()


def compare(tag, Wg, Ug, Wo, Uo, bitwise=False, tol=TOL, mask=None):
    """Every field within `tol` (relative to the field's largest magnitude);
    bitwise when requested.  Returns whether all fields are bitwise equal."""
    same = True
    for name, A, B in (("W", Wg, Wo), ("U", Ug, Uo)):
        for k in range(A.shape[0]):
            a, b = A[k], B[k]
            if mask is not None:
                a, b = a[mask], b[mask]
            same_nan = np.array_equal(np.isnan(a), np.isnan(b))
            assert same_nan, f"{tag}: NaN pattern differs in {name}[{k}]"
            err = K.rel_err(a, b)
            assert err <= tol, f"{tag}: {name}[{k}] relative error {err:.3e} > {tol}"
            eq = np.array_equal(a, b, equal_nan=True)
            same = same and eq
            if bitwise:
                assert eq, f"{tag}: {name}[{k}] not bitwise equal (max rel {err:.3e})"
    print(f"{tag}: within {tol:g}; bitwise={same}")
    return same


def interior_and_bc_mask(sim):
    """Cells whose content is specified by the reference's post-step contract:
    everything except the four ghost corner blocks."""
    g = sim.g
    m = np.ones((sim.ty, sim.tx), dtype=bool)
    m[:g, :g] = m[:g, -g:] = m[-g:, :g] = m[-g:, -g:] = False
    return m


def run_pair(gpulib, setup, limiter, solver="riemann_solver_hllc", steps=11, cfl=0.8, tol=1e-6,
             fixed_dt=None, exact_time=False):
    sim = setup(gpulib)
    grid = K.oracle_grid(sim, limiter, solver, tol=tol)
    res = sim.launch("muscl_hancock_2d", solver, limiter, cfl, tf=np.inf, tol=tol, max_steps=steps,
                     fixed_dt=fixed_dt)
    r = grid.run(cfl, max_steps=steps, fixed_dt=fixed_dt or 0.0)
    assert r.error == 0
    assert res.num_steps == r.num_steps == steps
    if exact_time:
        assert res.t == r.t and res.dt == r.dt
    else:
        assert abs(res.t - r.t) <= TOL * r.t and abs(res.dt - r.dt) <= TOL * r.dt
    W, U = K.state(sim)
    return sim, W, U, grid.W, grid.U


@pytest.mark.parametrize("shape", K.SHAPES, ids=lambda s: f"{s[0]}x{s[1]}g{s[2]}")
@pytest.mark.parametrize("bcs", list(K.BC_SETS), ids=str)
@pytest.mark.parametrize("limiter", K.LIMITERS)
def test_tiny_grids_bitwise(gpulib, shape, bcs, limiter):
    nx, ny, g = shape
    seed = 100 + nx * 7 + ny + g
    setup = lambda lib: K.setup_random_2d(lib, nx, ny, g, K.BC_SETS[bcs], seed)
    sim, W, U, Wo, Uo = run_pair(gpulib, setup, limiter, steps=11)
    compare(f"{shape} {bcs} {limiter}", W, U, Wo, Uo)
    sim.close()


@pytest.mark.parametrize("steps", [1, 2])
def test_first_steps_bitwise(gpulib, steps):
    setup = lambda lib: K.setup_random_2d(lib, 13, 7, 2, K.BC_SETS["mixed"], 7)
    sim, W, U, Wo, Uo = run_pair(gpulib, setup, "monotonized_central", steps=steps)
    compare(f"steps={steps}", W, U, Wo, Uo)


def test_kh_64_50_steps_bitwise(gpulib):
    sim, W, U, Wo, Uo = run_pair(gpulib, lambda lib: K.setup_kh(lib, 64), "monotonized_central", steps=50)
    compare("KH 64^2 x 50", W, U, Wo, Uo)


def test_kh_wide_strips_bitwise(gpulib):
    # 130 columns: three 60-column strips, the last one partly filled; several row strips
    sim, W, U, Wo, Uo = run_pair(gpulib, lambda lib: K.setup_kh(lib, 130), "minmod", steps=12)
    compare("KH 130^2 x 12", W, U, Wo, Uo)


@pytest.mark.parametrize("solver", [2, 1], ids=["hllc_kh130", "exact_sedov96"])
def test_strip_heights_identical(gpulib, solver):
    """Strips of 1, 2, 3, 5 and 37 rows (shorter than the three-row window of
    conserved rows a strip keeps in LDS for its updates, and taller than the
    default) give the bits of the library's default height: HLLC on a shear
    layer, the exact solver's queued faces on a blast; the default run is
    held against the oracle by the other tests."""
    from hydrofv_amd import problems
    if solver == 2:
        sim, lim, cfl, tol = K.setup_kh(gpulib, 130), 3, 0.8, 1e-6
    else:
        sim, lim, cfl, tol = problems.sedov(gpulib, 96), 2, 0.5, 1e-8
    W0, U0 = sim.primitive(), sim.conserved()
    runs = {}
    for rows in (0, 1, 2, 3, 5, 37):
        cfg = problems.device_config(sim, limiter=lim, solver=solver, cfl=cfl, max_steps=12, tol=tol, strip_rows=rows)
        m = gpu.MH2D(gpulib, cfg)
        m.upload(W0, U0)
        m.start()
        m.advance(11, use_graph=False)
        st = m.status()
        assert st.err_kind == 0 and st.num_steps == 12, (rows, st.err_kind, st.num_steps)
        runs[rows] = (st.t, np.stack(m.download()[1]))
        m.close()
    sim.close()
    t0, U0d = runs[0]
    for rows, (t, U) in runs.items():
        assert t == t0, rows
        assert np.array_equal(U.view(np.uint64), U0d.view(np.uint64)), rows


@pytest.mark.parametrize("case", ["kh130", "quadrants"])
def test_hllc_forms_identical(gpulib, case):
    """The two HLLC code forms (selects, branches) give the same bits on the
    single-grid engine, on a shear layer and on a four-quadrant problem."""
    from hydrofv_amd import problems, workloads
    if case == "kh130":
        sim, limiter = K.setup_kh(gpulib, 130), 3
    else:
        sim, limiter = problems.four_quadrant(gpulib, 150, workloads.ensemble_specs(1, seed=3)[1][0]), 1
    W0, U0 = sim.primitive(), sim.conserved()
    out = []
    for form in (gpu.HLLC_SELECTS, gpu.HLLC_BRANCHES):
        cfg = problems.device_config(sim, limiter=limiter, solver=2, cfl=0.8, max_steps=30, hllc_form=form)
        m = gpu.MH2D(gpulib, cfg)
        m.upload(W0, U0)
        m.start()
        m.advance(40, use_graph=True)
        st = m.status()
        assert st.err_kind == 0 and st.num_steps == 30
        out.append((st.t, np.stack(m.download()[1])))
        m.close()
    sim.close()
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1])


def far_from_disc(sim, r=0.3):
    """Cells (ghost cells included) farther than r from the centre of [0,1]^2."""
    X, Y = np.meshgrid(sim.mid_x, sim.mid_y)
    return (X - 0.5) ** 2 + (Y - 0.5) ** 2 > r * r


@pytest.mark.parametrize("bcs", ["periodic", "transmissive", "reflective", "mixed", "none"], ids=str)
@pytest.mark.parametrize("limiter", K.LIMITERS)
def test_uniform_regions_bitwise(gpulib, bcs, limiter):
    """A uniform flow (rho 1, u 0.4, v -0.3, p 1) around a denser disc, 190 x 150
    cells: the strip tiles away from the disc hold one state, so each computes
    one row and gives its results to its other rows (ghost columns and rows
    included); the disc's tiles take the full path.  Within the tolerance
    everywhere (the disc's two-rarefaction faces take pow), bitwise away from
    the disc, where every face has equal states."""
    setup = lambda lib: K.setup_blob_2d(lib, 190, 150, 2, K.BC_SETS[bcs])
    sim, W, U, Wo, Uo = run_pair(gpulib, setup, limiter, steps=12)
    compare(f"disc {bcs} {limiter}", W, U, Wo, Uo)
    compare(f"disc {bcs} {limiter}, far field", W, U, Wo, Uo, bitwise=True, mask=far_from_disc(sim))
    sim.close()


@pytest.mark.parametrize("case", ["sedov96_exact", "quadrants150_hllc_selects", "quadrants150_hllc_branches",
                                  "disc_gravity_first_order"])
def test_uniform_tiles_identical(gpulib, case):
    """The uniform-tile form (the default) against every row computed
    (full_tiles = 1), 40 steps: the same bits and the same time, on a blast
    in a quiescent medium (the exact solver's split passes and queued
    faces), a four-quadrant problem (both HLLC forms) and a disc in a
    uniform flow with gravity (first-order Godunov)."""
    from hydrofv_amd import problems, workloads
    kw = {}
    if case.startswith("sedov"):
        sim, kw = problems.sedov(gpulib, 96), dict(limiter=2, solver=1, cfl=0.5, tol=1e-8)
    elif case.startswith("quadrants"):
        sim = problems.four_quadrant(gpulib, 150, workloads.ensemble_specs(1, seed=3)[1][0])
        kw = dict(limiter=1, solver=2, cfl=0.8,
                  hllc_form=gpu.HLLC_BRANCHES if case.endswith("branches") else gpu.HLLC_SELECTS)
    else:
        sim = K.setup_blob_2d(gpulib, 130, 70, 1, K.BC_SETS["reflective"], gravity=0.7)
        kw = dict(limiter=1, solver=2, cfl=0.8)
    W0, U0 = sim.primitive(), sim.conserved()
    out = []
    for full in (0, 1):
        cfg = problems.device_config(sim, max_steps=40, full_tiles=full, **kw)
        if case.startswith("disc"):
            cfg.first_order = 1
        m = gpu.MH2D(gpulib, cfg)
        m.upload(W0, U0)
        m.start()
        m.advance(39, use_graph=False)
        st = m.status()
        assert st.err_kind == 0 and st.num_steps == 40, (full, st.err_kind, st.num_steps)
        W, U = m.download()
        out.append((st.t, np.stack(W), np.stack(U)))
        m.close()
    sim.close()
    assert out[0][0] == out[1][0]
    for k in (1, 2):
        assert np.array_equal(out[0][k].view(np.uint64), out[1][k].view(np.uint64)), k


@pytest.mark.parametrize("integrator", ["muscl_hancock_2d", "godunov_first_order_2d"])
def test_uniform_regions_gravity_and_first_order(gpulib, integrator):
    """The same with gravity 0.7 and reflective walls (the source term on the
    replicated rows), for both 2D integrators."""
    sim = K.setup_blob_2d(gpulib, 130, 70, 2, K.BC_SETS["reflective"], gravity=0.7)
    grid = K.oracle_grid(sim, "van_leer")
    res = sim.launch(integrator, "riemann_solver_hllc", "van_leer", 0.8, tf=np.inf, max_steps=11)
    r = grid.run(0.8, max_steps=11, first_order=integrator.startswith("godunov"))
    assert r.error == 0 and res.num_steps == r.num_steps == 11
    assert abs(res.t - r.t) <= TOL * r.t
    W, U = K.state(sim)
    compare(f"disc gravity {integrator}", W, U, grid.W, grid.U)
    compare(f"disc gravity {integrator}, far field", W, U, grid.W, grid.U, bitwise=True, mask=far_from_disc(sim))
    sim.close()


def test_gravity_bitwise(gpulib):
    setup = lambda lib: K.setup_random_2d(lib, 13, 9, 2, K.BC_SETS["reflective"], 11, gravity=0.7)
    sim, W, U, Wo, Uo = run_pair(gpulib, setup, "van_leer", steps=11)
    compare("gravity", W, U, Wo, Uo)


def test_fixed_dt_bitwise(gpulib):
    setup = lambda lib: K.setup_random_2d(lib, 13, 7, 2, K.BC_SETS["periodic"], 5)
    sim, W, U, Wo, Uo = run_pair(gpulib, setup, "minmod", steps=9, fixed_dt=1e-3)
    compare("fixed dt", W, U, Wo, Uo)


def test_run_to_tf_matches_reference_loop(gpulib):
    """launch_simulation to tf: step count, final t/dt and state."""
    sim = K.setup_random_2d(gpulib, 13, 7, 2, K.BC_SETS["transmissive"], 21)
    grid = K.oracle_grid(sim, "monotonized_central")
    tf = 0.05
    res = sim.launch("muscl_hancock_2d", "riemann_solver_hllc", "monotonized_central", 0.8, tf)
    r = grid.run(0.8, tf=tf)
    assert r.error == 0
    assert (res.num_steps, res.t, res.dt) == (r.num_steps, r.t, r.dt)
    W, U = K.state(sim)
    compare("to tf", W, U, grid.W, grid.U)


def test_exact_solver_sedov_small(gpulib):
    """Exact Riemann solver (2D extension): pow() may differ from glibc in the
    last place, so this case is held to the tolerance, not to bitwise."""
    sim, W, U, Wo, Uo = run_pair(gpulib, lambda lib: K.setup_sedov(lib, 32), "van_leer",
                                 solver="riemann_solver_exact", steps=20, cfl=0.5, tol=1e-8, exact_time=False)
    compare("sedov 32", W, U, Wo, Uo, bitwise=False, mask=interior_and_bc_mask(sim))


def test_dt_zero_error(gpulib):
    # an infinite pressure makes the CFL maximum infinite, hence dt = 0
    sim = driver.Simulation(gpulib, "cartesian_2d", 1.4, 8, 8, 2)
    sim.set_boundary("periodic", "periodic", "periodic", "periodic")
    W = K.random_state_2d(8, 8, 4)
    W[3][2, 5] = np.inf
    sim.set_primitive(W, interior_only=True)
    sim.apply_boundary()
    sim.primitive_to_conserved()
    with pytest.raises(driver.HydroError) as e:
        sim.launch("muscl_hancock_2d", "riemann_solver_hllc", "minmod", 0.5, tf=1.0)
    assert e.value.code == 2
    assert "dt is zero." in e.value.text


def test_hllc_invalid_wave_speed_error(gpulib):
    # NaN pressure in one interior cell makes every HLLC wave-speed test fail
    sim = driver.Simulation(gpulib, "cartesian_2d", 1.4, 8, 8, 2)
    sim.set_boundary("periodic", "periodic", "periodic", "periodic")
    W = K.random_state_2d(8, 8, 9)
    W[3][4, 4] = np.nan
    sim.set_primitive(W, interior_only=True)
    sim.apply_boundary()
    sim.primitive_to_conserved()
    W0, U0 = K.state(sim)
    with pytest.raises(driver.HydroError) as e:
        sim.launch("muscl_hancock_2d", "riemann_solver_hllc", "minmod", 0.5, tf=1.0)
    assert e.value.code == 2
    assert "Invalid wave speed for the HLLC riemann solver" in e.value.text
    # the reference stops inside the failing first step with W as given; the
    # device returns the state at the start of that step
    W1, U1 = K.state(sim)
    assert np.array_equal(W1.view(np.uint64), W0.view(np.uint64))
    assert np.array_equal(U1.view(np.uint64), U0.view(np.uint64))


def test_flux_failure_returns_the_state_before_the_failing_step(gpulib):
    """A failure in a later step: a run of 4 steps continued from a 7-step
    state whose conserved energy is poisoned in one cell fails in its
    second step (the first reads the caller's clean primitive state); the
    state returned is the one after its first step, which a 1-step run from
    the same poisoned state gives, and the status counts that one step."""
    setup = lambda lib: K.setup_random_2d(lib, 13, 9, 2, K.BC_SETS["periodic"], 17)
    runs = []
    for steps in (1, 4):
        sim = setup(gpulib)
        sim.launch("muscl_hancock_2d", "riemann_solver_hllc", "minmod", 0.5, tf=np.inf, max_steps=7)
        sim.system.energy_[5 * 17 + 6] = np.nan  # row 5, column 6: its pressure becomes NaN
        if steps == 1:
            res = sim.launch("muscl_hancock_2d", "riemann_solver_hllc", "minmod", 0.5, tf=np.inf, max_steps=1)
            assert res.num_steps == 1
        else:
            with pytest.raises(driver.HydroError) as e:
                sim.launch("muscl_hancock_2d", "riemann_solver_hllc", "minmod", 0.5, tf=np.inf, max_steps=4)
            assert e.value.code == 2
            assert sim.last_status.num_steps == 1
        runs.append(K.state(sim))
        sim.close()
    (Wa, Ua), (Wb, Ub) = runs
    assert np.array_equal(Ua.view(np.uint64), Ub.view(np.uint64))
    assert np.array_equal(Wa.view(np.uint64), Wb.view(np.uint64))


def test_probe_sqrt_grouped_root_is_the_library_root(gpulib):
    """The kernels' square roots (sqrt_rn: the library's refinement without its
    range scaling when every argument of a wavefront is in [2^-767, +inf))
    give the bits of the device library sqrt and of the correctly rounded
    root, for wavefronts entirely in range and for wavefronts holding an
    argument outside it (0, denormals, values just below 2^-767, +inf, NaN,
    negatives), which take the library sequence."""
    r = np.random.default_rng(11)
    n = 64 * 4096
    # bit patterns uniform over [2^-767, max finite]
    lo = np.float64(2.0 ** -767).view(np.int64)
    hi = np.float64(np.finfo(np.float64).max).view(np.int64)
    inr = r.integers(lo, hi, size=n, endpoint=True).view(np.float64)
    inr[:8] = [2.0 ** -767, np.nextafter(2.0 ** -767, 1.0), 1.0, 2.0, 0.5, np.finfo(np.float64).max, 4.0, 1e-5]
    # physical ranges the kernels see: gamma*p/rho and 1 + kq*(ps/p - 1)
    phys = np.exp(r.uniform(np.log(1e-12), np.log(1e12), n))
    special = np.array([0.0, -0.0, 5e-324, 2.0 ** -1022, np.nextafter(2.0 ** -767, 0.0), 2.0 ** -800, np.inf,
                        np.nan, -1.0, -np.inf, 1e-310])
    # wavefronts that mix one out-of-range argument into in-range ones
    mixed = r.uniform(0.1, 10.0, 64 * len(special))
    mixed[::64] = special
    for x in (inr, phys, mixed, special):
        lib = gpu.probe_sqrt(gpulib, 0, x)
        grp = gpu.probe_sqrt(gpulib, 1, x)
        with np.errstate(invalid="ignore"):
            ref = np.sqrt(x)
        np.testing.assert_array_equal(grp.view(np.int64), lib.view(np.int64))
        np.testing.assert_array_equal(lib, ref)  # NaN positions compare equal
    for x in (inr, phys):
        core = gpu.probe_sqrt(gpulib, 2, x)
        np.testing.assert_array_equal(core.view(np.int64), np.sqrt(x).view(np.int64))


def test_probe_division_core_is_the_library_division(gpulib):
    """The kernels' divisions (rcp_core / div_core: the device division's
    refinement without its scalings and fix-up when every operand of a
    wavefront has magnitude in [2^-300, 2^300]) give the bits of the
    device division and of the correctly rounded quotient, for wavefronts
    entirely in range and for wavefronts holding an operand outside it
    (0, -0, denormals, the range's neighbours, huge and tiny values,
    infinities, NaN), which take the division itself."""
    r = np.random.default_rng(13)
    n = 64 * 4096
    def in_range(k):
        lo = np.float64(2.0 ** -300).view(np.int64)
        hi = np.float64(2.0 ** 300).view(np.int64)
        x = r.integers(lo, hi, size=k, endpoint=True).view(np.float64)
        return x * r.choice([-1.0, 1.0], k)
    a, b = in_range(n), in_range(n)
    a[:6] = [2.0 ** -300, 2.0 ** 300, -2.0 ** 300, 1.0, 3.0, 1e-5]
    b[:6] = [2.0 ** 300, 2.0 ** -300, 2.0 ** -300, 3.0, 1.0, 7.0]
    # physical ranges: densities, pressures, momenta, dt A / V
    pa = np.exp(r.uniform(np.log(1e-12), np.log(1e12), n)) * r.choice([-1.0, 1.0], n)
    pb = np.exp(r.uniform(np.log(1e-8), np.log(1e8), n))
    special = np.array([0.0, -0.0, 5e-324, 2.0 ** -1022, np.nextafter(2.0 ** -300, 0.0), 2.0 ** 301, 2.0 ** 1000,
                        np.inf, -np.inf, np.nan, 1e-310, -2.0 ** -400])
    ma, mb = r.uniform(0.1, 10.0, 64 * len(special)), r.uniform(0.1, 10.0, 64 * len(special))
    ma[::64] = special
    mb[32::64] = special
    for x, y in ((a, b), (pa, pb), (ma, mb), (special, special[::-1].copy())):
        lib = gpu.probe_div(gpulib, 0, x, y)
        grp = gpu.probe_div(gpulib, 1, x, y)
        with np.errstate(all="ignore"):
            ref = x / y
        np.testing.assert_array_equal(grp.view(np.int64), lib.view(np.int64))
        np.testing.assert_array_equal(lib, ref)  # NaN positions compare equal
    for x, y in ((a, b), (pa, pb)):
        core = gpu.probe_div(gpulib, 2, x, y)
        np.testing.assert_array_equal(core.view(np.int64), (x / y).view(np.int64))


def test_probe_pow_pos_matches_host_build(gpulib):
    """The kernels' pow (pow_pos) gives the same bits on the
    device as its host build (tests/native), which test_adjoint_math checks
    against glibc and against correctly rounded values; the device library
    pow is reported beside it."""
    import ctypes as C
    import math
    import os
    r = np.random.default_rng(5)
    n = 200_000
    x = np.exp(r.uniform(np.log(1e-6), np.log(1e4), n))
    gam = r.uniform(1.05, 3.0, n)
    h = 0.5 * (gam - 1.0)
    y = np.where(r.uniform(size=n) < 0.7, h / gam, gam / h)
    keep = np.abs(y * np.log(x)) < 690
    x, y = x[keep], y[keep]
    dev = gpu.probe_pow(gpulib, 1, x, y)
    lib_pow = gpu.probe_pow(gpulib, 0, x, y)
    ref = np.array([math.pow(a, b) for a, b in zip(x.tolist(), y.tolist())])
    so = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "libhfv_hostcheck.so")
    if os.path.exists(so):
        L = C.CDLL(so)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        L.hc_pow_pos.argtypes = [C.c_longlong, dp, dp, dp]
        L.hc_pow_pos.restype = None
        host = np.empty_like(x)
        L.hc_pow_pos(len(x), x, y, host)
        assert np.array_equal(dev, host)
    assert np.all(np.abs(dev - ref) <= np.spacing(np.abs(ref)))
    # zero, negative, infinite, NaN and extreme arguments
    sx, sy = K.pow_special_grid()
    sdev = gpu.probe_pow(gpulib, 1, sx, sy)
    K.assert_pow_special(sdev, sx, sy)
    if os.path.exists(so):
        shost = np.empty_like(sx)
        L.hc_pow_pos(len(sx), sx, sy, shost)
        np.testing.assert_array_equal(sdev, shost)
    print(f"pow vs glibc over {len(x)} pressure-guess arguments: pow_pos differs in {np.mean(dev != ref):.2e}, "
          f"device library pow in {np.mean(lib_pow != ref):.2e}")


def test_step_limit_far_above_the_run_ends_at_tf(gpulib):
    """A step limit used as a safety cap (1e9) with a finite tf: the run
    stops at tf after its own step count, as the run to tf does."""
    import time
    setup = lambda lib: K.setup_random_2d(lib, 13, 7, 2, K.BC_SETS["transmissive"], 21)
    a, b = setup(gpulib), setup(gpulib)
    ra = a.launch("muscl_hancock_2d", "riemann_solver_hllc", "van_leer", 0.8, 0.05)
    t0 = time.perf_counter()
    rb = b.launch("muscl_hancock_2d", "riemann_solver_hllc", "van_leer", 0.8, 0.05, max_steps=10 ** 9)
    assert time.perf_counter() - t0 < 30.0
    assert (ra.num_steps, ra.t, ra.dt) == (rb.num_steps, rb.t, rb.dt)
    for x, y in zip(K.state(a), K.state(b)):
        assert np.array_equal(x.view(np.uint64), y.view(np.uint64))
