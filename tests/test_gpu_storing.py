"""Snapshots written during device runs (launch_simulation with storing on):
the files, their steps and times follow the reference loop's storing rule
(integrator_muscl_hancock.c:1146-1155: after every step that ends at
t >= interval * (stored - initial + 1), at most one per step), and each file
holds the state of that step, checked against the oracle (1e-12 relative,
mostly bitwise) and, where the reference build is present, against the
reference's own snapshot files (h5diff for the layout, then the fields to
1e-12 relative; 1e-10 with the exact solver)."""
import glob
import os
import re

import numpy as np
import pytest

import _cases as K
import _h5 as H
import _oracle as O

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not H.available(), reason="HDF5 library missing")]


def snapshot_files(d):
    files = glob.glob(os.path.join(d, "snapshot_*.h5"))
    return sorted(files, key=lambda f: int(re.search(r"snapshot_(\d+)\.h5$", f).group(1)))


def fields_of(path, dim):
    names = ["density", "velocity_x", "velocity_y", "pressure"] if dim == 2 else ["density", "velocity_x", "pressure"]
    return np.stack([H.read(path, f"/fields/{n}") for n in names])


def check_run(gpulib, tmp_path, mk, integrator, solver, limiter, cfl, tf, interval, store_initial=True,
              ref_compare=True):
    out = str(tmp_path / "ours") + "/"
    sim = mk(gpulib)
    dim = sim.dim
    first_order = integrator.startswith("godunov")
    random_choice = integrator.startswith("random_choice")
    res = sim.launch(integrator, solver, limiter, cfl, tf, store_dir=out, store_interval=interval,
                     store_initial=store_initial)
    files = snapshot_files(out)
    assert len(files) >= 3
    offset = 1 if store_initial else 0
    prev_n = 0
    for k, f in enumerate(files):
        n = int(H.read(f, "/simulation_status/num_steps", "i8")[0])
        t = float(H.read(f, "/simulation_status/simulation_time")[0])
        dt = float(H.read(f, "/simulation_status/dt")[0])
        start = mk(gpulib)
        grid = K.oracle_grid(start, limiter, solver)
        r = grid.run(cfl, tf=tf, max_steps=n, first_order=first_order, random_choice=random_choice)
        assert r.error == 0 and r.num_steps == n
        tol = 1e-10 if solver == "riemann_solver_exact" else 1e-12
        if solver == "riemann_solver_exact":  # pow() rounding may reach dt
            assert abs(t - r.t) <= tol * max(r.t, 1e-300) and abs(dt - r.dt) <= tol * max(r.dt, 1e-300)
        else:
            assert (t, dt) == (r.t, r.dt), (f, t, r.t, dt, r.dt)
        got = fields_of(f, dim)
        want = grid.W.reshape(got.shape)
        for q in range(got.shape[0]):
            assert K.rel_err(got[q], want[q]) <= tol, (f, q)
        if k < offset:
            assert n == 0
            continue
        # the storing rule: the first step after the previous snapshot that
        # ends at t >= interval * (k - offset + 1)
        due = interval * (k - offset + 1)
        assert t >= due
        if n - 1 > prev_n:
            g2 = K.oracle_grid(mk(gpulib), limiter, solver)
            r2 = g2.run(cfl, tf=tf, max_steps=n - 1, first_order=first_order, random_choice=random_choice)
            assert r2.t < due
        prev_n = n
    assert prev_n <= res.num_steps
    if ref_compare and O.ref_available() and H.h5diff():
        import ctypes as C
        from hydrofv_amd import abi
        ref = abi.bind(C.CDLL(O.REF_SO))
        ref_out = str(tmp_path / "ref") + "/"
        rsim = mk(ref)
        rres = rsim.launch(integrator, solver, limiter, cfl, tf, store_dir=ref_out, store_interval=interval,
                           store_initial=store_initial)
        assert rres.num_steps == res.num_steps
        ref_files = snapshot_files(ref_out)
        assert [os.path.basename(f) for f in ref_files] == [os.path.basename(f) for f in files]
        tol = 1e-10 if solver == "riemann_solver_exact" else 1e-12
        for a, b in zip(files, ref_files):
            same, text = H.same_files(a, b, delta=1e300)  # groups, names, types, shapes
            assert same, (a, text)
            assert H.read(a, "/simulation_status/num_steps", "i8")[0] == H.read(b, "/simulation_status/num_steps", "i8")[0]
            fa, fb = fields_of(a, dim), fields_of(b, dim)
            for q in range(fa.shape[0]):
                assert K.rel_err(fa[q], fb[q]) <= tol, (a, q)
    return files


def test_kh_2d_snapshots(gpulib, tmp_path):
    files = check_run(gpulib, tmp_path, lambda L: K.setup_kh(L, 64), "muscl_hancock_2d", "riemann_solver_hllc",
                      "monotonized_central", 0.8, 0.05, 0.01)
    assert len(files) == 6


def test_godunov_2d_snapshots_without_initial(gpulib, tmp_path):
    check_run(gpulib, tmp_path, lambda L: K.setup_random_2d(L, 40, 30, 1, K.BC_SETS["mixed"], 8),
              "godunov_first_order_2d", "riemann_solver_hllc", "minmod", 0.8, 0.06, 0.02, store_initial=False)


def test_sod_1d_snapshots(gpulib, tmp_path):
    files = check_run(gpulib, tmp_path, lambda L: K.setup_sod(L, 200), "muscl_hancock_1d", "riemann_solver_hllc",
                      "minmod", 0.9, 0.2, 0.05)
    assert len(files) == 5


def test_godunov_1d_spherical_snapshots(gpulib, tmp_path):
    check_run(gpulib, tmp_path,
              lambda L: K.setup_sod(L, 64, coord="spherical_1d", bc=("reflective", "transmissive")),
              "godunov_first_order_1d", "riemann_solver_exact", "minmod", 0.9, 0.2, 0.03)


def test_random_choice_snapshots(gpulib, tmp_path):
    check_run(gpulib, tmp_path, lambda L: K.setup_sod(L, 100), "random_choice_1d", "riemann_solver_exact",
              "minmod", 0.5, 0.2, 0.05)


def test_exact_2d_snapshots(gpulib, tmp_path):
    """The exact solver's queued-face step paused at every snapshot (the
    reference has no 2D exact solver, so the oracle is the only check)."""
    files = check_run(gpulib, tmp_path, lambda L: K.setup_sedov(L, 48), "muscl_hancock_2d", "riemann_solver_exact",
                      "van_leer", 0.5, 0.05, 0.01, ref_compare=False)
    assert len(files) == 6
