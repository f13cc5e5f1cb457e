"""Custom 2D boundary hooks in device form (hfv_register_device_bc_2d,
hfv_mh2d_set_device_bc).  The step's primitive state is filled on the device:
c2p of every cell, then the built-in sides and the hooks' own kernels in the
reference's side order (boundary.c:513-777).  Each run is compared with the
oracle running the host forms of the same hooks (pinned to the reference
build in the CPU suite), and the host forms are not called during the run."""
import ctypes as C

import numpy as np
import pytest

import _cases as K
import _devhooks
from hydrofv_amd import gpu
from test_gpu_parity_2d import compare

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not _devhooks.available(), reason="tests/native/libhfv_devhooks.so not built")]

CASES = [
    ("cavity", lambda L, **kw: K.setup_cavity(L, 24, **kw), "muscl_hancock_2d", "minmod", 30),
    ("cavity_periodic_x", lambda L, **kw: K.setup_cavity(L, 24, bcs=("periodic", "periodic", "reflective"), **kw),
     "muscl_hancock_2d", "monotonized_central", 30),
    ("cavity_transmissive_g3",
     lambda L, **kw: K.setup_cavity(L, 20, g=3, bcs=("transmissive", "reflective", "transmissive"), **kw),
     "muscl_hancock_2d", "van_leer", 25),
    ("cavity_130", lambda L, **kw: K.setup_cavity(L, 130, **kw), "muscl_hancock_2d", "monotonized_central", 12),
    ("all_custom", lambda L, **kw: K.setup_all_custom(L, 17, 11, 2, 5, **kw), "muscl_hancock_2d", "van_leer", 20),
    ("all_custom_godunov", lambda L, **kw: K.setup_all_custom(L, 17, 11, 1, 6, **kw), "godunov_first_order_2d",
     "minmod", 20),
]


def run_and_compare(gpulib, name, mk, integrator, limiter, steps, expect_host_calls=False):
    calls = []
    sim = mk(gpulib, device=True, calls=calls)
    grid = K.oracle_grid(sim, limiter, "riemann_solver_hllc")
    before = len(calls)
    res = sim.launch(integrator, "riemann_solver_hllc", limiter, 0.8, np.inf, max_steps=steps)
    host_calls = len(calls) - before
    r = grid.run(0.8, max_steps=steps, first_order=integrator.startswith("godunov"))
    assert r.error == 0 and (res.num_steps, res.t, res.dt) == (r.num_steps, r.t, r.dt)
    W, U = K.state(sim)
    compare(f"device hooks {name}", W, U, grid.W.reshape(W.shape), grid.U.reshape(U.shape))
    sim.close()
    return host_calls


@pytest.mark.parametrize("name,mk,integrator,limiter,steps", CASES, ids=[c[0] for c in CASES])
def test_device_hooks_match_oracle(gpulib, name, mk, integrator, limiter, steps):
    assert run_and_compare(gpulib, name, mk, integrator, limiter, steps) == 0


def test_device_hooks_with_the_exact_solver(gpulib):
    """The exact solver's queued-face step under device hooks: every step
    reads the primitive state the device fill has just refreshed."""
    calls = []
    sim = K.setup_all_custom(gpulib, 40, 23, 2, 9, device=True, calls=calls)
    grid = K.oracle_grid(sim, "van_leer", "riemann_solver_exact")
    before = len(calls)
    res = sim.launch("muscl_hancock_2d", "riemann_solver_exact", "van_leer", 0.5, np.inf, max_steps=12)
    assert len(calls) == before
    r = grid.run(0.5, max_steps=12)
    assert r.error == 0 and res.num_steps == r.num_steps == 12
    assert abs(res.t - r.t) <= 1e-10 * r.t
    W, U = K.state(sim)
    compare("device hooks exact", W, U, grid.W.reshape(W.shape), grid.U.reshape(U.shape), tol=1e-10)
    sim.close()


def test_partly_registered_hooks_keep_the_host_path(gpulib):
    """Two of four custom sides with a device form: the run takes the host
    path, one host call per custom side and step."""
    mk = lambda L, **kw: K.setup_all_custom(L, 17, 11, 2, 7, device_sides=(0, 1), **kw)
    calls = run_and_compare(gpulib, "partial", mk, "muscl_hancock_2d", "minmod", 9)
    assert calls == 4 * 9


def test_device_hooks_to_tf_with_snapshots(gpulib, tmp_path):
    """The run-to-tf loop with snapshots: every file equals the oracle's state
    at its step, and the lid is filled on the device."""
    import glob
    import _h5 as H
    if not H.available():
        pytest.skip("HDF5 library missing")
    calls = []
    sim = K.setup_cavity(gpulib, 32, device=True, calls=calls)
    grid = K.oracle_grid(sim, "monotonized_central", "riemann_solver_hllc")
    out = str(tmp_path) + "/"
    before = len(calls)
    res = sim.launch("muscl_hancock_2d", "riemann_solver_hllc", "monotonized_central", 0.8, 0.05,
                     store_dir=out, store_interval=0.02, store_initial=True)
    assert len(calls) == before
    r = grid.run(0.8, tf=0.05)
    assert r.error == 0 and (res.num_steps, res.t, res.dt) == (r.num_steps, r.t, r.dt)
    W, U = K.state(sim)
    compare("device hooks cavity to tf", W, U, grid.W, grid.U)
    files = sorted(glob.glob(out + "snapshot_*.h5"))
    assert len(files) == 3
    for f in files[1:]:
        n = int(H.read(f, "/simulation_status/num_steps", "i8")[0])
        g2 = K.oracle_grid(K.setup_cavity(gpulib, 32), "monotonized_central", "riemann_solver_hllc")
        assert g2.run(0.8, tf=0.05, max_steps=n).num_steps == n
        got = np.stack([H.read(f, f"/fields/{q}") for q in ("density", "velocity_x", "velocity_y", "pressure")])
        want = g2.W.reshape(got.shape)
        assert np.array_equal(got, want, equal_nan=True), f


def test_set_device_bc_checks(gpulib):
    """hfv_mh2d_set_device_bc accepts custom sides only, and a run with a
    custom side left without a device hook refuses to advance."""
    gpu.bind(gpulib)
    sim = K.setup_cavity(gpulib, 16)
    cfg = gpu.MH2DConfig()
    cfg.num_cells_x = cfg.num_cells_y = 16
    cfg.num_ghost_cells_side = 2
    cfg.gamma = 1.4
    cfg.dx = cfg.dy = sim.system.dx_
    cfg.area_x = cfg.area_y = sim.system.dx_
    cfg.volume = sim.system.dx_ * sim.system.dy_
    cfg.slope_limiter = 1
    cfg.riemann_solver = 2
    cfg.tol = 1e-6
    cfg.bc[:] = [5, 2, 5, 5]
    cfg.cfl = 0.8
    cfg.cfl_initial_shrink_factor = 0.2
    cfg.cfl_initial_shrink_num_steps = 10
    cfg.tf = 1.0
    cfg.max_steps = -1
    h = C.c_void_p()
    assert gpulib.hfv_mh2d_create(C.byref(cfg), C.byref(h)) == 0
    try:
        fn, user = _devhooks.lid()
        assert gpulib.hfv_mh2d_set_device_bc(h, 1, fn, user) == -1  # reflective side
        assert gpulib.hfv_mh2d_set_device_bc(h, 4, fn, user) == -1
        assert gpulib.hfv_mh2d_set_device_bc(h, 3, fn, user) == 0
        W = [np.ascontiguousarray(a) for a in sim.primitive()]
        U = [np.ascontiguousarray(a) for a in sim.conserved()]
        wp = (gpu.dp * 4)(*[a.ctypes.data_as(gpu.dp) for a in W])
        up = (gpu.dp * 4)(*[a.ctypes.data_as(gpu.dp) for a in U])
        assert gpulib.hfv_mh2d_upload(h, wp, up, None) == 0
        assert gpulib.hfv_mh2d_start(h, None) == 0
        assert gpulib.hfv_mh2d_advance(h, 2, 0, None) == -1  # x_min and y_min have none
        fn2, user2 = _devhooks.copy(2)
        assert gpulib.hfv_mh2d_set_device_bc(h, 0, fn2, user2) == 0
        assert gpulib.hfv_mh2d_set_device_bc(h, 2, fn2, user2) == 0
        assert gpulib.hfv_mh2d_advance(h, 2, 0, None) == 0
        st = gpu.RunStatus()
        assert gpulib.hfv_mh2d_status(h, C.byref(st), None) == 0
        assert st.err_kind == 0 and st.num_steps == 3
    finally:
        gpulib.hfv_mh2d_destroy(h)
        sim.close()
