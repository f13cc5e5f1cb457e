"""The C ABI: libhydrofv_amd.so loads without a GPU and exports every function
that include/hydrofv/*.h declares; the struct layouts match the reference's
(checked by driving both libraries through the same ctypes structures in the
other tests).  No device work is started here."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", "hydrofv", h) for h in ("hydro.h", "hfv_gpu.h")]


def declared_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"#define[^\n]*(\\\n[^\n]*)*", "", text)
    names = set()
    for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{}]*)\)\s*;", text):
        name = m.group(1)
        before = text[:m.start()].rstrip()
        if name in ("sizeof",) or before.endswith("typedef") or "(*" in m.group(0):
            continue
        names.add(name)
    return names


def test_headers_declare_the_hot_path():
    names = set().union(*(declared_functions(h) for h in HEADERS))
    for required in ("muscl_hancock_2d", "muscl_hancock_1d", "godunov_first_order_2d", "godunov_first_order_1d",
                     "random_choice_1d",
                     "launch_simulation", "system_init",
                     "hfv_mh2d_create", "hfv_mh2d_advance", "hfv_mh1d_run", "hfv_probe_flux"):
        assert required in names


def test_library_exports_every_declared_symbol():
    import hydrofv_amd
    lib = C.CDLL(hydrofv_amd.LIB_PATH)
    missing = []
    for h in HEADERS:
        for name in sorted(declared_functions(h)):
            if not hasattr(lib, name):
                missing.append((os.path.basename(h), name))
    assert not missing, missing


def test_struct_sizes_match_x86_64_layout():
    from hydrofv_amd import abi
    assert C.sizeof(abi.ErrorStatus) == 24
    assert C.sizeof(abi.System) == 8 + 8 * 8 + 4 * 5 + 4 + 3 * 8 + 17 * 8
    assert C.sizeof(abi.BoundaryConditionParam) == 6 * 8 + 6 * 8 + 6 * 4
    assert C.sizeof(abi.IntegratorParam) == 3 * 8 + 8 + 8 + 8 + 8 + 3 * 4 + 4
    assert C.sizeof(abi.SimulationStatus) == 24


def test_reference_and_library_agree_on_host_setup(reflib):
    """system_init geometry, boundary fill and primitive->conserved conversion
    of the library equal the reference's bit for bit."""
    import numpy as np
    import _cases as K
    import hydrofv_amd
    lib = hydrofv_amd.load()
    makers = [lambda L: K.setup_kh(L, 40), lambda L: K.setup_random_2d(L, 13, 7, 2, K.BC_SETS["mixed"], 5),
              lambda L: K.setup_random_2d(L, 9, 11, 3, K.BC_SETS["none"], 5),
              lambda L: K.setup_sod(L, 100, coord="spherical_1d"), lambda L: K.setup_sod(L, 100, coord="cylindrical_1d"),
              lambda L: K.setup_sedov(L, 16)]
    for mk in makers:
        a, b = mk(lib), mk(reflib)
        Wa, Ua = K.state(a)
        Wb, Ub = K.state(b)
        assert np.array_equal(Wa, Wb, equal_nan=True) and np.array_equal(Ua, Ub, equal_nan=True)
        assert a.system.dx_ == b.system.dx_ and a.system.dy_ == b.system.dy_
        for name in ("mid_points_x_", "surface_area_x_", "volume_"):
            n = a.tx if name != "volume_" else a.tx * a.ty
            va = np.ctypeslib.as_array(getattr(a.system, name), (n,))
            vb = np.ctypeslib.as_array(getattr(b.system, name), (n,))
            lo = 1 if name == "surface_area_x_" else 0  # entry 0 is unset by the reference in cylindrical 1D
            assert np.array_equal(va[lo:], vb[lo:]), name


@pytest.mark.parametrize("field,value,code,text", [
    ("gamma", 0.5, 2, "Gamma must be greater than 1"),
    ("num_ghost_cells_side", 0, 2, "Number of ghost cells per side must be at least 1"),
])
def test_system_init_errors(field, value, code, text):
    import hydrofv_amd
    from hydrofv_amd import abi, driver
    lib = hydrofv_amd.load()
    s = lib.get_new_system_struct()
    s.coord_sys = b"cartesian_2d"
    s.gamma, s.x_min, s.x_max, s.y_min, s.y_max = 1.4, 0.0, 1.0, 0.0, 1.0
    s.num_cells_x = s.num_cells_y = 8
    s.num_ghost_cells_side = 2
    setattr(s, field, value)
    st = lib.system_init(C.byref(s))
    assert st.return_code == code
    assert text in abi.traceback_text(lib, st)


def test_launch_errors_before_any_device_work():
    """Parameter validation happens on the host with the reference's messages."""
    import numpy as np
    import hydrofv_amd
    import _cases as K
    from hydrofv_amd import driver
    lib = hydrofv_amd.load()
    sim = K.setup_random_2d(lib, 8, 8, 2, K.BC_SETS["periodic"], 1)
    for kwargs, code, text in [
        (dict(integrator="muscl_hancock_2d", riemann_solver="riemann_solver_hllc", slope_limiter="superbee"), 2,
         "Slope limiter not recognized."),
        (dict(integrator="muscl_hancock_3d", riemann_solver="riemann_solver_hllc", slope_limiter="minmod"), 2,
         "Integrator not recognized."),
        (dict(integrator="muscl_hancock_1d", riemann_solver="riemann_solver_hllc", slope_limiter="minmod"), 2,
         "Wrong coordinate system"),
        (dict(integrator="random_choice_1d", riemann_solver="riemann_solver_exact", slope_limiter="minmod"), 2,
         "Wrong coordinate system"),
    ]:
        with pytest.raises(driver.HydroError) as e:
            sim.launch(cfl=0.5, tf=1.0, **kwargs)
        assert e.value.code == code and text in e.value.text
    sim1 = K.setup_inflow_1d(lib, 16)
    with pytest.raises(driver.HydroError) as e:
        sim1.launch("random_choice_1d", "riemann_solver_exact", "minmod", 0.5, 0.1)
    assert e.value.code == 6 and "random-choice" in e.value.text
    small_g = driver.Simulation(lib, "cartesian_2d", 1.4, 8, 8, 1)
    small_g.set_boundary("periodic", "periodic", "periodic", "periodic")
    with pytest.raises(driver.HydroError) as e:
        small_g.launch("muscl_hancock_2d", "riemann_solver_hllc", "minmod", 0.5, 1.0)
    assert "Number of ghost cells must be at least 2." in e.value.text
    with pytest.raises(driver.HydroError) as e:
        sim.set_boundary("periodic", "reflective", "periodic", "periodic")
    assert "Periodic boundary condition mismatch for x_min and x_max." in e.value.text


def test_ensemble_rejects_first_order_configs():
    """The batched engine runs MUSCL-Hancock only; a first-order config is
    refused before any device work."""
    import hydrofv_amd
    import _cases as K
    from hydrofv_amd import ensemble, problems
    lib = hydrofv_amd.load()
    sim = K.setup_random_2d(lib, 8, 8, 2, K.BC_SETS["periodic"], 1)
    cfg = problems.device_config(sim, limiter=1, solver=2, cfl=0.5)
    cfg.first_order = 1
    with pytest.raises(RuntimeError):
        ensemble.Ensemble(lib, [cfg])


def test_device_config_mirrors_match_the_header():
    """The ctypes mirrors of hfv_gpu.h's structs (hydrofv_amd.gpu,
    hydrofv_amd.adjoint) have the C layout, as compiled by the host build in
    tests/native."""
    import ctypes as C
    import os
    from hydrofv_amd import adjoint, gpu
    so = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "libhfv_hostcheck.so")
    if not os.path.exists(so):
        pytest.fail(f"{so} is not built (make -C tests/native)")
    L = C.CDLL(so)
    out = (C.c_longlong * 8)()
    L.hc_struct_layout(out)
    assert list(out[:4]) == [C.sizeof(gpu.MH2DConfig), C.sizeof(gpu.MH1DConfig), C.sizeof(gpu.RunStatus),
                             C.sizeof(adjoint.Adj2DConfig)]
    assert out[4] == gpu.MH2DConfig.full_tiles.offset
    assert out[5] == gpu.MH1DConfig.t_pause.offset
    assert out[6] == gpu.MH1DConfig.random_choice.offset
    assert out[7] == adjoint.Adj2DConfig.slope_limiter.offset


def test_strip_height_fills_the_rounds_of_resident_wavefronts():
    """The 2D step's default strip height (hfv_strips.hpp, host build in
    tests/native): the height near the 32-rows-halved rule (in [28, 40] for a
    large grid, within 25% of the halved height for a smaller one) whose strip
    count best fills the rounds of resident wavefronts; an unknown device
    keeps the rule."""
    import ctypes as C
    import math
    import os
    so = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "libhfv_hostcheck.so")
    if not os.path.exists(so):
        pytest.fail(f"{so} is not built (make -C tests/native)")
    L = C.CDLL(so)
    L.hc_strip_rows.argtypes = [C.c_int, C.c_int, C.c_longlong]
    L.hc_strip_rows.restype = C.c_int

    def strips(n, r):
        return math.ceil(n / 60) * math.ceil(n / r)

    def fill(t, slots):
        return t / (math.ceil(t / slots) * slots)

    # MI355X: 256 CUs x 8 resident one-wavefront workgroups
    assert L.hc_strip_rows(4096, 4096, 2048) == 35          # 8142 strips: 99.4% of 4 rounds
    assert L.hc_strip_rows(2048, 2048, 2048) == 18          # 3990 strips, 97% of 2 rounds (16 rows: 73% of 3)
    assert L.hc_strip_rows(4096, 4096, 0) == 32             # device unknown
    assert L.hc_strip_rows(256, 256, 2048) == 6             # 215 strips: more, shorter strips
    assert L.hc_strip_rows(1024, 1024, 2048) == 10          # 1854 strips in one round
    assert L.hc_strip_rows(2048, 2048, 0) == 16
    for n in (3000, 4096, 5000, 8192):
        r = L.hc_strip_rows(n, n, 2048)
        assert 28 <= r <= 40
        assert fill(strips(n, r), 2048) >= max(fill(strips(n, q), 2048) for q in range(28, 41)) - 1e-3
    for n in (600, 1024, 1500, 2048):
        base = L.hc_strip_rows(n, n, 0)
        r = L.hc_strip_rows(n, n, 2048)
        window = range(base - base // 4, base + base // 4 + 1)
        assert r in window
        assert fill(strips(n, r), 2048) >= max(fill(strips(n, q), 2048) for q in window) - 1e-3
    # the exact solver's default: [28, 40] rows once a grid has half a round of 32-row strips
    L.hc_exact_strip_rows.argtypes = [C.c_int, C.c_int, C.c_longlong]
    L.hc_exact_strip_rows.restype = C.c_int
    assert L.hc_exact_strip_rows(2048, 2048, 2048) == 36    # 1995 strips: 97% of one round
    assert L.hc_exact_strip_rows(256, 256, 2048) == L.hc_strip_rows(256, 256, 2048)  # small: the general rule
    assert L.hc_exact_strip_rows(2048, 2048, 0) == 16       # device unknown
    for n in (1500, 2048, 4096):
        r = L.hc_exact_strip_rows(n, n, 2048)
        assert 28 <= r <= 40
        assert fill(strips(n, r), 2048) >= max(fill(strips(n, q), 2048) for q in range(28, 41)) - 1e-3


def test_device_bc_registry():
    """hfv_register_device_bc_2d: host-side table of device hook forms (no
    device work): NULL host hook refused, re-registration replaces, NULL
    device hook removes, at most 64 entries."""
    import hydrofv_amd
    lib = hydrofv_amd.load()
    reg = lib.hfv_register_device_bc_2d
    reg.restype = C.c_int
    reg.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    base = 0x1000
    assert reg(None, 0x10, None) == -1
    try:
        for k in range(64):
            assert reg(base + 8 * k, 0x10, None) == 0
        assert reg(base + 8 * 64, 0x10, None) == -1          # full
        assert reg(base, 0x20, 0x30) == 0                     # replaces, no new entry
        assert reg(base + 8, None, None) == 0                 # removes one
        assert reg(base + 8 * 64, 0x10, None) == 0            # room again
        assert reg(base + 8 * 65, 0x10, None) == -1
        assert reg(base + 8 * 999, None, None) == 0           # removing an absent entry is harmless
    finally:
        for k in range(66):
            reg(base + 8 * k, None, None)
    for k in range(64):
        assert reg(base + 8 * k, 0x10, None) == 0
    for k in range(64):
        reg(base + 8 * k, None, None)
