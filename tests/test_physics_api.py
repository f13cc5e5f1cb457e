"""The reference's per-interface and per-cell API exported by libhydrofv_amd.so
(limit_slope, solve_flux_*, guess_p, riemann_A/B_L_or_R, solve_vacuum_*,
solve_exact_1d, get_sound_speed, get_time_step_*, the source terms), called
through its C ABI on the CPU:
  * against the committed fixtures that the reference build wrote
    (tests/golden/scalar_vectors.npz), bit for bit;
  * where the reference build (oracle/_ref) is present, against the same
    functions of that build on the same inputs, bit for bit, including the
    functions no fixture covers.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest

import _cases as K
import _oracle as O
from hydrofv_amd import abi, driver

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scalar_vectors.npz")
LIMITERS = {"minmod": 1, "van_leer": 2, "monotonized_central": 3}


@pytest.fixture(scope="module")
def lib():
    import hydrofv_amd
    return abi.bind(C.CDLL(hydrofv_amd.LIB_PATH))


@pytest.fixture(scope="module")
def ref():
    if not O.ref_available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    return abi.bind(C.CDLL(O.REF_SO))


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _free(L, st):
    if st.traceback:
        L.free_traceback(C.byref(st))


def limits(L, flag, a, b):
    ip = abi.IntegratorParam()
    ip.slope_limiter_flag_ = flag
    v = C.c_double()
    out = np.empty(len(a))
    for k in range(len(a)):
        st = L.limit_slope(C.byref(v), C.byref(ip), float(a[k]), float(b[k]))
        assert st.return_code == 0
        out[k] = v.value
    return out


def fluxes(L, kind, gamma, tol, wl, wr, verbose=0):
    """kind: hllc1d, hllc2d, exact1d (the named functions) or disp1d/disp2d
    (solve_flux_1d/2d with the flags of `tol`'s solver: HLLC)."""
    nf = 4 if kind.endswith("2d") else 3
    n = wl.shape[0]
    f = np.full((n, nf), np.nan)
    err = np.zeros(n, dtype=np.int32)
    out = [C.c_double() for _ in range(nf)]
    ptrs = [C.byref(o) for o in out]
    ip = abi.IntegratorParam()
    ip.riemann_solver_flag_ = 2
    ip.tol = tol
    se = abi.Settings(verbose=verbose)
    for k in range(n):
        l = [float(x) for x in wl[k]]
        r = [float(x) for x in wr[k]]
        for o in out:
            o.value = np.nan
        if kind == "hllc1d":
            st = L.solve_flux_hllc_1d(*ptrs, gamma, *l, *r, tol)
        elif kind == "hllc2d":
            st = L.solve_flux_hllc_2d(*ptrs, gamma, *l, *r, tol)
        elif kind == "exact1d":
            st = L.solve_flux_exact_1d(*ptrs, gamma, *l, *r, tol, 0.0, verbose)
        elif kind == "disp1d":
            st = L.solve_flux_1d(C.byref(ip), C.byref(se), *ptrs, gamma, *l, *r)
        else:
            st = L.solve_flux_2d(C.byref(ip), C.byref(se), *ptrs, gamma, *l, *r)
        err[k] = st.return_code
        _free(L, st)
        f[k] = [o.value for o in out]
    return f, err


def test_limit_slope_equals_the_reference_fixtures(lib):
    g = np.load(GOLD)
    for name, flag in LIMITERS.items():
        got = limits(lib, flag, g["lim_a"], g["lim_b"])
        assert np.array_equal(bits(got), bits(g[f"lim_{name}"])), name
    ip = abi.IntegratorParam()
    ip.slope_limiter_flag_ = 7
    v = C.c_double(1.5)
    st = lib.limit_slope(C.byref(v), C.byref(ip), 1.0, 2.0)
    assert st.return_code == abi.VALUE_ERROR
    assert "Limiter not recognized." in abi.traceback_text(lib, st)


@pytest.mark.parametrize("kind,key,gamma,tol", [
    ("hllc2d", "hllc2d", 1.4, 1e-6),
    ("disp2d", "hllc2d", 1.4, 1e-6),
    ("hllc1d", "hllc1d", 1.4, 1e-6),
    ("disp1d", "hllc1d", 1.4, 1e-6),
    ("exact1d", "exact1d", 1.4, 1e-6),
    ("exact1d", "exact1d_g53_tol8", 5.0 / 3.0, 1e-8),
])
def test_fluxes_equal_the_reference_fixtures(lib, kind, key, gamma, tol):
    g = np.load(GOLD)
    wl, wr = g["wl"], g["wr"]
    if not kind.endswith("2d"):
        wl, wr = wl[:, [0, 1, 3]], wr[:, [0, 1, 3]]
    f, err = fluxes(lib, kind, gamma, tol, wl, wr)
    assert np.array_equal(err, g[f"{key}_err"])
    ok = err == 0  # a failed exact solve leaves the reference's outputs unset
    assert np.array_equal(bits(f[ok]), bits(g[f"{key}_f"][ok])), key
    assert (~ok).sum() >= 1  # the NaN-pressure state fails in both


def test_flux_dispatch_errors(lib):
    ip = abi.IntegratorParam()
    ip.riemann_solver_flag_ = 9
    se = abi.Settings()
    out = [C.c_double() for _ in range(4)]
    st = lib.solve_flux_1d(C.byref(ip), C.byref(se), *[C.byref(o) for o in out[:3]], 1.4, 1, 0, 1, 0.125, 0, 0.1)
    assert st.return_code == abi.VALUE_ERROR
    assert "Riemann solver flag not recognized." in abi.traceback_text(lib, st)
    st = lib.solve_flux_hllc_1d(*[C.byref(o) for o in out[:3]], 1.4, 1.0, 0.0, float("nan"), 1.0, 0.0, 1.0, 1e-6)
    assert st.return_code == abi.VALUE_ERROR
    assert "Invalid wave speed for the HLLC riemann solver." in abi.traceback_text(lib, st)


def test_exact_solver_in_2d_extends_the_1d_flux(lib):
    """solve_flux_2d with the exact solver (an extension: the reference
    rejects it in 2D): normal components equal the 1D exact flux bit for bit
    at v = 0, the transverse momentum is the mass flux times the upwind v."""
    g = np.load(GOLD)
    wl, wr = g["wl"][:400].copy(), g["wr"][:400].copy()
    ip = abi.IntegratorParam()
    ip.riemann_solver_flag_ = 1
    ip.tol = 1e-6
    se = abi.Settings()
    out = [C.c_double() for _ in range(4)]
    f1, e1 = fluxes(lib, "exact1d", 1.4, 1e-6, wl[:, [0, 1, 3]], wr[:, [0, 1, 3]])
    for k in range(len(wl)):
        if e1[k] != 0:
            continue
        l0, r0 = wl[k].copy(), wr[k].copy()
        l0[2] = r0[2] = 0.0
        st = lib.solve_flux_2d(C.byref(ip), C.byref(se), *[C.byref(o) for o in out], 1.4, *map(float, l0),
                               *map(float, r0))
        assert st.return_code == 0
        v = [o.value for o in out]
        assert np.array_equal(bits([v[0], v[1], v[3]]), bits(f1[k])), k
        st = lib.solve_flux_2d(C.byref(ip), C.byref(se), *[C.byref(o) for o in out], 1.4, *map(float, wl[k]),
                               *map(float, wr[k]))
        v2 = [o.value for o in out]
        assert v2[0] == v[0] and v2[1] == v[1]
        assert v2[2] in (v[0] * wl[k, 2], v[0] * wr[k, 2])


# ---- against the reference build --------------------------------------------

def scalar_inputs(n=2000, seed=29):
    wl, wr = K.riemann_states(n, seed)
    gam = np.where(np.arange(n) % 2 == 0, 1.4, 5.0 / 3.0)
    return wl, wr, gam


def test_guess_and_coefficients_equal_the_reference(lib, ref):
    wl, wr, gam = scalar_inputs()
    for k in range(len(wl)):
        gk = float(gam[k])
        rl, ul, _, pl = map(float, wl[k])
        rr, ur, _, pr = map(float, wr[k])
        al = lib.get_sound_speed(gk, rl, pl) if rl > 0 else 0.0
        ar = lib.get_sound_speed(gk, rr, pr) if rr > 0 else 0.0
        if rl > 0:
            assert bits(al) == bits(ref.get_sound_speed(gk, rl, pl))
        assert bits(lib.riemann_A_L_or_R(gk, rl)) == bits(ref.riemann_A_L_or_R(gk, rl))
        assert bits(lib.riemann_B_L_or_R(gk, pl)) == bits(ref.riemann_B_L_or_R(gk, pl))
        for tol in (1e-6, 1e-8):
            a = lib.guess_p(gk, rl, ul, pl, al, rr, ur, pr, ar, tol)
            b = ref.guess_p(gk, rl, ul, pl, al, rr, ur, pr, ar, tol)
            assert bits(a) == bits(b), (k, a, b)


def test_vacuum_and_exact_samplers_equal_the_reference(lib, ref):
    """solve_vacuum_1d/2d and solve_exact_1d at sample speeds across every
    fan, shock and contact (x/t from -4 to 4)."""
    wl, wr, gam = scalar_inputs(600, 31)
    speeds = np.linspace(-4.0, 4.0, 17)
    o1 = [C.c_double() for _ in range(4)]
    o2 = [C.c_double() for _ in range(4)]
    for k in range(len(wl)):
        gk = float(gam[k])
        rl, ul, vl, pl = map(float, wl[k])
        rr, ur, vr, pr = map(float, wr[k])
        al = ref.get_sound_speed(gk, rl, pl) if rl > 0 else 0.0
        ar = ref.get_sound_speed(gk, rr, pr) if rr > 0 else 0.0
        for s in speeds:
            s = float(s)
            lib.solve_vacuum_1d(*[C.byref(o) for o in o1[:3]], gk, rl, ul, pl, al, rr, ur, pr, ar, s)
            ref.solve_vacuum_1d(*[C.byref(o) for o in o2[:3]], gk, rl, ul, pl, al, rr, ur, pr, ar, s)
            assert np.array_equal(bits([o.value for o in o1[:3]]), bits([o.value for o in o2[:3]])), (k, s)
            lib.solve_vacuum_2d(*[C.byref(o) for o in o1], gk, rl, ul, vl, pl, al, rr, ur, vr, pr, ar, s)
            ref.solve_vacuum_2d(*[C.byref(o) for o in o2], gk, rl, ul, vl, pl, al, rr, ur, vr, pr, ar, s)
            assert np.array_equal(bits([o.value for o in o1]), bits([o.value for o in o2])), (k, s)
            for o in o1 + o2:
                o.value = 0.0
            st1 = lib.solve_exact_1d(*[C.byref(o) for o in o1[:3]], gk, rl, ul, pl, rr, ur, pr, 1e-6, s, 0)
            st2 = ref.solve_exact_1d(*[C.byref(o) for o in o2[:3]], gk, rl, ul, pl, rr, ur, pr, 1e-6, s, 0)
            assert st1.return_code == st2.return_code, (k, s)
            if st1.return_code == 0:
                assert np.array_equal(bits([o.value for o in o1[:3]]), bits([o.value for o in o2[:3]])), (k, s)
            _free(lib, st1)
            _free(ref, st2)


def _pair(lib, ref, make):
    return make(lib), make(ref)


def test_time_steps_equal_the_reference(lib, ref):
    a, b = _pair(lib, ref, lambda L: K.setup_random_2d(L, 37, 23, 2, K.BC_SETS["mixed"], 5))
    for cfl in (0.8, 0.16):
        assert bits(lib.get_time_step_2d(C.byref(a.system), cfl)) == bits(ref.get_time_step_2d(C.byref(b.system), cfl))
    a, b = _pair(lib, ref, lambda L: K.setup_sod(L, 200))
    assert bits(lib.get_time_step_1d(C.byref(a.system), 0.9)) == bits(ref.get_time_step_1d(C.byref(b.system), 0.9))


@pytest.mark.parametrize("coord", ["cylindrical_1d", "spherical_1d"])
def test_geometry_source_equals_the_reference(lib, ref, coord):
    a, b = _pair(lib, ref, lambda L: K.setup_sod(L, 128, coord=coord, bc=("reflective", "transmissive")))
    for dt in (1e-3, 2.5e-4):
        sa = lib.add_geometry_source_term(C.byref(a.bc), C.byref(a.system), dt)
        sb = ref.add_geometry_source_term(C.byref(b.bc), C.byref(b.system), dt)
        assert sa.return_code == sb.return_code == 0
        for x, y in zip(K.state(a), K.state(b)):
            assert np.array_equal(bits(x), bits(y))


def test_geometry_source_rejects_cartesian(lib):
    a = K.setup_sod(lib, 16)
    st = lib.add_geometry_source_term(C.byref(a.bc), C.byref(a.system), 1e-3)
    assert st.return_code == abi.VALUE_ERROR
    assert "Wrong coordinate system, expected cylindrical 1D or spherical 1D." in abi.traceback_text(lib, st)


def test_gravity_source_equals_the_reference(lib, ref):
    a, b = _pair(lib, ref, lambda L: K.setup_random_2d(L, 19, 13, 2, K.BC_SETS["reflective"], 9, gravity=0.7))
    for dt in (1e-3, 4e-3):
        sa = lib.add_gravity_source_term_2d(C.byref(a.bc), C.byref(a.system), dt)
        sb = ref.add_gravity_source_term_2d(C.byref(b.bc), C.byref(b.system), dt)
        assert sa.return_code == sb.return_code == 0
        for x, y in zip(K.state(a), K.state(b)):
            assert np.array_equal(bits(x), bits(y))
