"""Problem set-ups shared by the golden-fixture generator, the oracle tests and
the device parity tests.  Every initial state is built through the library
API (system_init, boundary fill, primitive->conserved conversion) exactly the
way the reference's example drivers do it, so a case can be run by the
reference build, by libhydrofv_amd.so, or handed to the oracle."""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from hydrofv_amd import driver, workloads  # noqa: E402

LIMITERS = ["minmod", "van_leer", "monotonized_central"]
BC_SETS = {
    "periodic": ("periodic",) * 4,
    "transmissive": ("transmissive",) * 4,
    "reflective": ("reflective",) * 4,
    "mixed": ("reflective", "transmissive", "periodic", "periodic"),
    "none": ("none", "none", "transmissive", "reflective"),
}
SHAPES = [(13, 7, 2), (9, 11, 3)]


def random_state_2d(nx, ny, seed, rho=(0.5, 1.5), p=(0.5, 1.5), vel=0.3):
    """Cell-wise random primitive state (splitmix64 stream ``seed``)."""
    r = workloads.uniform01(seed, 4 * nx * ny).reshape(4, ny, nx)
    return [rho[0] + (rho[1] - rho[0]) * r[0], vel * (2 * r[1] - 1), vel * (2 * r[2] - 1),
            p[0] + (p[1] - p[0]) * r[3]]


def setup_random_2d(lib, nx, ny, g, bcs, seed, gamma=1.4, gravity=0.0):
    """Random interior state; every ghost cell too when a side has no
    boundary fill ("none"), as a driver must then provide it."""
    s = driver.Simulation(lib, "cartesian_2d", gamma, nx, ny, g, gravity=gravity)
    s.set_boundary(*bcs)
    if "none" in bcs:
        s.set_primitive(random_state_2d(nx + 2 * g, ny + 2 * g, seed))
    else:
        s.set_primitive(random_state_2d(nx, ny, seed), interior_only=True)
    s.apply_boundary()
    s.primitive_to_conserved()
    return s


def setup_blob_2d(lib, nx, ny, g, bcs, gamma=1.4, gravity=0.0, flow=(1.0, 0.4, -0.3, 1.0), radius=0.08):
    """A uniform flow `flow` (rho, u, v, p) with a denser, hotter disc at the
    centre of [0,1]^2: away from the disc every cell holds the same state."""
    s = driver.Simulation(lib, "cartesian_2d", gamma, nx, ny, g, gravity=gravity)
    s.set_boundary(*bcs)
    # every cell, ghost cells included, when a side has no boundary fill
    X, Y = np.meshgrid(s.mid_x, s.mid_y)
    W = [np.full(X.shape, flow[k]) for k in range(4)]
    disc = (X - 0.5) ** 2 + (Y - 0.5) ** 2 < radius ** 2
    W[0] = np.where(disc, 2.0 * flow[0], W[0])
    W[3] = np.where(disc, 3.0 * flow[3], W[3])
    if "none" in bcs:
        s.set_primitive(W)
    else:
        s.set_primitive([w[g:g + ny, g:g + nx] for w in W], interior_only=True)
    s.apply_boundary()
    s.primitive_to_conserved()
    return s


from hydrofv_amd import problems  # noqa: E402

setup_kh = problems.kelvin_helmholtz
setup_sedov = problems.sedov
setup_sod = problems.sod
setup_quadrant = problems.four_quadrant


def state(sim):
    return np.stack(sim.primitive()), np.stack(sim.conserved())


def oracle_grid(sim, limiter, solver="riemann_solver_hllc", bcs=None, tol=1e-6, nthreads=1):
    """Oracle state object holding a copy of a Simulation's current state."""
    import _oracle as O
    sysm = sim.system
    coord = sysm.coord_sys.decode()
    if bcs is None:
        names = [sim.bc.boundary_condition_x_min, sim.bc.boundary_condition_x_max,
                 sim.bc.boundary_condition_y_min, sim.bc.boundary_condition_y_max]
        bcs = tuple(n.decode() if n else None for n in names)
    b = sim.bc
    if sim.dim == 2:
        hooks = (b.set_boundary_condition_cartesian_2d_x_min, b.set_boundary_condition_cartesian_2d_x_max,
                 b.set_boundary_condition_cartesian_2d_y_min, b.set_boundary_condition_cartesian_2d_y_max)
    else:
        hooks = (b.set_boundary_condition_1d_x_min, b.set_boundary_condition_1d_x_max, None, None)
    grid = O.Grid(coord, sim.nx, sim.ny, sim.g, sysm.gamma, x=(sysm.x_min, sysm.x_max),
                  y=(sysm.y_min, sysm.y_max), bc=bcs, limiter=limiter, solver=solver, tol=tol,
                  gravity=sysm.gravity, nthreads=nthreads, hooks=hooks)
    W, U = state(sim)
    grid.W[...] = W
    grid.U[...] = U
    return grid


def riemann_states(n, seed):
    """Interface state pairs (rho, u, v, p) covering the solver branches:
    mild jumps (PVRS guess), strong rarefactions (two-rarefaction guess and
    vacuum generation), strong shocks (two-shock guess), vacuum on either
    side, supersonic flow both ways, equal states and a NaN pressure."""
    from hydrofv_amd import workloads
    u = workloads.uniform01(seed, n * 8).reshape(n, 8)
    wl = np.empty((n, 4))
    wr = np.empty((n, 4))
    wl[:, 0] = 0.1 + 2.0 * u[:, 0]
    wr[:, 0] = 0.1 + 2.0 * u[:, 1]
    wl[:, 1] = 2.0 * u[:, 2] - 1.0
    wr[:, 1] = 2.0 * u[:, 3] - 1.0
    wl[:, 2] = 2.0 * u[:, 4] - 1.0
    wr[:, 2] = 2.0 * u[:, 5] - 1.0
    wl[:, 3] = 0.1 + 2.0 * u[:, 6]
    wr[:, 3] = 0.1 + 2.0 * u[:, 7]
    cat = np.arange(n) % 10
    # 1: strong rarefaction, 2: vacuum generation, 3: strong shock / huge pressure ratio
    wl[cat == 1, 1] -= 2.0
    wr[cat == 1, 1] += 2.0
    wl[cat == 2, 1] -= 8.0
    wr[cat == 2, 1] += 8.0
    wl[cat == 3, 1] += 3.0
    wr[cat == 3, 1] -= 3.0
    wl[cat == 3, 3] *= 1e3
    # 4/5: vacuum on the left / right; 6: both sides vacuum
    wl[cat == 4, 0] = 0.0
    wr[cat == 5, 0] = 0.0
    wl[cat == 6, 0] = 0.0
    wr[cat == 6, 0] = 0.0
    # 7: supersonic to the right, 8: supersonic to the left
    wl[cat == 7, 1] += 6.0
    wr[cat == 7, 1] += 6.0
    wl[cat == 8, 1] -= 6.0
    wr[cat == 8, 1] -= 6.0
    # 9: identical states (and one NaN pressure)
    wr[cat == 9] = wl[cat == 9]
    if n > 19:
        wl[19, 3] = np.nan
    return wl, wr


def rel_err(a, b):
    """Largest absolute difference over the finite entries of b, relative to
    the largest magnitude among them."""
    a = np.asarray(a)
    b = np.asarray(b)
    fin = np.isfinite(b)
    scale = np.maximum(np.abs(b[fin]).max() if fin.any() else 1.0, 1e-300)
    d = np.abs(a[fin] - b[fin]).max() if fin.any() else 0.0
    return d / scale


def pow_special_grid():
    """Every pairing of special and ordinary x with special and ordinary y."""
    xs = [0.0, 1e-310, 2.5e-308, 0.5, 1.0, 2.0, 3.0, 1e300, 1.7976931348623157e308, np.inf]
    xs = xs + [-v for v in xs] + [np.nan]
    ys = [0.0, 0.2, 0.5, 1.0, 2.0, 3.0, 5.0, 7.0, 5.000000000000001, 1e20, 2.0 ** 53 + 2, np.inf, 2000.0]
    ys = ys + [-v for v in ys] + [np.nan]
    gx, gy = np.meshgrid(np.array(xs), np.array(ys))
    return gx.ravel(), gy.ravel()


def assert_pow_special(out, x, y):
    """out against glibc's pow: the same zero, infinity, NaN and sign; finite
    nonzero results within 1 ulp."""
    import ctypes as C
    libm = C.CDLL("libm.so.6")
    libm.pow.restype = C.c_double
    libm.pow.argtypes = [C.c_double, C.c_double]
    ref = np.array([libm.pow(a, b) for a, b in zip(x.tolist(), y.tolist())])
    same_kind = (np.isnan(out) == np.isnan(ref)) & (np.isinf(out) == np.isinf(ref)) & ((out == 0) == (ref == 0))
    same_sign = np.isnan(ref) | (np.signbit(out) == np.signbit(ref))
    bad = ~(same_kind & same_sign)
    assert not bad.any(), list(zip(x[bad], y[bad], out[bad], ref[bad]))[:10]
    fin = np.isfinite(ref) & (ref != 0)
    with np.errstate(over="ignore"):
        assert np.all(np.abs(out[fin] - ref[fin]) <= np.spacing(np.abs(ref[fin])))
    special = ~fin
    np.testing.assert_array_equal(out[special], ref[special])


# -- custom boundary hooks (boundary.h:17-33) ------------------------------
def _plane(ptr, g, nx, ny):
    tx, ty = nx + 2 * g, ny + 2 * g
    return np.ctypeslib.as_array(ptr, shape=(ty * tx,)).reshape(ty, tx)


def lid_hook(lid_u=1.0, lid_v=0.0):
    """y_max hook of a lid-driven cavity: every ghost row above the top wall,
    corner columns included, mirrors density and pressure of the rows below
    and takes the lid's velocity.  The corner columns read x-ghost cells, so
    the result depends on the x sides being filled first."""
    def hook(rho, u, v, p, g, nx, ny):
        R, U, V, P = (_plane(a, g, nx, ny) for a in (rho, u, v, p))
        for j in range(g):
            R[g + ny + j, :] = R[g + ny - 1 - j, :]
            P[g + ny + j, :] = P[g + ny - 1 - j, :]
            U[g + ny + j, :] = lid_u
            V[g + ny + j, :] = lid_v
    return hook


def inflow_hook_2d(state=(1.2, 0.8, 0.1, 1.1)):
    """x_min hook: a fixed inflow state in every ghost column of the
    interior rows."""
    def hook(rho, u, v, p, g, nx, ny):
        for a, val in zip((rho, u, v, p), state):
            _plane(a, g, nx, ny)[g:g + ny, :g] = val
    return hook


def copy_hook_2d(side):
    """A hook that copies the nearest interior layer outwards (zero
    gradient), for side 1 (x_max), 2 (y_min) or 3 (y_max)."""
    def hook(rho, u, v, p, g, nx, ny):
        for a in (rho, u, v, p):
            A = _plane(a, g, nx, ny)
            for i in range(g):
                if side == 1:
                    A[g:g + ny, g + nx + i] = A[g:g + ny, g + nx - 1]
                elif side == 2:
                    A[g - 1 - i, g:g + nx] = A[g, g:g + nx]
                else:
                    A[g + ny + i, g:g + nx] = A[g + ny - 1, g:g + nx]
    return hook


def _counted(hook, calls):
    if calls is None:
        return hook

    def wrapped(*args):
        calls.append(1)
        hook(*args)
    return wrapped


def setup_cavity(lib, n, g=2, device=False, calls=None, bcs=("reflective", "reflective", "reflective")):
    """Lid-driven cavity: walls `bcs` (x_min, x_max, y_min), a custom moving
    lid on top.  device: register the lid's device form as well (2D runs
    then fill it on the device); calls: a list that grows by one per host
    hook call."""
    s = driver.Simulation(lib, "cartesian_2d", 1.4, n, n, g)
    name = "set_boundary_condition_cartesian_2d_y_max"
    dev = None
    if device:
        import _devhooks
        dev = {name: _devhooks.lid()}
    s.set_boundary(*bcs, "custom", custom={name: _counted(lid_hook(), calls)}, device=dev)
    one = np.ones((n, n))
    s.set_primitive([one, 0 * one, 0 * one, one / 1.4], interior_only=True)
    s.apply_boundary()
    s.primitive_to_conserved()
    return s


def setup_all_custom(lib, nx, ny, g, seed, device=False, calls=None, device_sides=(0, 1, 2, 3)):
    """Random interior state with every side filled by a hook; device and
    calls as for setup_cavity (device forms for `device_sides` only)."""
    s = driver.Simulation(lib, "cartesian_2d", 1.4, nx, ny, g)
    names = ["set_boundary_condition_cartesian_2d_" + k for k in ("x_min", "x_max", "y_min", "y_max")]
    hooks = [inflow_hook_2d(), copy_hook_2d(1), copy_hook_2d(2), copy_hook_2d(3)]
    dev = None
    if device:
        import _devhooks
        forms = [_devhooks.inflow(), _devhooks.copy(1), _devhooks.copy(2), _devhooks.copy(3)]
        dev = {names[k]: forms[k] for k in device_sides}
    s.set_boundary("custom", "custom", "custom", "custom",
                   custom={n: _counted(h, calls) for n, h in zip(names, hooks)}, device=dev)
    s.set_primitive(random_state_2d(nx, ny, seed), interior_only=True)
    s.apply_boundary()
    s.primitive_to_conserved()
    return s


def inflow_hook_1d(state=(1.5, 0.9, 1.3)):
    def hook(rho, u, p, g, nx):
        for a, val in zip((rho, u, p), state):
            np.ctypeslib.as_array(a, shape=(nx + 2 * g,))[:g] = val
    return hook


def setup_inflow_1d(lib, n, g=2):
    """Sod tube whose left ghost cells hold a fixed inflow state (custom)."""
    s = driver.Simulation(lib, "cartesian_1d", 1.4, n, g=g)
    s.set_primitive(workloads.sod(n, g))
    s.set_boundary("custom", "transmissive", custom={"set_boundary_condition_1d_x_min": inflow_hook_1d()})
    s.apply_boundary()
    s.primitive_to_conserved()
    return s


def field_errors(A, B):
    """Per-field differences of A against B (arrays of fields): the L1
    relative difference sum|A-B| / sum|B|, the largest difference relative
    to the field's largest magnitude, and the largest elementwise relative
    difference |A-B| / |B| over the entries with B != 0 (an entry where B is
    0 and A is not counts as infinite).  Entries where B is not finite are
    compared for equality of their bit patterns instead, and a mismatch
    there counts as infinite too."""
    out = []
    for a, b in zip(A, B):
        a = np.asarray(a, dtype=np.float64)
        b = np.asarray(b, dtype=np.float64)
        fin = np.isfinite(b)
        bad = bool((a[~fin].view(np.uint64) != b[~fin].view(np.uint64)).any())
        d = np.abs(a[fin] - b[fin])
        bb = np.abs(b[fin])
        l1 = d.sum() / max(bb.sum(), 1e-300)
        mx = d.max() / max(bb.max(), 1e-300) if d.size else 0.0
        nz = bb > 0
        el = (d[nz] / bb[nz]).max() if nz.any() else 0.0
        if bad or bool((d[~nz] > 0).any()):
            el = np.inf
        out.append((float(l1), float(mx), float(el)))
    return out


def format_field_errors(errs, names=("m", "mx", "my", "E")):
    return ", ".join(f"{n}: L1 {e[0]:.1e} max {e[1]:.1e} elem {e[2]:.1e}" for n, e in zip(names, errs))
