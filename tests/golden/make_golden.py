"""Generate the golden fixtures in tests/golden/ from the reference build.

Run in a container that has the reference sources (make -C oracle ref first):

    python tests/golden/make_golden.py

The reference library (oracle/_ref/libhydro_ref.so, compiled from the
reference's own C sources) is driven through its public API only.  Outputs:

  scalar_vectors.npz  limiter pairs, interface states and the reference's
                      HLLC (1D, 2D) and exact (1D) fluxes and return codes
  runs_2d.npz         initial and final states of full 2D runs
  runs_1d.npz         the same for 1D runs (Sod, curvilinear, boundaries)
  runs_godunov.npz    the same for godunov_first_order_{1d,2d} runs
  runs_random_choice.npz  the same for random_choice_1d runs
  sod_exact_t0.2.npz  copy of examples/sod_shock_1d/sol_cartesian_1d_0.2.npz,
                      the reference's exact Sod solution at t = 0.2 (data)
"""
from __future__ import annotations

import ctypes as C
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = os.path.dirname(HERE)
ROOT = os.path.dirname(TESTS)
sys.path[:0] = [ROOT, TESTS]

import _cases as K  # noqa: E402
import _oracle as O  # noqa: E402
from hydrofv_amd import abi  # noqa: E402

REF_EXAMPLES = "/root/reference/examples"


def scalar_vectors():
    out = {}
    wl, wr = K.riemann_states(3000, 11)
    out["wl"], out["wr"] = wl, wr
    f, e = O.ref_flux(2, "riemann_solver_hllc", 1.4, 1e-6, wl, wr)
    out["hllc2d_f"], out["hllc2d_err"] = f, e
    l3, r3 = wl[:, [0, 1, 3]], wr[:, [0, 1, 3]]
    f, e = O.ref_flux(1, "riemann_solver_hllc", 1.4, 1e-6, l3, r3)
    out["hllc1d_f"], out["hllc1d_err"] = f, e
    f, e = O.ref_flux(1, "riemann_solver_exact", 1.4, 1e-6, l3, r3)
    out["exact1d_f"], out["exact1d_err"] = f, e
    f, e = O.ref_flux(1, "riemann_solver_exact", 5.0 / 3.0, 1e-8, l3, r3)
    out["exact1d_g53_tol8_f"], out["exact1d_g53_tol8_err"] = f, e
    # limiter pairs incl. signs, ties, zeros and the Van Leer denominator guard
    a = np.concatenate([wl[:, 1] - wr[:, 1], [0.0, -0.0, 1.0, -1.0, 2.0, -2.0, 1e-300, -5e-11, 5e-11, 3.0]])
    b = np.concatenate([wr[:, 2], [1.0, 1.0, 1.0, -1.0, -2.0, -2.0, 1e-300, -5e-11, 5e-11, 3.0]])
    out["lim_a"], out["lim_b"] = a, b
    for name in K.LIMITERS:
        out[f"lim_{name}"] = O.ref_limit(name, a, b)
    return out


def _steps_then_ref(mk, limiter, steps, cfl, solver="riemann_solver_hllc", tol=1e-6, integrator="muscl_hancock_2d"):
    """Run the reference for exactly `steps` steps: the oracle gives t_N, then
    the reference runs to tf = t_N (its loop then stops after N steps)."""
    ref = abi.bind(C.CDLL(O.REF_SO))
    sim = mk(ref)
    W0, U0 = K.state(sim)
    grid = K.oracle_grid(sim, limiter, solver, tol=tol)
    r = grid.run(cfl, max_steps=steps, first_order=integrator.startswith("godunov"))
    assert r.error == 0, r.error
    res = sim.launch(integrator, solver, limiter, cfl, r.t, tol=tol)
    assert res.num_steps == steps, (res.num_steps, steps)
    W1, U1 = K.state(sim)
    return dict(W0=W0, U0=U0, W1=W1, U1=U1, t=res.t, dt=res.dt, n=res.num_steps,
                gamma=sim.system.gamma)


def runs_2d():
    out = {}
    for (nx, ny, g) in K.SHAPES:
        for bname, bcs in K.BC_SETS.items():
            for lim in K.LIMITERS:
                seed = 100 + nx * 7 + ny + g
                d = _steps_then_ref(lambda L: K.setup_random_2d(L, nx, ny, g, bcs, seed), lim, 11, 0.8)
                key = f"{nx}x{ny}g{g}_{bname}_{lim}"
                for k, v in d.items():
                    out[f"{key}/{k}"] = v
    d = _steps_then_ref(lambda L: K.setup_kh(L, 64), "monotonized_central", 50, 0.8)
    for k, v in d.items():
        out[f"kh64_mc_50/{k}"] = v
    d = _steps_then_ref(lambda L: K.setup_random_2d(L, 13, 9, 2, K.BC_SETS["reflective"], 11, gravity=0.7),
                        "van_leer", 11, 0.8)
    for k, v in d.items():
        out[f"gravity_13x9_vl/{k}"] = v
    return out


def runs_1d():
    out = {}
    ref = abi.bind(C.CDLL(O.REF_SO))
    cases = {
        "sod1000_hllc_minmod": (lambda L: K.setup_sod(L, 1000), "minmod", "riemann_solver_hllc", 0.9, 0.2),
        "sod200_exact_vl": (lambda L: K.setup_sod(L, 200), "van_leer", "riemann_solver_exact", 0.9, 0.2),
        "cyl64_hllc_mc": (lambda L: K.setup_sod(L, 64, coord="cylindrical_1d", bc=("reflective", "transmissive")),
                          "monotonized_central", "riemann_solver_hllc", 0.9, 0.2),
        "sph64_hllc_mc": (lambda L: K.setup_sod(L, 64, coord="spherical_1d", bc=("reflective", "transmissive")),
                          "monotonized_central", "riemann_solver_hllc", 0.9, 0.2),
        "sod50g3_periodic_minmod": (lambda L: K.setup_sod(L, 50, g=3, bc=("periodic", "periodic")), "minmod",
                                    "riemann_solver_hllc", 0.8, 0.05),
    }
    for key, (mk, lim, solver, cfl, tf) in cases.items():
        sim = mk(ref)
        W0, U0 = K.state(sim)
        res = sim.launch("muscl_hancock_1d", solver, lim, cfl, tf)
        W1, U1 = K.state(sim)
        for k, v in dict(W0=W0, U0=U0, W1=W1, U1=U1, t=res.t, dt=res.dt, n=res.num_steps).items():
            out[f"{key}/{k}"] = v
    return out


GODUNOV_SHAPES = [(13, 7, 1), (9, 11, 2)]


def runs_godunov():
    """First-order Godunov runs: 2D on tiny grids (1 and 2 ghost layers,
    every boundary set, 11 steps across the CFL-shrink boundary), a KH run
    and a gravity run; 1D shock tubes with both solvers, curvilinear
    geometry and one ghost layer."""
    out = {}
    for (nx, ny, g) in GODUNOV_SHAPES:
        for bname, bcs in K.BC_SETS.items():
            seed = 300 + nx * 7 + ny + g
            d = _steps_then_ref(lambda L: K.setup_random_2d(L, nx, ny, g, bcs, seed), "minmod", 11, 0.8,
                                integrator="godunov_first_order_2d")
            for k, v in d.items():
                out[f"2d_{nx}x{ny}g{g}_{bname}/{k}"] = v
    d = _steps_then_ref(lambda L: K.setup_kh(L, 64), "minmod", 40, 0.8, integrator="godunov_first_order_2d")
    for k, v in d.items():
        out[f"2d_kh64_40/{k}"] = v
    d = _steps_then_ref(lambda L: K.setup_random_2d(L, 13, 9, 1, K.BC_SETS["reflective"], 12, gravity=0.7),
                        "minmod", 11, 0.8, integrator="godunov_first_order_2d")
    for k, v in d.items():
        out[f"2d_gravity_13x9g1/{k}"] = v
    ref = abi.bind(C.CDLL(O.REF_SO))
    cases = {
        "1d_sod400_hllc": (lambda L: K.setup_sod(L, 400), "riemann_solver_hllc", 0.9, 0.2),
        "1d_sod200_exact": (lambda L: K.setup_sod(L, 200), "riemann_solver_exact", 0.9, 0.2),
        "1d_cyl64_hllc": (lambda L: K.setup_sod(L, 64, coord="cylindrical_1d", bc=("reflective", "transmissive")),
                          "riemann_solver_hllc", 0.9, 0.2),
        "1d_sph64_exact": (lambda L: K.setup_sod(L, 64, coord="spherical_1d", bc=("reflective", "transmissive")),
                           "riemann_solver_exact", 0.9, 0.2),
        "1d_sod50g1_periodic": (lambda L: K.setup_sod(L, 50, g=1, bc=("periodic", "periodic")),
                                "riemann_solver_hllc", 0.8, 0.05),
    }
    for key, (mk, solver, cfl, tf) in cases.items():
        sim = mk(ref)
        W0, U0 = K.state(sim)
        res = sim.launch("godunov_first_order_1d", solver, "minmod", cfl, tf)
        W1, U1 = K.state(sim)
        for k, v in dict(W0=W0, U0=U0, W1=W1, U1=U1, t=res.t, dt=res.dt, n=res.num_steps).items():
            out[f"{key}/{k}"] = v
    return out


def runs_random_choice():
    """random_choice_1d runs (exact solver sampling; CFL 0.5): Sod in the
    three geometries and a periodic tube, to t = 0.2 / 0.05."""
    out = {}
    ref = abi.bind(C.CDLL(O.REF_SO))
    cases = {
        "rc_sod200": (lambda L: K.setup_sod(L, 200), 0.5, 0.2, 1e-6),
        "rc_sod100_tol8": (lambda L: K.setup_sod(L, 100), 0.5, 0.2, 1e-8),
        "rc_cyl64": (lambda L: K.setup_sod(L, 64, coord="cylindrical_1d", bc=("reflective", "transmissive")),
                     0.5, 0.2, 1e-6),
        "rc_sph64": (lambda L: K.setup_sod(L, 64, coord="spherical_1d", bc=("reflective", "transmissive")),
                     0.5, 0.2, 1e-6),
        "rc_sod50g3_periodic": (lambda L: K.setup_sod(L, 50, g=3, bc=("periodic", "periodic")), 0.5, 0.05, 1e-6),
    }
    for key, (mk, cfl, tf, tol) in cases.items():
        sim = mk(ref)
        W0, U0 = K.state(sim)
        res = sim.launch("random_choice_1d", "riemann_solver_exact", "minmod", cfl, tf, tol=tol)
        W1, U1 = K.state(sim)
        for k, v in dict(W0=W0, U0=U0, W1=W1, U1=U1, t=res.t, dt=res.dt, n=res.num_steps).items():
            out[f"{key}/{k}"] = v
    return out


def main():
    if not O.ref_available():
        sys.exit("build the reference first: make -C oracle ref")
    if sys.argv[1:] == ["godunov"]:
        np.savez_compressed(os.path.join(HERE, "runs_godunov.npz"), **runs_godunov())
        return
    if sys.argv[1:] == ["random_choice"]:
        np.savez_compressed(os.path.join(HERE, "runs_random_choice.npz"), **runs_random_choice())
        return
    np.savez_compressed(os.path.join(HERE, "scalar_vectors.npz"), **scalar_vectors())
    np.savez_compressed(os.path.join(HERE, "runs_2d.npz"), **runs_2d())
    np.savez_compressed(os.path.join(HERE, "runs_1d.npz"), **runs_1d())
    np.savez_compressed(os.path.join(HERE, "runs_godunov.npz"), **runs_godunov())
    np.savez_compressed(os.path.join(HERE, "runs_random_choice.npz"), **runs_random_choice())
    src = os.path.join(REF_EXAMPLES, "sod_shock_1d", "sol_cartesian_1d_0.2.npz")
    if os.path.exists(src):
        shutil.copyfile(src, os.path.join(HERE, "sod_exact_t0.2.npz"))
    for f in sorted(os.listdir(HERE)):
        print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
