"""Device parity of random_choice_1d (integrator_random_choice_1d.c) against
the oracle: the sampled states come from the exact Riemann solver, whose
pow() may differ from glibc's in the last place, so runs are held to 1e-10
relative (most are bitwise)."""
import numpy as np
import pytest

import _cases as K
from test_gpu_parity_2d import compare

pytestmark = pytest.mark.gpu

CASES = [
    ("sod200", lambda L: K.setup_sod(L, 200), 0.5, 0.2, 1e-6),
    ("sod100_tol8", lambda L: K.setup_sod(L, 100), 0.5, 0.2, 1e-8),
    ("cyl64", lambda L: K.setup_sod(L, 64, coord="cylindrical_1d", bc=("reflective", "transmissive")), 0.5, 0.2, 1e-6),
    ("sph64", lambda L: K.setup_sod(L, 64, coord="spherical_1d", bc=("reflective", "transmissive")), 0.5, 0.2, 1e-6),
    ("sod50g3_periodic", lambda L: K.setup_sod(L, 50, g=3, bc=("periodic", "periodic")), 0.5, 0.05, 1e-6),
    ("sod60g1", lambda L: K.setup_sod(L, 60, g=1), 0.5, 0.1, 1e-6),
    ("sod40g1_none", lambda L: K.setup_sod(L, 40, g=1, bc=("none", "reflective")), 0.5, 0.1, 1e-6),
]


@pytest.mark.parametrize("key,mk,cfl,tf,tol", CASES, ids=[c[0] for c in CASES])
def test_random_choice_matches_oracle(gpulib, key, mk, cfl, tf, tol):
    sim = mk(gpulib)
    grid = K.oracle_grid(sim, "minmod", "riemann_solver_exact", tol=tol)
    res = sim.launch("random_choice_1d", "riemann_solver_exact", "minmod", cfl, tf, tol=tol)
    r = grid.run(cfl, tf=tf, random_choice=True)
    assert r.error == 0
    assert res.num_steps == r.num_steps and res.t == r.t
    assert abs(res.dt - r.dt) <= 1e-10 * r.dt
    W, U = K.state(sim)
    compare(f"random choice {key}", W[:, None], U[:, None], grid.W[:, None], grid.U[:, None], tol=1e-10)


def test_random_choice_step_limited(gpulib):
    sim = K.setup_sod(gpulib, 120)
    grid = K.oracle_grid(sim, "minmod", "riemann_solver_exact")
    res = sim.launch("random_choice_1d", "riemann_solver_exact", "minmod", 0.5, np.inf, max_steps=37)
    r = grid.run(0.5, max_steps=37, random_choice=True)
    assert res.num_steps == r.num_steps == 37
    W, U = K.state(sim)
    compare("random choice 37 steps", W[:, None], U[:, None], grid.W[:, None], grid.U[:, None], tol=1e-10)
