"""The oracle's custom boundary hooks (boundary.h:17-33, called side by side
in the reference's order, boundary.c:411-777) against the reference build
driven with the same hooks: bitwise equal states after every run."""
import ctypes as C

import numpy as np
import pytest

import _cases as K
import _oracle as O

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) missing")

CASES = [
    ("cavity", lambda L: K.setup_cavity(L, 24), "muscl_hancock_2d", "minmod", 30),
    ("all_custom", lambda L: K.setup_all_custom(L, 17, 11, 2, 5), "muscl_hancock_2d", "van_leer", 20),
    ("all_custom_godunov", lambda L: K.setup_all_custom(L, 17, 11, 1, 6), "godunov_first_order_2d", "minmod", 20),
    ("inflow_1d", lambda L: K.setup_inflow_1d(L, 60), "muscl_hancock_1d", "monotonized_central", 40),
    ("inflow_1d_godunov", lambda L: K.setup_inflow_1d(L, 60, g=1), "godunov_first_order_1d", "minmod", 40),
]


@pytest.mark.parametrize("name,mk,integrator,limiter,steps", CASES, ids=[c[0] for c in CASES])
def test_oracle_hooks_equal_reference(name, mk, integrator, limiter, steps):
    from hydrofv_amd import abi
    ref = abi.bind(C.CDLL(O.REF_SO))
    sim = mk(ref)
    grid = K.oracle_grid(sim, limiter, "riemann_solver_hllc")
    r = grid.run(0.8, max_steps=steps, first_order=integrator.startswith("godunov"))
    assert r.error == 0
    res = sim.launch(integrator, "riemann_solver_hllc", limiter, 0.8, r.t)
    assert (res.num_steps, res.t) == (r.num_steps, r.t)
    W, U = K.state(sim)
    assert np.array_equal(U, grid.U.reshape(U.shape), equal_nan=True)
    assert np.array_equal(W, grid.W.reshape(W.shape), equal_nan=True)
