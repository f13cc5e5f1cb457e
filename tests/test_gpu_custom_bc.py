"""Custom boundary hooks on the device runs (boundary.h:17-33): every step's
primitive state comes to the host, where the reference's boundary fill runs
with the hooks, and goes back for the next step.  Each run is compared with
the oracle running the same hooks (which the CPU suite pins to the reference
build), through launch_simulation and the step-limited entry points."""
import numpy as np
import pytest

import _cases as K
from test_gpu_parity_2d import compare
from test_oracle_custom_bc import CASES

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,mk,integrator,limiter,steps", CASES, ids=[c[0] for c in CASES])
def test_custom_hooks_match_oracle(gpulib, name, mk, integrator, limiter, steps):
    sim = mk(gpulib)
    grid = K.oracle_grid(sim, limiter, "riemann_solver_hllc")
    res = sim.launch(integrator, "riemann_solver_hllc", limiter, 0.8, np.inf, max_steps=steps)
    r = grid.run(0.8, max_steps=steps, first_order=integrator.startswith("godunov"))
    assert r.error == 0 and (res.num_steps, res.t, res.dt) == (r.num_steps, r.t, r.dt)
    W, U = K.state(sim)
    if sim.dim == 1:
        W, U = W[:, None], U[:, None]
    compare(f"custom {name}", W, U, grid.W.reshape(W.shape), grid.U.reshape(U.shape))


def test_cavity_to_tf_with_snapshots(gpulib, tmp_path):
    """The reference's lid-driven cavity set-up: the run-to-tf loop with
    snapshots, the lid hook applied between every two steps."""
    import _h5 as H
    if not H.available():
        pytest.skip("HDF5 library missing")
    sim = K.setup_cavity(gpulib, 32)
    grid = K.oracle_grid(sim, "monotonized_central", "riemann_solver_hllc")
    out = str(tmp_path) + "/"
    res = sim.launch("muscl_hancock_2d", "riemann_solver_hllc", "monotonized_central", 0.8, 0.05,
                     store_dir=out, store_interval=0.02, store_initial=True)
    r = grid.run(0.8, tf=0.05)
    assert r.error == 0 and (res.num_steps, res.t, res.dt) == (r.num_steps, r.t, r.dt)
    W, U = K.state(sim)
    compare("cavity to tf", W, U, grid.W, grid.U)
    import glob
    assert len(glob.glob(out + "snapshot_*.h5")) == 3
