"""A C program written against include/hydrofv/hydro.h alone
(tests/native/kh_driver.c, built by tests/native/Makefile) runs a
Kelvin-Helmholtz shear layer through launch_simulation on the device, in the
call order of the reference's example drivers; its result is checked
against the oracle run from the same initial state."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

import _cases as K
import _oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "native", "kh_driver")
TOL = 1e-12  # relative, per field (a power may differ from glibc's in the last place)


def read_output(path):
    with open(path, "rb") as f:
        nx, ny, g = np.frombuffer(f.read(12), dtype=np.int32)
        steps = int(np.frombuffer(f.read(8), dtype=np.int64)[0])
        t, dt, dt0 = np.frombuffer(f.read(24), dtype=np.float64)
        shape = (4, ny + 2 * g, nx + 2 * g)
        n = int(np.prod(shape))
        arrs = [np.frombuffer(f.read(8 * n), dtype=np.float64).reshape(shape) for _ in range(4)]
        assert f.read() == b""
    return int(nx), int(g), steps, t, dt, dt0, arrs


@pytest.mark.parametrize("n,tf", [(96, 0.05), (130, 0.02)])
def test_c_driver_matches_the_oracle(tmp_path, n, tf):
    assert os.path.exists(DRIVER), "tests/native/kh_driver is not built (make -C tests/native)"
    out = str(tmp_path / "kh.bin")
    print(f"C driver: KH {n}^2 to t={tf}", flush=True)
    r = subprocess.run([DRIVER, str(n), str(tf), out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    nx, g, steps, t, dt, dt0, (W0, U0, W1, U1) = read_output(out)
    def grid():
        G = O.Grid("cartesian_2d", nx, nx, g, 5.0 / 3.0, bc=("periodic",) * 4, limiter="monotonized_central",
                   solver="riemann_solver_hllc", tol=1e-6)
        G.W[...] = W0
        G.U[...] = U0
        return G

    # get_time_step_2d of the initial state, called by the driver, is the
    # oracle's first dt
    assert dt0 == grid().run(0.8, max_steps=1).dt
    grid = grid()
    res = grid.run(0.8, tf=tf)
    assert res.error == 0
    assert steps == res.num_steps and steps > 10
    assert abs(t - res.t) <= TOL * res.t and abs(dt - res.dt) <= TOL * res.dt
    worst = max(max(K.rel_err(W1[k], grid.W[k]), K.rel_err(U1[k], grid.U[k])) for k in range(4))
    bitwise = np.array_equal(W1.view(np.uint64), grid.W.view(np.uint64)) and \
        np.array_equal(U1.view(np.uint64), grid.U.view(np.uint64))
    print(f"  {steps} steps, max rel diff {worst:.2e}, bitwise {bitwise}", flush=True)
    assert worst <= TOL
