"""Batched ensemble engine: every instance of a batched run equals a
single-grid device run of the same instance bit for bit, whatever tile shapes
the batch uses, and matches the oracle to 1e-12; instances keep their own
CFL dt."""
import numpy as np
import pytest

import _cases as K
from hydrofv_amd import ensemble, gpu
from test_gpu_parity_2d import compare

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_batch(gpulib):
    sides, cfgs, init = ensemble.build_instances(gpulib, count=10, seed=7, side_lo=8, side_hi=150, steps=25)
    return sides, cfgs, init


def single_run(lib, cfg, init, steps):
    m = gpu.MH2D(lib, cfg)
    m.upload(*init)
    m.start()
    m.advance(steps - 1, use_graph=False)
    st = m.status()
    W, U = m.download()
    m.close()
    return st, np.stack(W), np.stack(U)


@pytest.mark.parametrize("rows", ["default", "prop8", "fixed16"])
def test_batch_equals_single_runs(gpulib, small_batch, rows):
    sides, cfgs, init = small_batch
    ens = ensemble.Ensemble(gpulib, cfgs, init)
    if rows != "default":
        kind, value = ("prop", 8) if rows == "prop8" else ("fixed", 16)
        ens.set_strip_rows(ensemble.strip_rule(sides, kind, value))
    ens.start()
    ens.advance(60, use_graph=True)  # more launches than steps: finished instances only forward
    st = ens.status()
    for i in range(len(cfgs)):
        assert st[i].err_kind == 0 and st[i].num_steps == 25 and st[i].finished
        s1, W1, U1 = single_run(gpulib, cfgs[i], init[i], 25)
        assert st[i].t == s1.t and st[i].dt == s1.dt
        W, U = ens.download(i)
        assert np.array_equal(np.stack(U), U1), i
        assert np.array_equal(np.stack(W), W1, equal_nan=True), i
    # instances of different sizes take different time steps
    assert len({st[i].dt for i in range(len(cfgs))}) > 1
    ens.close()


def test_batch_matches_oracle(gpulib, small_batch):
    sides, cfgs, init = small_batch
    ens = ensemble.Ensemble(gpulib, cfgs, init)
    ens.start()
    ens.advance(24, use_graph=False)
    for i in (0, 3):
        W, U = ens.download(i)
        n = int(sides[i])
        grid = K.oracle_grid(_sim_like(gpulib, n, init[i]), "minmod")
        r = grid.run(0.8, max_steps=25)
        assert r.error == 0
        compare(f"ensemble instance {i}", np.stack(W), np.stack(U), grid.W, grid.U)
    ens.close()


def _sim_like(lib, n, init):
    from hydrofv_amd import driver
    s = driver.Simulation(lib, "cartesian_2d", 1.4, n, n, 2)
    s.set_boundary("transmissive", "transmissive", "transmissive", "transmissive")
    W, U = init
    s.set_primitive(W)
    for name, a in zip(["mass", "momentum_x", "momentum_y", "energy"], U):
        s.field(name)[...] = a
    return s
