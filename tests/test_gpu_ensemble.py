"""Batched ensemble engine: every instance of a batched run equals a
single-grid device run of the same instance bit for bit, whatever tile shapes
the batch uses and in either HLLC code form, and matches the oracle to
1e-12; instances keep their own CFL dt."""
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


@pytest.mark.parametrize("rows,form", [("default", 0), ("prop8", 0), ("fixed16", 0), ("default", 1), ("prop8", 1)])
def test_batch_equals_single_runs(gpulib, small_batch, rows, form):
    """form: the batch's HLLC code form (selects 0, branches 1); the single
    runs keep the selects, so the forms are checked against each other too."""
    sides, cfgs, init = small_batch
    ens = ensemble.Ensemble(gpulib, cfgs, init)
    if rows != "default":
        kind, value = ("prop", 8) if rows == "prop8" else ("fixed", 16)
        ens.set_strip_rows(ensemble.strip_rule(sides, kind, value))
    ens.set_hllc_form(form)
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


def test_checkpoint_restore_replays_the_same_run(gpulib, small_batch):
    """A state kept by checkpoint() and put back by restore() continues
    exactly as the first time: the same steps, times and bits, through graph
    replays and plain launches alike."""
    sides, cfgs, init = small_batch
    ens = ensemble.Ensemble(gpulib, cfgs, init)
    with pytest.raises(RuntimeError):
        ens.restore()  # nothing kept yet
    ens.start()
    ens.advance(5, use_graph=False)
    ens.checkpoint()
    ens.advance(50, use_graph=True)
    st1 = ens.status()
    U1 = [np.stack(ens.download(i)[1]) for i in range(len(cfgs))]
    ens.restore()
    ens.advance(50, use_graph=True)
    st2 = ens.status()
    for i in range(len(cfgs)):
        assert (st1[i].num_steps, st1[i].t, st1[i].dt) == (st2[i].num_steps, st2[i].t, st2[i].dt), i
        assert np.array_equal(U1[i].view(np.uint64), np.stack(ens.download(i)[1]).view(np.uint64)), i
    ens.close()


def test_autotune_picks_a_timed_candidate(gpulib, small_batch):
    """The autotuner leaves the batch in its fastest configuration (strips
    and code form) with the initial state restored, so a run after it
    equals a plain run."""
    import torch
    sides, cfgs, init = small_batch
    ens = ensemble.Ensemble(gpulib, cfgs, init)
    best, results = ensemble.autotune(ens, sides, torch.cuda.Stream(), trial_launches=2)
    assert len(results) == 2 * len(ensemble.STRIP_RULES) and {r[0] for r in results} == {0, 1}
    assert best[3] == min(r[3] for r in results)
    ens.start()
    ens.advance(24, use_graph=False)
    st = ens.status()
    ref = ensemble.Ensemble(gpulib, cfgs, init)
    ref.start()
    ref.advance(24, use_graph=False)
    st_ref = ref.status()
    for i in range(len(cfgs)):
        assert st[i].num_steps == st_ref[i].num_steps and st[i].t == st_ref[i].t
        assert np.array_equal(np.stack(ens.download(i)[1]), np.stack(ref.download(i)[1])), i
    ens.close()
    ref.close()
