"""hfv_mh2d_run_noalloc, the 2D run that allocates nothing, against the
allocating entry points (hfv_mh2d_create / upload / start / advance /
download) on the same input: the same state bit for bit and the same step
count, t and dt, over two calls in a row, the second on a smaller grid."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

import _cases as K
from hydrofv_amd import gpu, problems

pytestmark = pytest.mark.gpu


def allocating_run(lib, cfg, W0, U0, launches):
    m = gpu.MH2D(lib, cfg)
    m.upload(W0, U0)
    m.start()
    m.advance(launches - 1, use_graph=True)
    st = m.status()
    W, U = m.download()
    m.close()
    return st, np.stack(W), np.stack(U)


def noalloc_run(lib, cfg, W0, U0, launches, pad, sync_status):
    """The caller's device arrays are torch tensors with rows `tx + pad`
    doubles apart; the entry point gets their addresses."""
    ty, tx = W0[0].shape
    pitch = tx + pad
    dev = []
    for a in list(W0) + list(U0):
        t = torch.full((ty, pitch), float("nan"), dtype=torch.float64, device="cuda")
        t[:, :tx] = torch.from_numpy(np.ascontiguousarray(a))
        dev.append(t)
    Wp = (C.c_void_p * 4)(*[t.data_ptr() for t in dev[:4]])
    Up = (C.c_void_p * 4)(*[t.data_ptr() for t in dev[4:]])
    torch.cuda.synchronize()
    st = gpu.RunStatus()
    rc = lib.hfv_mh2d_run_noalloc(C.byref(cfg), Wp, Up, pitch, launches, C.byref(st) if sync_status else None,
                                  None)
    gpu.check(lib, rc, "hfv_mh2d_run_noalloc")
    torch.cuda.synchronize()
    W = np.stack([t[:, :tx].cpu().numpy() for t in dev[:4]])
    U = np.stack([t[:, :tx].cpu().numpy() for t in dev[4:]])
    # the padding columns are not the run's to write
    assert all(bool(torch.isnan(t[:, tx:]).all()) for t in dev)
    return (st if sync_status else None), W, U


CASES = [
    # (label, set-up, limiter flag, launches, extra config, pad, status read)
    ("KH 150^2, MC, 30 steps", lambda lib: K.setup_kh(lib, 150), 3, 36, dict(max_steps=30), 5, True),
    ("random 41x23, mixed sides, minmod, to tf", lambda lib: K.setup_random_2d(lib, 41, 23, 2, K.BC_SETS["mixed"], 7),
     1, 60, dict(tf=0.02), 0, False),
]


def test_noalloc_run_equals_allocating_run(gpulib):
    for label, setup, limiter, launches, extra, pad, sync_status in CASES:
        print(f"no-alloc case: {label}", flush=True)
        sim = setup(gpulib)
        W0, U0 = sim.primitive(), sim.conserved()
        cfg = problems.device_config(sim, limiter=limiter, solver=2, cfl=0.8, **extra)
        st_a, Wa, Ua = allocating_run(gpulib, cfg, W0, U0, launches)
        st_b, Wb, Ub = noalloc_run(gpulib, cfg, W0, U0, launches, pad, sync_status)
        sim.close()
        assert st_a.err_kind == 0 and st_a.finished == 1, label
        # bit patterns: the corner primitives are c2p of zero states (NaN)
        assert np.array_equal(Ua.view(np.uint64), Ub.view(np.uint64)), label
        assert np.array_equal(Wa.view(np.uint64), Wb.view(np.uint64)), label
        if st_b is not None:
            assert (st_b.num_steps, st_b.t, st_b.dt, st_b.err_kind, st_b.finished) == \
                (st_a.num_steps, st_a.t, st_a.dt, st_a.err_kind, st_a.finished), label


def test_noalloc_rejects_grids_beyond_its_scratch(gpulib):
    sim = K.setup_random_2d(gpulib, 13, 7, 2, K.BC_SETS["periodic"], 3)
    cfg = problems.device_config(sim, limiter=1, solver=2, cfl=0.8, max_steps=1)
    sim.close()
    cfg.num_cells_x = cfg.num_cells_y = 2000   # 4 x 2004 x 2004 doubles > 64 MiB
    pitch = 2004
    null = (C.c_void_p * 4)(*([1] * 4))  # never dereferenced: the size check comes first
    rc = gpulib.hfv_mh2d_run_noalloc(C.byref(cfg), null, null, pitch, 1, None, None)
    assert rc == gpu.HFV_ETOOBIG
    assert 4 * pitch * 2004 * 8 > gpu.HFV_NOALLOC_SCRATCH_BYTES
    # the exact solver and device-side custom hooks are refused
    cfg.num_cells_x = cfg.num_cells_y = 13
    cfg.riemann_solver = 1
    assert gpulib.hfv_mh2d_run_noalloc(C.byref(cfg), null, null, 32, 1, None, None) == gpu.HFV_EINVAL
