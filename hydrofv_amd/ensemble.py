"""Batched ensemble of independent 2D problems (hfv_ens_*) and its autotuner.

The ensemble workload: `count` four-quadrant Riemann problems on square grids
whose sides are drawn log-uniform in [64, 1024]; each instance has its own
quadrant states, its own CFL time step and the same limiter/solver.  All
instances advance in one launch per step (one hipGraph replay per 48 steps).
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import gpu, problems, workloads
from .gpu import MH2DConfig, RunStatus, _ptr_array, check

_P = C.POINTER
PROTOTYPES = {
    "hfv_ens_create": (C.c_int, [C.c_int, _P(MH2DConfig), _P(C.c_void_p)]),
    "hfv_ens_destroy": (None, [C.c_void_p]),
    "hfv_ens_set_strip_rows": (C.c_int, [C.c_void_p, _P(C.c_int)]),
    "hfv_ens_set_hllc_form": (C.c_int, [C.c_void_p, C.c_int]),
    "hfv_ens_num_tiles": (C.c_int, [C.c_void_p]),
    "hfv_ens_upload": (C.c_int, [C.c_void_p, C.c_int, _P(gpu.dp), _P(gpu.dp), C.c_void_p]),
    "hfv_ens_start": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hfv_ens_advance": (C.c_int, [C.c_void_p, C.c_longlong, C.c_int, C.c_void_p]),
    "hfv_ens_status": (C.c_int, [C.c_void_p, _P(RunStatus), C.c_void_p]),
    "hfv_ens_download": (C.c_int, [C.c_void_p, C.c_int, _P(gpu.dp), _P(gpu.dp), C.c_void_p]),
    "hfv_ens_checkpoint": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
}


def bind(lib):
    for name, (restype, argtypes) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    return lib


def build_instances(lib, count=256, seed=2, side_lo=64, side_hi=1024, steps=500, cfl=0.8,
                    limiter=1, solver=2):
    """Initial states (host arrays) and device configs of the ensemble."""
    sides, states = workloads.ensemble_specs(count, seed, side_lo, side_hi)
    cfgs, init = [], []
    for i in range(count):
        sim = problems.four_quadrant(lib, int(sides[i]), states[i])
        cfgs.append(problems.device_config(sim, limiter=limiter, solver=solver, cfl=cfl, max_steps=steps))
        init.append((sim.primitive(), sim.conserved()))
        sim.close()
    return sides, cfgs, init


class Ensemble:
    def __init__(self, lib, cfgs, init=None, stream=None):
        self.lib = bind(lib)
        self.count = len(cfgs)
        arr = (MH2DConfig * self.count)(*cfgs)
        h = C.c_void_p()
        check(lib, lib.hfv_ens_create(self.count, arr, C.byref(h)), "hfv_ens_create")
        self.h = h
        self.cfgs = list(cfgs)
        self.init = init
        if init is not None:
            self.upload(init, stream)

    def upload(self, init, stream=None):
        for i, (W, U) in enumerate(init):
            W = [np.ascontiguousarray(w) for w in W]
            U = [np.ascontiguousarray(u) for u in U]
            check(self.lib, self.lib.hfv_ens_upload(self.h, i, _ptr_array(W), _ptr_array(U), stream), "upload")
            # keep host arrays alive until the copies have run
            if stream is not None and i % 16 == 15:
                self.status(stream)
        self.status(stream)

    def set_strip_rows(self, rows):
        a = (C.c_int * self.count)(*[int(r) for r in rows])
        check(self.lib, self.lib.hfv_ens_set_strip_rows(self.h, a), "set_strip_rows")

    def set_hllc_form(self, form):
        check(self.lib, self.lib.hfv_ens_set_hllc_form(self.h, int(form)), "set_hllc_form")

    @property
    def num_tiles(self):
        return self.lib.hfv_ens_num_tiles(self.h)

    def start(self, stream=None):
        check(self.lib, self.lib.hfv_ens_start(self.h, stream), "start")

    def advance(self, launches, use_graph=True, stream=None):
        check(self.lib, self.lib.hfv_ens_advance(self.h, launches, int(use_graph), stream), "advance")

    def status(self, stream=None):
        st = (RunStatus * self.count)()
        check(self.lib, self.lib.hfv_ens_status(self.h, st, stream), "status")
        return st

    def checkpoint(self, stream=None):
        """Keep a device copy of the whole state (fields and time control)."""
        check(self.lib, self.lib.hfv_ens_checkpoint(self.h, 0, stream), "checkpoint")

    def restore(self, stream=None):
        """Put the state kept by checkpoint() back."""
        check(self.lib, self.lib.hfv_ens_checkpoint(self.h, 1, stream), "restore")

    def download(self, i, stream=None):
        c = self.cfgs[i]
        shape = (c.num_cells_y + 2 * c.num_ghost_cells_side, c.num_cells_x + 2 * c.num_ghost_cells_side)
        W = [np.empty(shape) for _ in range(4)]
        U = [np.empty(shape) for _ in range(4)]
        check(self.lib, self.lib.hfv_ens_download(self.h, i, _ptr_array(W), _ptr_array(U), stream), "download")
        return W, U

    def close(self):
        if self.h:
            self.lib.hfv_ens_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def strip_rule(sides, kind, value):
    """Candidate tile shapes: a fixed strip height, or a height proportional
    to the instance's side (so small grids get short strips and more waves)."""
    sides = np.asarray(sides)
    if kind == "fixed":
        return np.full(len(sides), value)
    return np.clip(np.rint(sides / value), 8, 512).astype(int)


# the strip-height rules the autotuner times (strip_rule)
STRIP_RULES = [("fixed", v) for v in (16, 24, 32, 48, 64)] + [("prop", d) for d in (4, 8, 16, 32, 48, 64)]


def autotune(ens: Ensemble, sides, stream, trial_launches=48, log=None):
    """Pick the strip heights and the HLLC code form (hfv_mh2d_config.hllc_form)
    that minimise the ensemble step time.  Every candidate starts from the
    same state (the device copy checkpoint() keeps of the state the call
    finds), runs its first step and one untimed graph replay of
    `trial_launches` steps (which captures the graph), is put back to that
    state, and is then timed with HIP events on `stream` (a torch stream)
    over the first step and one replay of the same graph.  The state is
    restored afterwards, and `stream` synchronised.  Returns ((form, kind, value, seconds per step,
    rows), all results)."""
    import torch

    sh = stream.cuda_stream
    rules = STRIP_RULES
    forms = (gpu.HLLC_SELECTS, gpu.HLLC_BRANCHES) if ens.cfgs[0].riemann_solver == 2 else (gpu.HLLC_SELECTS,)
    best = None
    results = []
    ens.checkpoint(sh)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for kind, value in rules:
        rows = strip_rule(sides, kind, value)
        ens.set_strip_rows(rows)
        for form in forms:
            ens.set_hllc_form(form)
            ens.restore(sh)
            ens.start(sh)
            ens.advance(trial_launches, use_graph=True, stream=sh)
            ens.restore(sh)
            e0.record(stream)
            ens.start(sh)
            ens.advance(trial_launches, use_graph=True, stream=sh)
            e1.record(stream)
            e1.synchronize()
            dt = e0.elapsed_time(e1) / 1e3 / (trial_launches + 1)
            results.append((form, kind, value, dt, ens.num_tiles))
            if log:
                log(f"autotune form={FORM_NAMES[form]} {kind}={value}: {dt * 1e3:.3f} ms/step, {ens.num_tiles} tiles")
            if best is None or dt < best[3]:
                best = (form, kind, value, dt, rows)
    ens.set_strip_rows(best[4])
    ens.set_hllc_form(best[0])
    ens.restore(sh)
    stream.synchronize()  # the caller may go on on another stream
    return best, results


FORM_NAMES = {gpu.HLLC_SELECTS: "selects", gpu.HLLC_BRANCHES: "branches"}


def benchmark(lib, count=256, seed=2, steps=500, log=None):
    """Aggregate cell-updates/s of the batched engine (autotuned strips and
    HLLC code form, one launch per step, hipGraph replays of 48 steps)
    against the same instances run one after another on the single-grid
    engine (the same code form, its own graph replays and default strips).
    Both are timed with HIP events over `steps` steps from the initial state
    (the first step, which reads the primitive state, included), data
    resident on the device, graphs captured beforehand."""
    import torch

    sides, cfgs, init = build_instances(lib, count=count, seed=seed, steps=-1)
    cells = int(np.sum(sides.astype(np.int64) ** 2))
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    ens = Ensemble(lib, cfgs, init, stream=sh)
    best, _ = autotune(ens, sides, stream, log=log)
    # capture the replays the timed run takes, then start over from the
    # initial state
    ens.checkpoint(sh)
    ens.start(sh)
    ens.advance(steps - 1, use_graph=True, stream=sh)
    ens.restore(sh)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    ens.start(sh)
    ens.advance(steps - 1, use_graph=True, stream=sh)
    e1.record(stream)
    torch.cuda.synchronize()
    t_batch = e0.elapsed_time(e1) / 1e3
    st = ens.status(sh)
    assert all(s.err_kind == 0 and s.num_steps == steps for s in st), "ensemble run failed"
    ens.close()

    singles = []
    for c, (W, U) in zip(cfgs, init):
        c.hllc_form = best[0]  # the sequential runs use the same code form
        m = gpu.MH2D(lib, c)
        m.upload(W, U, stream=sh)
        m.start(stream=sh)
        # graph capture for the controller phase the timed run replays at
        m.advance(48, use_graph=True, stream=sh)
        m.upload(W, U, stream=sh)
        singles.append(m)
    torch.cuda.synchronize()
    e0.record(stream)
    for m in singles:
        m.start(stream=sh)
        m.advance(steps - 1, use_graph=True, stream=sh)
    e1.record(stream)
    torch.cuda.synchronize()
    t_seq = e0.elapsed_time(e1) / 1e3
    for m in singles:
        st = m.status(stream=sh)
        assert st.err_kind == 0 and st.num_steps == steps, "sequential run failed"
        m.close()
    batch = cells * steps / t_batch
    seq = cells * steps / t_seq
    return {"instances": count, "cells": cells, "steps": steps, "strip_rule": f"{best[1]}={best[2]}",
            "hllc_form": FORM_NAMES[best[0]],
            "batched_cell_updates_per_s": batch, "sequential_cell_updates_per_s": seq,
            "speedup": batch / seq, "batched_ms_per_step": t_batch / steps * 1e3,
            "sequential_ms_per_step": t_seq / steps * 1e3}
