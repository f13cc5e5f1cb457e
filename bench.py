"""Benchmark: 2D MUSCL-Hancock + HLLC + MC (FP64) on Kelvin-Helmholtz 4096^2.

A step is one MUSCL-Hancock update of the whole grid (CFL time step, both
sweeps, conserved update, CFL reduction for the next step), i.e. one launch
of the fused step kernel.  Inputs are resident in HBM before the timed region;
W warm-up steps (including the first step, which reads the caller's primitive
state) run first, then exactly K steps are timed with HIP events on the
stream the kernels are launched on.

Prints one JSON line.  Extra diagnostics go to stderr.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, AMD public spec
FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 vector, AMD public spec
# wave64 VALU issue: 256 CUs x 4 SIMDs x 2.4 GHz, one FP64 wave64 instruction
# per SIMD every 4 clocks (16 FP64 lanes per SIMD, which is the 78.6 TFLOP/s)
VALU_PEAK_WAVE_INSTR = 256 * 4 * 2.4e9 / 4
ALGO_BYTES_PER_CELL = 64       # 4 FP64 conserved fields read once and written once
PMC_SUMMARY = os.path.join(ROOT, "profiles", "mh2d_step_kh4096_pmc.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic():
    """HBM bytes per step launch from the committed rocprofv3 PMC summary, if any."""
    try:
        with open(PMC_SUMMARY) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def measured_instructions():
    """Per-launch instruction counters of the step kernel from the committed
    rocprofv3 PMC summary: SQ_INSTS_VALU and its FP64 FMA/MUL/ADD parts."""
    out = {}
    try:
        with open(PMC_SUMMARY) as f:
            for line in f:
                cells = [c.strip() for c in line.split("|")]
                if len(cells) > 3 and cells[1] in ("SQ_INSTS_VALU", "SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_MUL_F64",
                                                   "SQ_INSTS_VALU_ADD_F64"):
                    out[cells[1]] = float(cells[2])
    except (OSError, ValueError):
        return None
    return out if len(out) == 4 else None


def compute_bound(ms_per_step):
    """The step kernel against the FP64 VALU limits, which bound it (not
    HBM): FP64 FLOP rate (FMA = 2) against the chip's FP64 vector peak, and
    VALU wave-instruction issue against one FP64 wave64 instruction per SIMD
    every 4 clocks, from the profiled instruction counts per launch."""
    c = measured_instructions()
    if c is None:
        return None
    sec = ms_per_step / 1e3
    flops = 64 * (2 * c["SQ_INSTS_VALU_FMA_F64"] + c["SQ_INSTS_VALU_MUL_F64"] + c["SQ_INSTS_VALU_ADD_F64"])
    tflops = flops / sec / 1e12
    issue = c["SQ_INSTS_VALU"] / sec
    return {"fp64_tflops": tflops, "fp64_peak_tflops": FP64_PEAK_TFLOPS, "fp64_frac": tflops / FP64_PEAK_TFLOPS,
            "valu_wave_instr_per_s": issue, "valu_issue_peak": VALU_PEAK_WAVE_INSTR,
            "valu_issue_frac": issue / VALU_PEAK_WAVE_INSTR, "valu_instr_per_launch": c["SQ_INSTS_VALU"],
            "source": os.path.relpath(PMC_SUMMARY, ROOT)}


def sod_run(lib, n=1000):
    """Benchmark configuration 'Sod 1D, N=1000, HLLC + minmod, CFL 0.9, to
    t=0.2' as one muscl_hancock_1d call (the whole run in one device
    workgroup)."""
    from hydrofv_amd import problems
    out = {"cells": n, "tf": 0.2}
    sim = problems.sod(lib, n)
    sim.launch("muscl_hancock_1d", "riemann_solver_hllc", "minmod", 0.9, 0.2)  # warm-up
    sim.close()
    sim = problems.sod(lib, n)
    t0 = time.perf_counter()
    res = sim.launch("muscl_hancock_1d", "riemann_solver_hllc", "minmod", 0.9, 0.2)
    out["device_ms"] = (time.perf_counter() - t0) * 1e3
    out["steps"] = res.num_steps
    sim.close()
    out["note"] = "wall time of the library call, host-device copies included"
    return out


def copy_bandwidth(stream, mib=1024, reps=5):
    """Achievable HBM bandwidth: a device-to-device copy of `mib` MiB on the
    bench's stream, timed with HIP events; bytes read plus bytes written per
    second, best of `reps`."""
    import torch
    src = torch.ones(mib * (1 << 20) // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    best = float("inf")
    with torch.cuda.stream(stream):
        dst.copy_(src)  # warm-up
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            dst.copy_(src)
            e1.record(stream)
            e1.synchronize()
            best = min(best, e0.elapsed_time(e1))
    del src, dst
    return 2 * mib * (1 << 20) / (best / 1e3) / 1e9


def cpu_baseline_sod(n=1000):
    """The reference's serial build (oracle/_ref) on the same Sod run, beside
    the device's: part of the CPU-baseline leg."""
    import ctypes as C
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    from hydrofv_amd import abi, problems
    if not os.path.exists(O.REF_SO):
        return None
    ref = abi.bind(C.CDLL(O.REF_SO))
    sim = problems.sod(ref, n)
    t0 = time.perf_counter()
    res = sim.launch("muscl_hancock_1d", "riemann_solver_hllc", "minmod", 0.9, 0.2)
    ms = (time.perf_counter() - t0) * 1e3
    sim.close()
    return {"cpu_reference_serial_ms": ms, "cpu_steps": res.num_steps, "cores": 1, "kind": "reference"}


def sedov_run(lib, n=2048, tf=0.05, use_graph=True):
    """Benchmark configuration 'Sedov 2D 2048^2, exact Riemann solver (tol
    1e-8) + Van Leer, CFL 0.5, run to t = 0.05': the whole run, state
    resident on the device, steps queued as captured graphs between status
    checks (use_graph False: plain launches).  Reports the wall time of the
    run (first step to the end) and its mean cell-updates/s."""
    import torch
    from hydrofv_amd import gpu, problems
    sim = problems.sedov(lib, n)
    cfg = problems.device_config(sim, limiter=2, solver=1, cfl=0.5, tf=tf, tol=1e-8)
    dev = gpu.MH2D(lib, cfg)
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    dev.upload(sim.primitive(), sim.conserved(), stream=sh)
    sim.close()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev.start(stream=sh)
    while True:
        dev.advance(960, use_graph=use_graph, stream=sh)
        st = dev.status(stream=sh)
        if st.finished:
            break
    sec = time.perf_counter() - t0
    dev.close()
    assert st.err_kind == 0 and st.t == tf, (st.err_kind, st.t)
    return {"cells": f"{n}x{n}", "solver": "exact (tol 1e-8)", "limiter": "van_leer", "tf": tf,
            "steps": st.num_steps, "run_s": sec, "ms_per_step": sec / st.num_steps * 1e3,
            "cell_updates_per_s": n * n * st.num_steps / sec}


def cpu_baseline(n: int, t_sample: float, steps_sample: int):
    """The reference's OpenMP build (oracle/_ref) on the same KH problem for a
    bounded number of steps (tf = time after `steps_sample` steps); falls back
    to the oracle's OpenMP port when the reference build cannot be loaded."""
    import ctypes as C
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    from hydrofv_amd import abi, problems
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    try:
        ref = abi.bind(C.CDLL(O.REF_OMP_SO))
        kind = "reference"
    except OSError as e:
        log(f"cpu_baseline: reference build not loadable ({e}); timing the oracle port")
        ref = None
        kind = "port"
    if ref is not None:
        sim = problems.kelvin_helmholtz(ref, n)
        t0 = time.perf_counter()
        res = sim.launch("muscl_hancock_2d", "riemann_solver_hllc", "monotonized_central", 0.8, t_sample)
        sec = time.perf_counter() - t0
        steps = res.num_steps
        sim.close()
    else:
        import hydrofv_amd
        sim = problems.kelvin_helmholtz(hydrofv_amd.load(), n)
        import _cases as K
        grid = K.oracle_grid(sim, "monotonized_central", nthreads=threads)
        sim.close()
        t0 = time.perf_counter()
        r = grid.run(0.8, max_steps=steps_sample)
        sec = time.perf_counter() - t0
        steps = r.num_steps
    value = n * n * steps / sec
    return {"value": value, "unit": "cell-updates/s", "cores": threads, "kind": kind,
            "sample": f"KH {n}x{n}, {steps} steps (the 10 CFL-shrink steps), {sec:.1f} s wall, "
                      f"OMP_NUM_THREADS={threads}",
            "cpu": host_cpu()}


def host_cpu():
    """Model name and logical CPU count of the host, from /proc/cpuinfo."""
    try:
        names = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")]
        return f"{names[0]} ({len(names)} logical CPUs)" if names else None
    except OSError:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--strip-rows", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="plain launches instead of captured graphs (the KH and Sedov legs)")
    ap.add_argument("--adjoint-solves", type=int, default=32,
                    help="gradient solves of the 1024^2, 500-step adjoint workload (0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=96,
                    help="steps of the 256-instance batched ensemble and of its sequential baseline (0: skip)")
    ap.add_argument("--no-extra-configs", action="store_true",
                    help="skip the Sod 1D and Sedov 2D configuration runs")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="after the timed steps, copy a buffer of this many MiB once on the device: a "
                         "kernel of known traffic that scales the profiler's FETCH_SIZE/WRITE_SIZE counters")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: this library has no multi-GPU path (--gpus {args.gpus} requested).",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import hydrofv_amd
    from hydrofv_amd import gpu, problems

    assert torch.cuda.is_available(), "no GPU visible"
    torch.cuda.init()
    lib = hydrofv_amd.load()
    n = args.n
    t_setup = time.perf_counter()
    sim = problems.kelvin_helmholtz(lib, n)
    cfg = problems.device_config(sim, limiter=3, solver=2, cfl=0.8, strip_rows=args.strip_rows)
    dev = gpu.MH2D(lib, cfg)
    W = sim.primitive()
    U = sim.conserved()
    sim.close()
    # A stream of our own: the library runs work handed no stream (handle 0)
    # on a private non-blocking stream, which events on torch's default
    # stream would not bracket.
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    assert sh, "expected a non-null stream handle"
    dev.upload(W, U, stream=sh)
    del W, U
    dev.start(stream=sh)                      # step 1 (from the primitive state)
    dev.advance(max(args.warmup - 1, 0), use_graph=not args.no_graph, stream=sh)
    st0 = dev.status(stream=sh)
    log(f"setup+warmup {time.perf_counter() - t_setup:.1f} s; steps={st0.num_steps} t={st0.t:.6g}")
    assert st0.err_kind == 0, gpu.ERR_NAMES.get(st0.err_kind)

    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    dev.advance(args.steps, use_graph=not args.no_graph, stream=sh)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    st = dev.status(stream=sh)
    assert st.err_kind == 0 and st.num_steps == st0.num_steps + args.steps, (st.err_kind, st.num_steps)
    ms_per_step = ms / args.steps
    cells = n * n
    value = cells * args.steps / (ms / 1e3)
    achieved = ALGO_BYTES_PER_CELL * cells / (ms_per_step / 1e3) / 1e9
    log(f"{ms_per_step:.4f} ms/step, {value / 1e9:.2f} G cell-updates/s, {achieved:.0f} GB/s algorithmic")
    dev.close()

    if args.calibrate_copy_mib:
        src = torch.ones(args.calibrate_copy_mib * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calibrate_copy_mib} MiB read + {args.calibrate_copy_mib} MiB written")
        del src, dst

    copy_gbs = None
    if not args.calibrate_copy_mib:  # a counter run keeps its one known copy
        try:
            copy_gbs = copy_bandwidth(stream)
            log(f"device copy: {copy_gbs:.0f} GB/s (read + write)")
        except Exception as e:  # the main figure stands without it
            log(f"copy bandwidth failed: {e!r}")

    cpu = None
    if not args.no_cpu_baseline:
        sample_steps = 10
        probe = gpu.MH2D(lib, cfg)
        sim2 = problems.kelvin_helmholtz(lib, n)
        probe.upload(sim2.primitive(), sim2.conserved())
        sim2.close()
        probe.start()
        probe.advance(sample_steps - 1, use_graph=False)
        t_sample = probe.status().t
        probe.close()
        try:
            cpu = cpu_baseline(n, t_sample, sample_steps)
            log(f"cpu baseline: {cpu}")
        except Exception as e:  # the GPU figure stands without it
            log(f"cpu baseline failed: {e!r}")

    ens = None
    if args.ensemble_steps:
        from hydrofv_amd import ensemble
        try:
            ens = ensemble.benchmark(lib, count=256, steps=args.ensemble_steps, log=log)
            log(f"ensemble: {ens}")
        except Exception as e:  # the main figure stands without it
            log(f"ensemble benchmark failed: {e!r}")

    adj = None
    if args.adjoint_solves:
        from hydrofv_amd import adjoint
        try:
            adj = adjoint.benchmark(n=1024, steps=500, solves=args.adjoint_solves, log=log)
            log(f"adjoint: {adj}")
        except Exception as e:  # the main figure stands without it
            log(f"adjoint benchmark failed: {e!r}")

    extra = {}
    if not args.no_extra_configs:
        legs = (("sod_1d", lambda: sod_run(lib)), ("sedov_2d", lambda: sedov_run(lib, use_graph=not args.no_graph)))
        for name, fn in legs:
            try:
                extra[name] = fn()
                log(f"{name}: {extra[name]}")
            except Exception as e:  # the main figure stands without them
                log(f"{name} failed: {e!r}")
        if "sod_1d" in extra and not args.no_cpu_baseline:
            try:
                extra["sod_1d"]["cpu_baseline"] = cpu_baseline_sod()
            except Exception as e:
                log(f"sod cpu baseline failed: {e!r}")

    traffic = measured_traffic()
    line = {
        "metric": "cell-updates/s FP64 MUSCL-HLLC 4096²; % HBM roofline; batched & adjoint ratio",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "kelvin_helmholtz_2d", "cells": f"{n}x{n}", "ghost_cells": 2, "gamma": "5/3",
                   "boundary": "periodic", "limiter": "monotonized_central", "riemann_solver": "hllc",
                   "cfl": 0.8, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "achievable_copy_gbs": copy_gbs},
        "compute_bound": compute_bound(ms_per_step),
        "cpu_baseline": cpu,
        "ensemble": ens,
        "adjoint": adj,
        **extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
