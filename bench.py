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
import subprocess
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


def clock_warmup(stream, ms):
    """FP64 matrix products (4096^2) on the bench's stream for about `ms`
    milliseconds of GPU time, before the timed steps: the GPU's clocks ramp
    up over the first tens of milliseconds of compute load, and a short
    warmup of steps alone leaves the first timed steps below the sustained
    clock.  The simulation state is not touched.  Returns the GPU time spent
    (ms)."""
    import torch
    if ms <= 0:
        return 0.0
    a = torch.rand(4096, 4096, dtype=torch.float64, device="cuda")
    spent = 0.0
    with torch.cuda.stream(stream):
        while spent < ms:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(4):
                b = a @ a
            e1.record(stream)
            e1.synchronize()
            spent += e0.elapsed_time(e1)
    del a, b
    return spent


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


def host_topology():
    """The host CPU as lscpu reports it (model, sockets, cores per socket,
    threads per core), and the physical cores this process may run on (its
    affinity mask, one logical CPU per (package, core) pair)."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    cpus = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in cpus:
        base = f"/sys/devices/system/cpu/cpu{c}/topology/"
        try:
            cores.add((open(base + "physical_package_id").read().strip(), open(base + "core_id").read().strip()))
        except OSError:
            cores.add(("?", str(c)))
    return {"model": info.get("Model name"), "sockets": info.get("Socket(s)"),
            "cores_per_socket": info.get("Core(s) per socket"), "threads_per_core": info.get("Thread(s) per core"),
            "logical_cpus_allowed": len(cpus), "physical_cores_allowed": len(cores)}


def cpu_child(n: int, times):
    """Child side of cpu_baseline: the reference's OpenMP build (or, when it
    cannot be loaded, the oracle's OpenMP port) on the KH problem, run from
    the initial state to each time in `times`; prints one JSON line of
    (steps, seconds) per run.  Runs in its own process so that the OpenMP
    placement variables reach the runtime before it starts."""
    import ctypes as C
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    from hydrofv_amd import abi, problems
    runs = []
    try:
        ref = abi.bind(C.CDLL(O.REF_OMP_SO))
        kind = "reference"
    except OSError:
        ref, kind = None, "port"
    for tf in times:
        if ref is not None:
            sim = problems.kelvin_helmholtz(ref, n)
            t0 = time.perf_counter()
            res = sim.launch("muscl_hancock_2d", "riemann_solver_hllc", "monotonized_central", 0.8, tf)
            runs.append((res.num_steps, time.perf_counter() - t0))
            sim.close()
        else:
            import hydrofv_amd
            import _cases as K
            sim = problems.kelvin_helmholtz(hydrofv_amd.load(), n)
            grid = K.oracle_grid(sim, "monotonized_central", nthreads=int(os.environ["OMP_NUM_THREADS"]))
            sim.close()
            t0 = time.perf_counter()
            r = grid.run(0.8, tf=tf)
            runs.append((r.num_steps, time.perf_counter() - t0))
    print(json.dumps({"kind": kind, "runs": runs}), flush=True)


def cpu_baseline(n: int, t_start: float, t_end: float, shrink_steps: int, window: int):
    """The reference's OpenMP build (oracle/_ref) on the same KH problem,
    timed over a window of `window` full-CFL steps: one run to t_start (the
    `shrink_steps` reduced-CFL steps), one to t_end (the window beyond), and
    the difference of their wall times.  One OpenMP thread per physical core
    this process may use, capped by OMP_NUM_THREADS when the host sets it;
    OMP_PLACES=cores, OMP_PROC_BIND=close."""
    topo = host_topology()
    threads = topo["physical_cores_allowed"]
    cap = os.environ.get("OMP_NUM_THREADS")
    if cap and cap.isdigit() and int(cap) > 0:
        threads = min(threads, int(cap))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PLACES="cores", OMP_PROC_BIND="close")
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-child", str(n), repr(t_start), repr(t_end)]
    out = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=1800)
    if out.returncode != 0:
        raise RuntimeError(f"cpu baseline child failed ({out.returncode}): {out.stderr[-2000:]}")
    child = json.loads(out.stdout.strip().splitlines()[-1])
    (s0, sec0), (s1, sec1) = child["runs"]
    if s0 != shrink_steps or s1 != shrink_steps + window:
        raise RuntimeError(f"cpu baseline took {s0} and {s1} steps, expected {shrink_steps} and "
                           f"{shrink_steps + window}")
    sec = sec1 - sec0
    return {"value": n * n * window / sec, "unit": "cell-updates/s", "cores": threads, "kind": child["kind"],
            "sample": f"KH {n}x{n}: steps {shrink_steps + 1}-{shrink_steps + window} (the {window} full-CFL steps "
                      f"after the {shrink_steps} shrink steps), {sec:.1f} s wall = run to step "
                      f"{shrink_steps + window} ({sec1:.1f} s) - run to step {shrink_steps} ({sec0:.1f} s)",
            "omp": {"OMP_NUM_THREADS": threads, "OMP_PLACES": "cores", "OMP_PROC_BIND": "close"},
            "cpu": topo}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--strip-rows", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--clock-warmup-ms", type=float, default=200.0,
                    help="GPU time of device copies run after the warmup steps and before the timed ones, "
                         "so that the timed steps run at the sustained clock (0: none)")
    ap.add_argument("--no-graph", action="store_true",
                    help="plain launches instead of captured graphs (the KH and Sedov legs)")
    ap.add_argument("--adjoint-solves", type=int, default=32,
                    help="gradient solves of the 1024^2, 500-step adjoint workload (0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=500,
                    help="steps of the 256-instance batched ensemble and of its sequential baseline, from the "
                         "initial state (the workload's 500 by default; 0: skip)")
    ap.add_argument("--no-extra-configs", action="store_true",
                    help="skip the Sod 1D and Sedov 2D configuration runs")
    ap.add_argument("--cpu-window", type=int, default=100,
                    help="full-CFL steps of the KH CPU baseline's timed window (after the 10 shrink steps)")
    ap.add_argument("--cpu-child", nargs="+", help=argparse.SUPPRESS)
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="after the timed steps, copy a buffer of this many MiB once on the device: a "
                         "kernel of known traffic that scales the profiler's FETCH_SIZE/WRITE_SIZE counters")
    args = ap.parse_args()

    if args.cpu_child:  # the CPU baseline's own process (no GPU)
        cpu_child(int(args.cpu_child[0]), [float(t) for t in args.cpu_child[1:]])
        return

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
    warm_ms = clock_warmup(stream, args.clock_warmup_ms)

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
        # the times after the shrink steps and after the window beyond them,
        # from the device run (bitwise the reference's, so the reference
        # reaches them in exactly those step counts)
        shrink = cfg.cfl_initial_shrink_num_steps
        probe = gpu.MH2D(lib, cfg)
        sim2 = problems.kelvin_helmholtz(lib, n)
        probe.upload(sim2.primitive(), sim2.conserved())
        sim2.close()
        probe.start()
        probe.advance(shrink - 1, use_graph=False)
        t_start = probe.status().t
        probe.advance(args.cpu_window, use_graph=False)
        t_end = probe.status().t
        probe.close()
        try:
            cpu = cpu_baseline(n, t_start, t_end, shrink, args.cpu_window)
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
        "clock_warmup_ms": round(warm_ms, 1),
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
