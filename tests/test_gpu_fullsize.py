"""Parity at the full sizes of the measured workloads, on the kernels and
launch shapes the benchmark times (default strip heights, the autotuned
ensemble tiles):

- Kelvin-Helmholtz 4096^2 (MC + HLLC, CFL 0.8): 12 steps on the device,
  crossing the end of the CFL shrink, against the reference's own OpenMP
  build run to the device's t_12 (it then takes the same 12 steps);
- the 256-instance four-quadrant ensemble with autotuned tiles: 10 steps,
  the three largest, three smallest and four seeded random instances bit
  for bit equal to single-grid runs, and the smallest against the oracle;
- Sedov 2048^2 (exact solver, tol 1e-8, Van Leer, CFL 0.5): 10 steps
  against the oracle on the host's cores.

Each comparison reports, per conserved field, the L1 relative difference,
the largest difference relative to the field's largest value and the
largest elementwise relative difference (tests/_cases.field_errors).  The
shear-layer and ensemble runs hold all three to 1e-12 (they are bit for bit
equal as run); the blast holds the first two to 1e-10."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest

import _cases as K
import _oracle as O
from hydrofv_amd import abi, ensemble, gpu, problems

pytestmark = pytest.mark.gpu

CHAOTIC_TOL = 1e-12   # relative, per field, over the short horizon of a shear-layer run
SHOCK_TOL = 1e-10     # relative, per field, blast and shock problems (exact solver powers)


def host_threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def device_run(lib, sim, steps, **kw):
    cfg = problems.device_config(sim, max_steps=steps, **kw)
    m = gpu.MH2D(lib, cfg)
    m.upload(sim.primitive(), sim.conserved())
    m.start()
    m.advance(steps + 2, use_graph=True)
    st = m.status()
    W, U = m.download()
    m.close()
    return st, np.stack(W), np.stack(U)


def check_fields(label, W, U, Wr, Ur, tol, elementwise=True):
    """L1 and max relative differences per field (the tolerance's metrics)
    within tol; with `elementwise`, every entry's relative difference too."""
    eu = K.field_errors(U, Ur)
    ew = K.field_errors(W, Wr)
    bitwise = np.array_equal(U.view(np.uint64), Ur.view(np.uint64)) and \
        np.array_equal(W.view(np.uint64), Wr.view(np.uint64))
    print(f"{label}: bitwise {bitwise}; U {K.format_field_errors(eu)}; "
          f"W {K.format_field_errors(ew, ('rho', 'u', 'v', 'p'))}", flush=True)
    for e in eu + ew:
        assert e[0] <= tol and e[1] <= tol, (label, e)
        if elementwise:
            assert e[2] <= tol, (label, e)
    return bitwise


@pytest.fixture(scope="module")
def ref_omp():
    if not O.ref_available() or not os.path.exists(O.REF_OMP_SO):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    return abi.bind(C.CDLL(O.REF_OMP_SO))


def test_kh_4096_against_the_reference_build(gpulib, ref_omp):
    n, steps = 4096, 12
    print(f"KH {n}^2: device, {steps} steps", flush=True)
    sim = problems.kelvin_helmholtz(gpulib, n)
    st, W, U = device_run(gpulib, sim, steps, limiter=3, solver=2, cfl=0.8)
    sim.close()
    assert st.err_kind == 0 and st.num_steps == steps
    print(f"KH {n}^2: reference OpenMP build ({host_threads()} host threads) to t={st.t!r}", flush=True)
    ref = problems.kelvin_helmholtz(ref_omp, n)
    res = ref.launch("muscl_hancock_2d", "riemann_solver_hllc", "monotonized_central", 0.8, st.t)
    assert res.num_steps == steps and res.t == st.t and res.dt == st.dt
    Wr, Ur = K.state(ref)
    ref.close()
    check_fields(f"KH {n}^2 x {steps}", W, U, Wr, Ur, CHAOTIC_TOL)


def test_ensemble_256_autotuned_equals_single_runs(gpulib):
    steps = 10
    print("ensemble: 256 instances, autotune", flush=True)
    # no step limit, as in the benchmark (the autotuner times live steps);
    # exactly `steps` launches below
    sides, cfgs, init = ensemble.build_instances(gpulib, count=256, seed=2, steps=-1)
    import torch
    ens = ensemble.Ensemble(gpulib, cfgs, init)
    best, _ = ensemble.autotune(ens, sides, torch.cuda.Stream(), trial_launches=2)
    print(f"ensemble: tiles {best[1]}={best[2]}, form {ensemble.FORM_NAMES[best[0]]}, {ens.num_tiles} tiles",
          flush=True)
    ens.start()
    ens.advance(steps - 1, use_graph=True)
    st = ens.status()
    assert all(s.err_kind == 0 and s.num_steps == steps for s in st)
    order = np.argsort(sides, kind="stable")
    rng = np.random.default_rng(5)
    rest = np.setdiff1d(np.arange(len(sides)), np.r_[order[:3], order[-3:]])
    picks = list(order[-3:]) + list(order[:3]) + list(rng.choice(rest, 4, replace=False))
    for i in picks:
        c = cfgs[i]
        c.hllc_form = best[0]
        m = gpu.MH2D(gpulib, c)
        m.upload(*init[i])
        m.start()
        m.advance(steps - 1, use_graph=False)
        s1 = m.status()
        W1, U1 = (np.stack(a) for a in m.download())
        m.close()
        W, U = (np.stack(a) for a in ens.download(i))
        assert (st[i].t, st[i].dt, st[i].num_steps) == (s1.t, s1.dt, s1.num_steps), i
        assert np.array_equal(U.view(np.uint64), U1.view(np.uint64)), (i, int(sides[i]))
        assert np.array_equal(W.view(np.uint64), W1.view(np.uint64)), (i, int(sides[i]))
    # the smallest instance against the oracle
    i = int(order[0])
    from test_gpu_ensemble import _sim_like
    grid = K.oracle_grid(_sim_like(gpulib, int(sides[i]), init[i]), "minmod")
    r = grid.run(0.8, max_steps=steps)
    assert r.error == 0 and r.t == st[i].t
    W, U = (np.stack(a) for a in ens.download(i))
    check_fields(f"ensemble instance {i} (side {int(sides[i])})", W, U, grid.W, grid.U, CHAOTIC_TOL)
    ens.close()


def test_sedov_2048_against_the_oracle(gpulib):
    n, steps = 2048, 10
    print(f"Sedov {n}^2: device, {steps} steps", flush=True)
    sim = problems.sedov(gpulib, n)
    grid = K.oracle_grid(sim, "van_leer", solver="riemann_solver_exact", tol=1e-8, nthreads=host_threads())
    st, W, U = device_run(gpulib, sim, steps, limiter=2, solver=1, cfl=0.5, tol=1e-8)
    sim.close()
    assert st.err_kind == 0 and st.num_steps == steps
    print(f"Sedov {n}^2: oracle on {host_threads()} host threads", flush=True)
    r = grid.run(0.5, max_steps=steps)
    assert r.error == 0 and r.num_steps == steps
    assert abs(r.t - st.t) <= SHOCK_TOL * r.t
    # The exact solver's powers (pow_pos) differ from glibc's pow in the last
    # place on about 1 argument in 1000, and in a blast such a difference
    # reaches cells whose momentum is a small remainder of large cancelling
    # fluxes: there the elementwise relative difference is large (up to
    # ~1e-1 on entries a few percent of the field's maximum, measured), while
    # the L1 and max relative differences per field stay near 1e-11.  The
    # elementwise figures are printed, the tolerance holds on the field
    # metrics.
    check_fields(f"Sedov {n}^2 x {steps}", W, U, grid.W, grid.U, SHOCK_TOL, elementwise=False)
