"""Differentiable fixed-dt 2D MUSCL-Hancock + HLLC rollouts on the device.

`Periodic2D` describes a doubly periodic Cartesian grid and binds the C ABI
entry points hfv_fwd2d_step (one step, V = step(U)) and hfv_adj2d_step (the
transposed Jacobian of that step applied to a cotangent).  `rollout(U0, p, n)`
is a torch.autograd.Function: forward runs n steps keeping every k-th state
(k ~ sqrt(n)); backward walks the segments in reverse, re-running each
segment's forward from its checkpoint and then the adjoint steps through it,
so memory holds about 2 sqrt(n) states instead of n.

States are float64 CUDA tensors of shape (4, ny + 2g, nx + 2g) holding
(m, mx, my, E) of every cell (extensive, as the reference's U).  Only the
interior is state: ghost entries are not read by a step and receive no
gradient.  Errors: a failed forward flux raises; faces where the adjoint met
the HLLC vacuum branch (derivative taken as zero) are counted in
`Periodic2D.last_flags`.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import torch

from . import load
from .gpu import check

FLUX_ERROR, VACUUM, INVALID = 1, 2, 4


class Adj2DConfig(C.Structure):
    """hfv_adj2d_config (include/hydrofv/hfv_gpu.h)"""
    _fields_ = [("num_cells_x", C.c_int), ("num_cells_y", C.c_int), ("num_ghost_cells_side", C.c_int),
                ("pitch", C.c_longlong), ("gamma", C.c_double), ("dt", C.c_double), ("tol", C.c_double),
                ("area_x", C.c_double), ("area_y", C.c_double), ("volume", C.c_double),
                ("slope_limiter", C.c_int)]


_P4 = C.c_void_p * 4
PROTOTYPES = {
    "hfv_fwd2d_step": (C.c_int, [C.POINTER(Adj2DConfig), _P4, _P4, C.c_void_p, C.c_void_p]),
    "hfv_adj2d_step": (C.c_int, [C.POINTER(Adj2DConfig), _P4, _P4, _P4, C.c_void_p, C.c_void_p]),
}


def bind(lib):
    for name, (restype, argtypes) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    return lib


def _fields(t: torch.Tensor):
    assert t.dtype == torch.float64 and t.is_cuda and t.is_contiguous() and t.dim() == 3 and t.shape[0] == 4
    stride = t.stride(0) * t.element_size()
    base = t.data_ptr()
    return _P4(*[base + f * stride for f in range(4)])


@dataclass
class Periodic2D:
    """A doubly periodic uniform Cartesian grid of nx x ny cells of size
    dx x dy with g ghost cells a side, and the fixed time step dt."""
    nx: int
    ny: int
    dx: float
    dy: float
    dt: float
    gamma: float = 1.4
    g: int = 2
    limiter: int = 1          # slope_limiter.h: minmod 1, Van Leer 2, MC 3
    tol: float = 1e-6

    def __post_init__(self):
        self.lib = bind(load())
        self.tx, self.ty = self.nx + 2 * self.g, self.ny + 2 * self.g
        self.cfg = Adj2DConfig(self.nx, self.ny, self.g, self.tx, self.gamma, self.dt, self.tol,
                               self.dy, self.dx, self.dx * self.dy, self.limiter)
        self.last_flags = 0
        self._flags = None

    @property
    def volume(self):
        return self.dx * self.dy

    def shape(self):
        return (4, self.ty, self.tx)

    def interior(self, t):
        g = self.g
        return t[..., g:g + self.ny, g:g + self.nx]

    def flags(self, device):
        if self._flags is None or self._flags.device != device:
            self._flags = torch.zeros(1, dtype=torch.int32, device=device)
        return self._flags

    def _stream(self, t):
        return torch.cuda.current_stream(t.device).cuda_stream

    def step(self, U: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """V = step(U) (no autograd)."""
        V = torch.empty_like(U) if out is None else out
        fl = self.flags(U.device)
        check(self.lib, self.lib.hfv_fwd2d_step(C.byref(self.cfg), _fields(U), _fields(V),
                                                C.c_void_p(fl.data_ptr()), C.c_void_p(self._stream(U))),
              "hfv_fwd2d_step")
        return V

    def adjoint_step(self, U: torch.Tensor, lam_in: torch.Tensor, lam_out: torch.Tensor) -> torch.Tensor:
        """lam_out = (d step(U)/dU)^T lam_in on the interior (ghosts of lam_out untouched)."""
        fl = self.flags(U.device)
        check(self.lib, self.lib.hfv_adj2d_step(C.byref(self.cfg), _fields(U), _fields(lam_in), _fields(lam_out),
                                                C.c_void_p(fl.data_ptr()), C.c_void_p(self._stream(U))),
              "hfv_adj2d_step")
        return lam_out

    def take_flags(self, device) -> int:
        """Read and clear the device flags (synchronises)."""
        fl = self.flags(device)
        v = int(fl.item())
        fl.zero_()
        self.last_flags |= v
        return v

    def run(self, U0: torch.Tensor, steps: int) -> torch.Tensor:
        """U after `steps` steps (no autograd).  Ghost entries of the result
        are defined (both buffers start as U0) but not part of the state."""
        a, b = U0.clone(), U0.clone()
        for _ in range(steps):
            self.step(a, b)
            a, b = b, a
        if self.take_flags(U0.device) & FLUX_ERROR:
            raise RuntimeError("HLLC flux failure in the forward rollout (invalid wave speeds)")
        return a


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U0, grid: Periodic2D, steps: int, every: int):
        ckpts = [U0.detach().clone()]
        # both buffers start as U0, so the result's ghost entries are defined;
        # they are not part of the state and carry no gradient
        a, b = U0.detach().clone(), U0.detach().clone()
        for n in range(steps):
            grid.step(a, b)
            a, b = b, a
            if (n + 1) % every == 0 and n + 1 < steps:
                ckpts.append(a.clone())
        if grid.take_flags(U0.device) & FLUX_ERROR:
            raise RuntimeError("HLLC flux failure in the forward rollout (invalid wave speeds)")
        ctx.grid, ctx.steps, ctx.every = grid, steps, every
        ctx.ckpts = ckpts
        return a

    @staticmethod
    def backward(ctx, grad_out):
        grid, steps, every = ctx.grid, ctx.steps, ctx.every
        if steps == 0:  # the identity
            ctx.ckpts = None
            return grad_out.clone(), None, None, None
        lam = torch.zeros(grid.shape(), dtype=torch.float64, device=grad_out.device)
        grid.interior(lam).copy_(grid.interior(grad_out))
        other = torch.zeros_like(lam)
        nseg = len(ctx.ckpts)
        for k in reversed(range(nseg)):
            start = k * every
            stop = min(start + every, steps)
            # re-run the segment: states[j] = U^{start + j}
            states = [ctx.ckpts[k]]
            for _ in range(start, stop - 1):
                states.append(grid.step(states[-1]))
            for U in reversed(states):
                grid.adjoint_step(U, lam, other)
                lam, other = other, lam
        ctx.ckpts = None
        if grid.take_flags(grad_out.device) & FLUX_ERROR:
            raise RuntimeError("HLLC flux failure while re-running the forward")
        return lam, None, None, None


def rollout(U0: torch.Tensor, grid: Periodic2D, steps: int, every: int | None = None) -> torch.Tensor:
    """U after `steps` fixed-dt steps, differentiable w.r.t. U0 through the
    discrete adjoint with checkpoints every `every` steps (default ~sqrt)."""
    if every is None:
        every = max(1, int(math.ceil(math.sqrt(steps))))
    return _Rollout.apply(U0, grid, steps, every)


def density_loss(U: torch.Tensor, rho_target: torch.Tensor, grid: Periodic2D) -> torch.Tensor:
    """1/2 sum over the interior of (rho - rho_target)^2, rho = m / V."""
    rho = grid.interior(U)[0] / grid.volume
    return 0.5 * ((rho - rho_target) ** 2).sum()


def conserved_from_primitive(W, grid: Periodic2D, device="cuda") -> torch.Tensor:
    """(rho, u, v, p) interior arrays -> U with zero ghost cells (m = rho V,
    momenta m u and m v, E = V (p/(gamma-1) + rho |v|^2 / 2), system.c:804-843)."""
    import numpy as np
    rho, u, v, p = (np.asarray(x, dtype=np.float64) for x in W)
    vol = grid.volume
    m = rho * vol
    U = torch.zeros(grid.shape(), dtype=torch.float64)
    inner = grid.interior(U)
    inner[0] = torch.from_numpy(m)
    inner[1] = torch.from_numpy(m * u)
    inner[2] = torch.from_numpy(m * v)
    inner[3] = torch.from_numpy(vol * (p / (grid.gamma - 1.0) + 0.5 * rho * (u * u + v * v)))
    return U.to(device)


def cfl_time_step(W, dx, dy, gamma, cfl=0.8):
    """The reference's CFL time step of a primitive state (utils.c:75-92)."""
    import numpy as np
    rho, u, v, p = W
    a = np.sqrt(gamma * p / rho)
    return cfl / ((a + np.abs(u)).max() / dx + (a + np.abs(v)).max() / dy)


def workload_grid(W, n, gamma=1.4, g=2, cfl=0.8, limiter=1) -> Periodic2D:
    """The unit square as the reference lays it out (dx = 1 / (n + 2g)) with
    the CFL-0.8 time step of W kept fixed."""
    dx = 1.0 / (n + 2 * g)
    return Periodic2D(n, n, dx, dx, cfl_time_step(W, dx, dx, gamma, cfl), gamma=gamma, g=g, limiter=limiter)


def benchmark(n=1024, steps=500, solves=32, sigma=16.0, log=None):
    """Adjoint workload: minmod + HLLC, periodic n x n, smooth random initial
    states, `steps` fixed-dt steps, density-mismatch loss against a target
    run from a 1e-2 density perturbation; the gradient w.r.t. all initial
    conserved variables by the checkpointed adjoint, for `solves` seeds.
    Reports the kernel-level adjoint/forward step cost and the time per
    gradient solve against a plain forward run."""
    from . import workloads

    dev = torch.device("cuda")
    W = workloads.adjoint_initial(n, seed=0, sigma=sigma)
    grid = workload_grid(W, n)
    U0 = conserved_from_primitive(W, grid)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 100
    a, b = U0.clone(), torch.empty_like(U0)
    for _ in range(10):
        grid.step(a, b)
        a, b = b, a
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        grid.step(a, b)
        a, b = b, a
    e1.record()
    torch.cuda.synchronize()
    fwd_ms = e0.elapsed_time(e1) / reps
    lam = torch.zeros_like(U0)
    grid.interior(lam).normal_()
    other = torch.zeros_like(U0)
    for _ in range(5):
        grid.adjoint_step(U0, lam, other)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        grid.adjoint_step(U0, lam, other)
        lam, other = other, lam
    e1.record()
    torch.cuda.synchronize()
    adj_ms = e0.elapsed_time(e1) / reps
    grid.take_flags(dev)
    if log:
        log(f"adjoint {n}^2: forward step {fwd_ms:.4f} ms, adjoint step {adj_ms:.4f} ms")

    t_grad = t_fwd = 0.0
    losses = []
    flags = 0
    import time
    for s in range(solves):
        W = workloads.adjoint_initial(n, seed=s, sigma=sigma)
        g = workload_grid(W, n)
        Wt = W.copy()
        Wt[0] = Wt[0] + workloads.adjoint_perturbation(n, seed=s, sigma=sigma)
        rho_t = g.interior(g.run(conserved_from_primitive(Wt, g), steps))[0] / g.volume
        U = conserved_from_primitive(W, g)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.run(U, steps)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        U.requires_grad_(True)
        loss = density_loss(rollout(U, g, steps), rho_t, g)
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        t_fwd += t1 - t0
        t_grad += t2 - t1
        losses.append(loss.item())
        flags |= g.last_flags
        assert torch.isfinite(U.grad).all()
    return {"cells": f"{n}x{n}", "steps": steps, "limiter": "minmod", "riemann_solver": "hllc",
            "forward_ms_per_step": fwd_ms, "adjoint_ms_per_step": adj_ms, "adjoint_to_forward": adj_ms / fwd_ms,
            "solves": solves, "gradient_s_per_solve": t_grad / solves, "forward_run_s": t_fwd / solves,
            "gradient_to_forward_run": t_grad / t_fwd, "vacuum_faces_flagged": bool(flags & VACUUM),
            "checkpoint_every": int(math.ceil(math.sqrt(steps))), "mean_loss": sum(losses) / len(losses)}
