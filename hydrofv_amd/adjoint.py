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
        """U after `steps` steps (no autograd)."""
        a, b = U0.clone(), torch.empty_like(U0)
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
        a, b = U0.detach().clone(), torch.empty_like(U0)
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
