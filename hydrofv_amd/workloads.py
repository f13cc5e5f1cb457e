"""Initial conditions of the benchmark workloads, as primitive-variable arrays.

Each generator returns arrays in the reference layout (row-major, x fastest,
ghost cells included) for the cells the corresponding example sets:
interior cells for the 2D problems (then the caller fills the ghosts with the
boundary condition and converts to conserved variables, as in
examples/kelvin_helmholtz_2d/kelvin_helmholtz_2d.c:118-128), every cell for the
shock tube (examples/sod_shock_1d/sod_shock_1d.c:47-62).

Random draws come from splitmix64 so that every consumer of a workload sees
the same numbers.
"""
from __future__ import annotations

import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, start: int = 0) -> np.ndarray:
    """Outputs start..start+n-1 of the splitmix64 stream seeded with ``seed``."""
    with np.errstate(over="ignore"):
        k = np.arange(start + 1, start + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, n: int, start: int = 0) -> np.ndarray:
    """Doubles in [0, 1): the top 53 bits of each splitmix64 output."""
    return (splitmix64(seed, n, start) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def cell_centres(lo: float, hi: float, n: int, g: int) -> np.ndarray:
    """Cell centres with the ghost cells inside the domain (src/system.c:302, 345)."""
    h = (hi - lo) / (n + 2 * g)
    return lo + (np.arange(n + 2 * g) + 0.5) * h


# ---------------------------------------------------------------------------
def kelvin_helmholtz(n: int, g: int = 2, seed: int = 1, noise: float = 1e-3):
    """KH shear layer on [0,1]^2, gamma 5/3: rho=2, vx=+0.5 where |y-0.5|<0.25,
    else rho=1, vx=-0.5; p=2.5; vy=0.01 sin(4 pi x); uniform noise of amplitude
    ``noise`` on vx then vy per cell, row-major over the interior."""
    xc = cell_centres(0.0, 1.0, n, g)[g:g + n]
    yc = cell_centres(0.0, 1.0, n, g)[g:g + n]
    X, Y = np.meshgrid(xc, yc)
    inside = np.abs(Y - 0.5) < 0.25
    rho = np.where(inside, 2.0, 1.0)
    vx = np.where(inside, 0.5, -0.5)
    vy = 0.01 * np.sin(4.0 * math.pi * X)
    p = np.full_like(rho, 2.5)
    if noise:
        r = uniform01(seed, 2 * n * n).reshape(n, n, 2)
        vx = vx + noise * (2.0 * r[..., 0] - 1.0)
        vy = vy + noise * (2.0 * r[..., 1] - 1.0)
    return [rho, vx, vy, p]


def sedov(n: int, g: int = 2, gamma: float = 1.4, energy: float = 1.0, p0: float = 1e-5,
          radius_cells: float = 3.5, lo: float = -1.0, hi: float = 1.0):
    """Sedov blast on [lo,hi]^2: rho=1, u=0, p=p0, with ``energy`` deposited
    uniformly as pressure in the cells whose centre lies within
    ``radius_cells`` cell widths of the origin."""
    c = cell_centres(lo, hi, n, g)
    h = (hi - lo) / (n + 2 * g)
    xc = c[g:g + n]
    X, Y = np.meshgrid(xc, xc)
    r = np.sqrt(X * X + Y * Y)
    hot = r < radius_cells * h
    n_hot = int(hot.sum())
    p = np.full((n, n), p0)
    p[hot] = (gamma - 1.0) * energy / (n_hot * h * h)
    z = np.zeros((n, n))
    return [np.ones((n, n)), z, z.copy(), p]


def sod(n: int = 1000, g: int = 2, split: float = 0.5):
    """Sod shock tube on [0,1]: (1,0,1) left of ``split``, (0.125,0,0.1) right,
    set on every cell including the ghosts."""
    xc = cell_centres(0.0, 1.0, n, g)
    left = xc < split
    return [np.where(left, 1.0, 0.125), np.zeros(n + 2 * g), np.where(left, 1.0, 0.1)]


def four_quadrant(n: int, states: np.ndarray, g: int = 2):
    """Four-quadrant Riemann problem on [0,1]^2.  ``states`` is (4, 4) with rows
    in quadrant order NE, NW, SW, SE and columns (rho, p, vx, vy)."""
    c = cell_centres(0.0, 1.0, n, g)[g:g + n]
    X, Y = np.meshgrid(c, c)
    east, north = X >= 0.5, Y >= 0.5
    quad = np.where(north, np.where(east, 0, 1), np.where(east, 3, 2))
    rho = states[quad, 0]
    p = states[quad, 1]
    vx = states[quad, 2]
    vy = states[quad, 3]
    return [rho, vx, vy, p]


def ensemble_specs(count: int = 256, seed: int = 2, side_lo: int = 64, side_hi: int = 1024):
    """Side lengths drawn log-uniform in [side_lo, side_hi] (one draw per
    instance), then per instance and per quadrant rho, p ~ U[0.1, 2] and
    vx, vy ~ U[-1, 1], all from one splitmix64 stream."""
    u = uniform01(seed, count)
    sides = np.rint(np.exp(math.log(side_lo) + u * math.log(side_hi / side_lo))).astype(np.int64)
    r = uniform01(seed, count * 16, start=count).reshape(count, 4, 4)
    states = np.empty_like(r)
    states[..., 0] = 0.1 + 1.9 * r[..., 0]
    states[..., 1] = 0.1 + 1.9 * r[..., 1]
    states[..., 2] = -1.0 + 2.0 * r[..., 2]
    states[..., 3] = -1.0 + 2.0 * r[..., 3]
    return sides, states


def normal01(seed: int, n: int) -> np.ndarray:
    """n standard normal deviates from one splitmix64 stream (Box-Muller)."""
    u = uniform01(seed, 2 * ((n + 1) // 2)).reshape(2, -1)
    r = np.sqrt(-2.0 * np.log(1.0 - u[0]))   # 1 - u in (0, 1]
    z = np.concatenate([r * np.cos(2.0 * math.pi * u[1]), r * np.sin(2.0 * math.pi * u[1])])
    return z[:n]


def smooth_field(n: int, seed: int, sigma: float = 16.0) -> np.ndarray:
    """White noise on an n x n periodic grid smoothed by a Gaussian of
    `sigma` cells (via FFT), scaled to max |f| = 1."""
    z = normal01(seed, n * n).reshape(n, n)
    k = np.fft.fftfreq(n)
    kk = k[:, None] ** 2 + k[None, :] ** 2
    f = np.fft.ifft2(np.fft.fft2(z) * np.exp(-2.0 * (math.pi * sigma) ** 2 * kk)).real
    return f / np.abs(f).max()


def adjoint_initial(n: int, seed: int = 0, sigma: float = 16.0):
    """Primitive interior state (rho, u, v, p) of the adjoint workload:
    rho = 1 + 0.5 f1, p = 1 + 0.5 f2 (both in [0.5, 1.5]), (u, v) = 0.3
    (f3, f4) / max |(f3, f4)|, from smooth fields of seeds 3..6 (+ 100 seed)."""
    f = [smooth_field(n, 3 + k + 100 * seed, sigma) for k in range(4)]
    speed = np.sqrt(f[2] ** 2 + f[3] ** 2).max()
    return np.stack([1.0 + 0.5 * f[0], 0.3 * f[2] / speed, 0.3 * f[3] / speed, 1.0 + 0.5 * f[1]])


def adjoint_perturbation(n: int, seed: int = 0, sigma: float = 16.0, amplitude: float = 1e-2) -> np.ndarray:
    """Density perturbation of the target run: amplitude * f5 (seed 7 + 100 seed)."""
    return amplitude * smooth_field(n, 7 + 100 * seed, sigma)
