"""Differentiable restatement, in torch float64 on the CPU, of the fixed-dt 2D
MUSCL-Hancock + HLLC step on a doubly periodic Cartesian grid.  Test
infrastructure: autograd of this code is the CPU reference gradient that the
device adjoint (hfv_adj2d_step) is checked against, and its forward is
checked against the oracle.

Arrays hold the interior only, shape (4, ny, nx), fields (m, mx, my, E) for
conserved and (rho, u, v, p) for primitive states; periodic neighbours come
from torch.roll.  Branches are selected with torch.where over inputs made
safe in the branches not taken, so autograd differentiates the branch the
forward evaluation takes (the one-sided convention of the device adjoint).
"""
from __future__ import annotations

import torch

MINMOD, VAN_LEER, MC = 1, 2, 3


def sound_speed(gamma, rho, p):
    """utils.c:23-26"""
    return torch.sqrt(gamma * p / rho)


def limit(lim, a, b):
    """slope_limiter.c:49-91, elementwise."""
    zero = a * b <= 0.0
    if lim == MINMOD:
        d = torch.where(a.abs() < b.abs(), a, b)
    elif lim == VAN_LEER:
        den = torch.where(zero, torch.ones_like(a), a + b + 1e-10)
        d = 2.0 * a * b / den
    else:
        cen = torch.where(a > 0, 0.5 * (a + b), -0.5 * (a + b))
        m2 = 2.0 * torch.minimum(a.abs(), b.abs())
        d = torch.where(cen <= m2, 0.5 * (a + b), torch.where(a.abs() <= b.abs(), 2.0 * a, 2.0 * b))
    return torch.where(zero, torch.zeros_like(a), d)


def c2p(gamma, vol, U):
    """system.c:763-802"""
    m, mx, my, E = U
    u = mx / m
    v = my / m
    return torch.stack([m / vol, u, v, (gamma - 1.0) * (E - 0.5 * m * (u * u + v * v)) / vol])


def predict_x(s, d, c, a):
    """integrator_muscl_hancock.c:692-740"""
    up = 1.0 + c * s[1]
    um = 1.0 - c * s[1]
    lo = torch.stack([
        s[0] - 0.5 * (up * d[0] + c * s[0] * d[1]),
        s[1] - 0.5 * (up * d[1] + (c / s[0]) * d[3]),
        s[2] - 0.5 * (up * d[2]),
        s[3] - 0.5 * (c * s[0] * a * a * d[1] + up * d[3])])
    hi = torch.stack([
        s[0] + 0.5 * (um * d[0] - c * s[0] * d[1]),
        s[1] + 0.5 * (um * d[1] - (c / s[0]) * d[3]),
        s[2] + 0.5 * (um * d[2]),
        s[3] + 0.5 * (-c * s[0] * a * a * d[1] + um * d[3])])
    return lo, hi


def predict_y(s, d, c, a):
    """integrator_muscl_hancock.c:977-1025 (the density row takes the slope
    of u, as the reference writes it)"""
    vp = 1.0 + c * s[2]
    vm = 1.0 - c * s[2]
    lo = torch.stack([
        s[0] - 0.5 * (vp * d[0] + c * s[0] * d[1]),
        s[1] - 0.5 * (vp * d[1]),
        s[2] - 0.5 * (vp * d[2] + (c / s[0]) * d[3]),
        s[3] - 0.5 * (c * s[0] * a * a * d[2] + vp * d[3])])
    hi = torch.stack([
        s[0] + 0.5 * (vm * d[0] - c * s[0] * d[1]),
        s[1] + 0.5 * (vm * d[1]),
        s[2] + 0.5 * (vm * d[2] - (c / s[0]) * d[3]),
        s[3] + 0.5 * (-c * s[0] * a * a * d[2] + vm * d[3])])
    return lo, hi


def hllc(gamma, tol, wl, wr):
    """riemann_solver_hllc.c:30-376 with the pressure guess of
    riemann_solver.c:154-224, for states (4, ...) in sweep-local order
    (rho, u_n, u_t, p); returns the flux (mass, normal, transverse, energy).
    Vacuum and invalid wave speeds raise (tests use states without them)."""
    rl, ul, vl, pl = wl
    rr, ur, vr, pr = wr
    al = sound_speed(gamma, rl, pl)
    ar = sound_speed(gamma, rr, pr)
    if bool(((al + ar) * 2.0 / (gamma + 1.0) <= (ur - ul)).any()):
        raise ValueError("vacuum generation in the torch reference")
    one = torch.ones_like(pl)
    plo = torch.where(pl > pr, pr, pl)
    phi = torch.where(pl > pr, pl, pr)
    ppv = 0.5 * (pl + pr) - 0.125 * (ur - ul) * (rl + rr) * (al + ar)
    pvrs = (phi / plo <= 2.0) & (plo <= ppv) & (ppv <= phi)
    trrs = ~pvrs & (ppv < plo)
    tsrs = ~pvrs & ~trrs
    h = 0.5 * (gamma - 1.0)
    x = (al + ar - h * (ur - ul)) / (al / pl ** (h / gamma) + ar / pr ** (h / gamma))
    g_trrs = torch.where(trrs, x, one) ** (gamma / h)
    bl = ppv + pl * ((gamma - 1.0) / (gamma + 1.0))
    br = ppv + pr * ((gamma - 1.0) / (gamma + 1.0))
    gl = torch.sqrt((2.0 / ((gamma + 1.0) * rl)) / torch.where(tsrs, bl, one))
    gr = torch.sqrt((2.0 / ((gamma + 1.0) * rr)) / torch.where(tsrs, br, one))
    g_tsrs = (gl * pl + gr * pr - (ur - ul)) / (gl + gr)
    ps = torch.where(pvrs, ppv, torch.where(trrs, g_trrs, g_tsrs))
    ps = torch.where(ps < tol, torch.full_like(ps, tol), ps)
    kq = 0.5 + 0.5 / gamma
    ql = torch.where(ps > pl, torch.sqrt(1.0 + kq * (ps / pl - 1.0)), one)
    qr = torch.where(ps > pr, torch.sqrt(1.0 + kq * (ps / pr - 1.0)), one)
    sl = ul - al * ql
    sr = ur + ar * qr
    sm = (pr - pl + rl * ul * (sl - ul) - rr * ur * (sr - ur)) / (rl * (sl - ul) - rr * (sr - ur))
    r0 = 0.0 <= sl
    r1 = ~r0 & (sl <= 0.0) & (0.0 <= sm)
    r2 = ~r0 & ~r1 & (sm <= 0.0) & (0.0 <= sr)
    r3 = ~r0 & ~r1 & ~r2 & (sr <= 0.0)
    if not bool((r0 | r1 | r2 | r3).all()):
        raise ValueError("invalid HLLC wave speeds in the torch reference")
    left = r0 | r1
    star = r1 | r2
    r = torch.where(left, rl, rr)
    u = torch.where(left, ul, ur)
    v = torch.where(left, vl, vr)
    p = torch.where(left, pl, pr)
    s = torch.where(left, sl, sr)
    e = r * (0.5 * (u * u + v * v) + p / (r * (gamma - 1.0)))
    fm = r * u
    fu = fm * u + p
    fe = u * (e + p)
    D = torch.where(star, s - sm, one)
    rs = r * (s - u) / D
    es = rs * (e / r + (sm - u) * (sm + p / (r * (s - u))))
    f_star = torch.stack([fm + s * (rs - r), fu + s * (rs * sm - r * u), fm * v + s * (rs * v - r * v),
                          fe + s * (es - e)])
    f_one = torch.stack([fm, fu, fm * v, fe])
    return torch.where(star, f_star, f_one)


def step(U, gamma, dt, dx, dy, lim=MINMOD, tol=1e-6):
    """One step U^n -> U^{n+1} (interior, periodic), uniform cells dx x dy
    (volume dx*dy, face areas dy for x faces and dx for y faces)."""
    vol = dx * dy
    ax, ay = dy, dx
    W = c2p(gamma, vol, U)
    # x sweep (integrator_muscl_hancock.c:600-815)
    d = limit(lim, W - torch.roll(W, 1, 2), torch.roll(W, -1, 2) - W)
    a = sound_speed(gamma, W[0], W[3])
    lo, hi = predict_x(W, d, dt * ax / vol, a)
    fxl = hllc(gamma, tol, torch.roll(hi, 1, 2), lo)        # flux through each cell's left face
    fxr = torch.roll(fxl, -1, 2)
    # y sweep (integrator_muscl_hancock.c:885-1101), solver order (rho, v, u, p)
    d = limit(lim, W - torch.roll(W, 1, 1), torch.roll(W, -1, 1) - W)
    lo, hi = predict_y(W, d, dt * ay / vol, a)
    perm = [0, 2, 1, 3]
    fyb = hllc(gamma, tol, torch.roll(hi, 1, 1)[perm], lo[perm])[perm]
    fyt = torch.roll(fyb, -1, 1)
    x = U + fxl * dt * ax
    x = x - fxr * dt * ax
    x = x + fyb * dt * ay
    return x - fyt * dt * ay
