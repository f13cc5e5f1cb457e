"""Discrete adjoint of the fixed-dt periodic 2D MUSCL-Hancock + HLLC step on
the device (hfv_fwd2d_step / hfv_adj2d_step through hydrofv_amd.adjoint):

- the forward step equals the oracle's fixed-dt step;
- one adjoint step equals the vector-Jacobian product of the CPU reference
  (torch float64 autograd of _torch_ref.step, itself pinned to the oracle),
  on grids that span several x-pass tiles and y-pass strips;
- gradients of multi-step rollouts with sqrt-N checkpointing equal the CPU
  reference gradient, and directional derivatives match central finite
  differences of the device forward on a 64 x 64 copy of the adjoint
  workload to 1e-6 relative.
"""
import numpy as np
import pytest
import torch

import _cases as K
import _torch_ref as T
from hydrofv_amd import adjoint, workloads

pytestmark = pytest.mark.gpu
GAMMA = 1.4


def conserved(W, vol, g):
    """Interior primitive (4, ny, nx) -> conserved with ghost cells (zeros)."""
    rho, u, v, p = W
    m = rho * vol
    U = np.stack([m, m * u, m * v, vol * (p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v))])
    out = np.zeros((4, U.shape[1] + 2 * g, U.shape[2] + 2 * g))
    out[:, g:-g, g:-g] = U
    return out


def cfl_dt(W, dx, dy, cfl=0.8):
    a = np.sqrt(GAMMA * W[3] / W[0])
    return cfl / ((a + np.abs(W[1])).max() / dx + (a + np.abs(W[2])).max() / dy)


def grid_for(nx, ny, W, limiter=1, cfl=0.8, g=2):
    # the unit square as the reference sets it up: its cell size counts the
    # ghost cells, dx = (x_max - x_min) / (nx + 2g) (system.c:400-415)
    dx, dy = 1.0 / (nx + 2 * g), 1.0 / (ny + 2 * g)
    return adjoint.Periodic2D(nx, ny, dx, dy, cfl_dt(W, dx, dy, cfl), gamma=GAMMA, limiter=limiter)


def random_W(nx, ny, seed):
    return np.stack(K.random_state_2d(nx, ny, seed))


def ref_vjp(grid, U_int, lam_int, steps=1):
    U = torch.tensor(U_int, requires_grad=True)
    x = U
    for _ in range(steps):
        x = T.step(x, GAMMA, grid.dt, grid.dx, grid.dy, lim=grid.limiter)
    (g,) = torch.autograd.grad(x, U, torch.tensor(lam_int))
    return x.detach().numpy(), g.numpy()


@pytest.mark.parametrize("nx,ny", [(70, 37), (13, 11)])
def test_forward_step_matches_oracle(gpulib, nx, ny):
    g = 2
    sim = K.setup_random_2d(gpulib, nx, ny, g, ("periodic",) * 4, seed=21, gamma=GAMMA)
    W = np.stack(sim.primitive())[:, g:-g, g:-g]
    grid = grid_for(nx, ny, W)
    assert (grid.dx, grid.dy) == (sim.system.dx_, sim.system.dy_)
    oracle = K.oracle_grid(sim, "minmod")
    r = oracle.run(0.8, max_steps=3, fixed_dt=grid.dt)
    assert r.error == 0
    U0 = torch.tensor(np.stack(sim.conserved()), device="cuda")
    U3 = grid.run(U0, 3)
    ref = oracle.U[:, g:-g, g:-g]
    got = grid.interior(U3).cpu().numpy()
    err = max(K.rel_err(got[k], ref[k]) for k in range(4))
    assert err <= 1e-12, err
    print(f"fixed-dt forward {nx}x{ny} x3: max rel diff {err:.2e}, bitwise={np.array_equal(got, ref)}")
    sim.close()


@pytest.mark.parametrize("nx,ny,lim", [(70, 37, 1), (13, 11, 1), (131, 20, 1), (40, 33, 3), (40, 33, 2)])
def test_adjoint_step_matches_reference(nx, ny, lim):
    W = random_W(nx, ny, 30 + nx)
    grid = grid_for(nx, ny, W, limiter=lim)
    U = conserved(W, grid.volume, grid.g)
    lam = np.zeros_like(U)
    lam[:, 2:-2, 2:-2] = np.random.default_rng(nx + ny).normal(size=(4, ny, nx))
    _, g_ref = ref_vjp(grid, U[:, 2:-2, 2:-2].copy(), lam[:, 2:-2, 2:-2].copy())
    Ut = torch.tensor(U, device="cuda")
    lt = torch.tensor(lam, device="cuda")
    out = torch.zeros_like(lt)
    grid.adjoint_step(Ut, lt, out)
    flags = grid.take_flags(Ut.device)
    assert flags == 0, flags
    got = out.cpu().numpy()
    assert np.all(got[:, :2, :] == 0) and np.all(got[:, :, :2] == 0)   # ghosts untouched
    got = got[:, 2:-2, 2:-2]
    for k in range(4):
        err = K.rel_err(got[k], g_ref[k])
        assert err <= 1e-10, (k, err)


def test_rollout_gradient_matches_reference():
    nx, ny, steps = 24, 18, 5
    W = random_W(nx, ny, 3)
    grid = grid_for(nx, ny, W, cfl=0.6)
    U0 = conserved(W, grid.volume, grid.g)
    target = W[0] + 0.01 * np.sin(np.linspace(0, 6, nx))[None, :]
    Ut = torch.tensor(U0, device="cuda", requires_grad=True)
    UN = adjoint.rollout(Ut, grid, steps, every=2)
    loss = adjoint.density_loss(UN, torch.tensor(target, device="cuda"), grid)
    loss.backward()
    got = Ut.grad.cpu().numpy()
    # CPU reference
    U = torch.tensor(U0[:, 2:-2, 2:-2].copy(), requires_grad=True)
    x = U
    for _ in range(steps):
        x = T.step(x, GAMMA, grid.dt, grid.dx, grid.dy)
    ref_loss = 0.5 * ((x[0] / grid.volume - torch.tensor(target)) ** 2).sum()
    (g_ref,) = torch.autograd.grad(ref_loss, U)
    assert abs(loss.item() - ref_loss.item()) <= 1e-12 * abs(ref_loss.item())
    assert np.all(got[:, :2, :] == 0)
    for k in range(4):
        err = K.rel_err(got[k, 2:-2, 2:-2], g_ref.numpy()[k])
        assert err <= 1e-10, (k, err)


def test_rollout_identity_and_ghost_entries():
    """Zero steps are the identity, gradient included.  After an odd number
    of steps the result's ghost entries are defined values (no memory left
    unwritten), and the gradient reaches the interior only."""
    nx, ny = 20, 14
    W = random_W(nx, ny, 8)
    grid = grid_for(nx, ny, W, cfl=0.6)
    U0 = conserved(W, grid.volume, grid.g)
    U0[:, :2, :] = 0.25  # distinct ghost values
    wt = torch.tensor(np.random.default_rng(4).standard_normal(U0.shape), device="cuda")
    ghost = torch.ones(U0.shape, dtype=torch.bool, device="cuda")
    grid.interior(ghost)[...] = False
    for steps in (0, 3):
        Ut = torch.tensor(U0, device="cuda", requires_grad=True)
        UN = adjoint.rollout(Ut, grid, steps, every=2)
        (UN * wt).sum().backward()
        if steps == 0:
            assert torch.equal(UN.detach(), Ut.detach())
            assert torch.equal(Ut.grad, wt)
        else:
            assert bool(torch.isfinite(UN.detach()).all())
            assert bool((Ut.grad[ghost] == 0).all())
            assert bool((grid.interior(Ut.grad) != 0).any())


def test_gradient_matches_finite_differences_64():
    """Central differences of the device forward along smooth random
    directions: the adjoint workload (smooth Gaussian random field) on a
    64 x 64 copy, 40 steps, density mismatch loss."""
    n, steps = 64, 40
    W = workloads.adjoint_initial(n, seed=0, sigma=4.0)
    grid = grid_for(n, n, W)
    U0 = conserved(W, grid.volume, grid.g)
    Wt = W.copy()
    Wt[0] += workloads.adjoint_perturbation(n, seed=0, sigma=4.0)
    Ut0 = torch.tensor(conserved(Wt, grid.volume, grid.g), device="cuda")
    rho_t = grid.interior(grid.run(Ut0, steps))[0] / grid.volume

    U = torch.tensor(U0, device="cuda", requires_grad=True)
    loss = adjoint.density_loss(adjoint.rollout(U, grid, steps), rho_t, grid)
    loss.backward()
    grad = U.grad.detach()
    assert grid.last_flags & adjoint.VACUUM == 0

    def L(Ux):
        return adjoint.density_loss(grid.run(Ux, steps), rho_t, grid).item()

    worst = 0.0
    for j in range(6):
        d = np.zeros_like(U0)
        for k in range(4):
            d[k, 2:-2, 2:-2] = workloads.smooth_field(n, 50 + 4 * j + k, 4.0) * np.abs(U0[k, 2:-2, 2:-2]).max()
        dt_ = torch.tensor(d, device="cuda")
        # The loss is piecewise smooth (limiter and HLLC branch points); the
        # central difference's error falls linearly with h from the kinks
        # crossed within +-h, down to the rounding floor near h = 1e-8.
        h = 1e-8
        fd = (L(U.detach() + h * dt_) - L(U.detach() - h * dt_)) / (2 * h)
        ad = float((grad * dt_).sum())
        rel = abs(fd - ad) / abs(ad)
        worst = max(worst, rel)
        print(f"direction {j}: adjoint {ad:.10e} central difference {fd:.10e} rel {rel:.2e}")
    assert worst <= 1e-6, worst
