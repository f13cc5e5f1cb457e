"""The oracle's random-choice loop (random_choice_1d,
integrator_random_choice_1d.c) against the fixtures of the reference build
(tests/golden/runs_random_choice.npz), bitwise; the van der Corput sequence
against its exact rational values; and the method's accuracy on the Sod tube
against the exact solution."""
import os
from fractions import Fraction

import numpy as np
import pytest

import _oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

CASES = [
    ("rc_sod200", "cartesian_1d", 200, 2, ("transmissive",) * 2, 0.5, 0.2, 1e-6),
    ("rc_sod100_tol8", "cartesian_1d", 100, 2, ("transmissive",) * 2, 0.5, 0.2, 1e-8),
    ("rc_cyl64", "cylindrical_1d", 64, 2, ("reflective", "transmissive"), 0.5, 0.2, 1e-6),
    ("rc_sph64", "spherical_1d", 64, 2, ("reflective", "transmissive"), 0.5, 0.2, 1e-6),
    ("rc_sod50g3_periodic", "cartesian_1d", 50, 3, ("periodic",) * 2, 0.5, 0.05, 1e-6),
]


@pytest.mark.parametrize("key,coord,n,g,bcs,cfl,tf,tol", CASES, ids=[c[0] for c in CASES])
def test_random_choice_golden(key, coord, n, g, bcs, cfl, tf, tol):
    d = np.load(os.path.join(GOLD, "runs_random_choice.npz"), allow_pickle=False)
    grid = O.Grid(coord, n, 1, g, 1.4, bc=bcs, solver="riemann_solver_exact", tol=tol)
    grid.W[...] = d[f"{key}/W0"]
    grid.U[...] = d[f"{key}/U0"]
    r = grid.run(cfl, tf=tf, random_choice=True)
    assert r.error == 0
    assert (r.num_steps, r.t, r.dt) == (int(d[f"{key}/n"]), float(d[f"{key}/t"]), float(d[f"{key}/dt"]))
    assert np.array_equal(grid.W, d[f"{key}/W1"], equal_nan=True)
    assert np.array_equal(grid.U, d[f"{key}/U1"], equal_nan=True)


def van_der_corput_exact(n, base=5, mult=3):
    """The sequence of integrator_random_choice_1d.c:28-43 in exact rational
    arithmetic: digit a_i of n in base 5 contributes ((3 a_i) mod 5) / 5^(i+1)."""
    x, bk = Fraction(0), Fraction(1, base)
    while n > 0:
        x += ((mult * (n % base)) % base) * bk
        n //= base
        bk /= base
    return x


def test_van_der_corput_values():
    L = O.lib()
    for n in list(range(0, 200)) + [5 ** 7 - 1, 5 ** 7, 123456789]:
        got = L.hfo_van_der_corput(n)
        want = van_der_corput_exact(n)
        assert abs(Fraction(got) - want) <= Fraction(1, 2 ** 50), (n, got, float(want))
    # the first terms, and the sequence fills [0, 1) evenly
    assert [round(L.hfo_van_der_corput(n), 12) for n in range(1, 6)] == [0.6, 0.2, 0.8, 0.4, 0.12]
    vals = np.array([L.hfo_van_der_corput(n) for n in range(1, 5 ** 4)])
    hist, _ = np.histogram(vals, bins=5, range=(0, 1))
    assert hist.min() >= 124 and hist.max() <= 125


def test_sod_random_choice_accuracy():
    """Glimm's method keeps shocks and contacts sharp: every cell holds a
    state of the exact solution, so away from the waves the error is zero
    and the L1 error comes from the wave positions only."""
    n, g = 400, 2
    grid = O.Grid("cartesian_1d", n, 1, g, 1.4, bc=("transmissive",) * 2, solver="riemann_solver_exact")
    x = (np.arange(n + 2 * g) + 0.5) / (n + 2 * g)
    left = x < 0.5
    grid.W[0] = np.where(left, 1.0, 0.125)
    grid.W[1] = 0.0
    grid.W[2] = np.where(left, 1.0, 0.1)
    grid.p2c()
    assert grid.run(0.5, tf=0.2, random_choice=True).error == 0
    xi = x[g:g + n]
    sol = O.oracle_exact_sample(1.4, 1e-12, [1.0, 0.0, 1.0], [0.125, 0.0, 0.1], (xi - 0.5) / 0.2)
    h = 1.0 / (n + 2 * g)
    l1 = np.sum(np.abs(grid.W[0][g:g + n] - sol[:, 0])) * h
    assert l1 < 4e-3, l1
    # far from every wave the state is exact to the solver's tolerance
    far = (xi < 0.2) | (xi > 0.95)
    assert np.allclose(grid.W[0][g:g + n][far], sol[far, 0], rtol=0, atol=1e-12)
