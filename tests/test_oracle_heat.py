"""Pins the heat oracle (oracle/heat.py) against the reference's golden
vectors, the exact discrete Fourier solution and the analytic decay."""
import os

import numpy as np
import pytest

from oracle import heat
from oracle import swept as sw

GOLD = os.path.join(os.path.dirname(__file__), "golden", "heat_step.npz")


def _run(step, u0, nstages, s):
    ring, fin = sw.standard_solve(step, u0[None], s, nstages)
    return ring.get(fin)[0]


@pytest.mark.parametrize("tag,fe,gts_list", [("fe", True, (1,)), ("rk2", False, (1, 2))])
def test_golden_vectors_bitwise(tag, fe, gts_list):
    """Reference heat.step (heat.py:21-64) on a seeded field: py order is
    bitwise equal, device order within 4 ulp."""
    g = np.load(GOLD)
    alpha, dt, dx = float(g["alpha"]), float(g["dt"]), float(g["dx"])
    c = heat.courant_py(alpha, dt, dx)
    for gts in gts_list:
        inp, ref = g[f"{tag}_g{gts}_in"], g[f"{tag}_g{gts}_out"][0]
        out = heat.step_py(inp[1, 0], inp[0, 0], gts, c, fe)
        assert np.array_equal(out, ref)
        outc = heat.step_cu(inp[1, 0], inp[0, 0], gts, alpha, dt, dx, dx, fe, 2)
        assert np.abs(outc - ref).max() <= 4e-16 * np.abs(ref).max() + 1e-300


@pytest.mark.parametrize("fe", [True, False])
def test_fourier_closed_form(fe):
    """u^n = IFFT(lambda^n FFT(u^0)) for the linear periodic scheme."""
    n, nsteps = 48, 60
    dx = 1.0 / n
    dt = 0.1 * dx * dx
    u0 = heat.initial_condition(n, seed=3)[0]
    c = heat.courant_py(1.0, dt, dx)
    s = 1 if fe else 2
    got = _run(lambda q, qb, g: heat.step_py(q[0], None if qb is None else qb[0], g, c, fe)[None],
               u0, s * nsteps, s)
    exact = heat.fourier_solution(u0, nsteps, c, fe)
    assert np.abs(got - exact).max() <= 1e-13 * np.abs(exact).max()


def test_planted_sign_error_fails_closed_form():
    """A wrong sign in the Laplacian (planted in a copy on the CPU) is caught."""
    n, nsteps = 32, 20
    dx = 1.0 / n
    dt = 0.1 * dx * dx
    u0 = heat.initial_condition(n, seed=4)[0]
    c = heat.courant_py(1.0, dt, dx)

    def bad(q, qb, g):
        u = q[0]
        xp, xm = np.roll(u, -1, 0), np.roll(u, 1, 0)
        yp, ym = np.roll(u, -1, 1), np.roll(u, 1, 1)
        return (c * (xp - 2 * u + xm) + c * (yp + 2 * u + ym) + u)[None]

    got = _run(bad, u0, nsteps, 1)
    exact = heat.fourier_solution(u0, nsteps, c, True)
    assert np.abs(got - exact).max() > 1e-6


@pytest.mark.parametrize("fe", [True, False])
def test_analytic_decay_second_order(fe):
    """exp(-8 pi^2 alpha t) decay of sin(2 pi x) sin(2 pi y) (heat.py:90-113):
    error falls ~4x when dx halves."""
    errs = []
    for n in (32, 64):
        dx = 1.0 / n
        dt = 0.1 * dx * dx
        c = heat.courant_py(1.0, dt, dx)
        nsteps = int(round(2e-3 / dt))
        u0 = heat.analytical(n, 0.0, 1.0, m=2)
        s = 1 if fe else 2
        got = _run(lambda q, qb, g: heat.step_py(q[0], None if qb is None else qb[0], g, c, fe)[None],
                   u0, s * nsteps, s)
        errs.append(np.abs(got - heat.analytical(n, nsteps * dt, 1.0, m=2)).max())
    assert 3.0 < errs[0] / errs[1] < 5.0


def test_py_and_cu_orders_agree():
    """Two statements of the same linear scheme stay within round-off."""
    n, nst = 32, 40
    dx = 1.0 / n
    dt = 0.1 * dx * dx
    u0 = heat.initial_condition(n, seed=5)[0]
    c = heat.courant_py(1.0, dt, dx)
    a = _run(lambda q, qb, g: heat.step_py(q[0], qb[0], g, c, False)[None], u0, nst, 2)
    b = _run(lambda q, qb, g: heat.step_cu(q[0], qb[0], g, 1.0, dt, dx, dx, False, 2)[None], u0, nst, 2)
    assert np.abs(a - b).max() <= 1e-13 * np.abs(a).max()
