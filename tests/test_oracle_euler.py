"""Pins the Euler oracle (oracle/euler.py) against the reference's golden
vectors and the invariants of the scheme."""
import os

import numpy as np
import pytest

from oracle import euler
from oracle import swept as sw

GOLD = os.path.join(os.path.dirname(__file__), "golden", "euler_step.npz")
GAMMA = 1.4


def _solve(stepfn, q0, stages, dx, dt):
    ring, fin = sw.standard_solve(lambda q, qb, g: stepfn(q, qb, g, GAMMA, dt, dx, dx), q0, 2, stages)
    return ring.get(fin)


@pytest.mark.parametrize("tag", ["random", "vortex"])
@pytest.mark.parametrize("gts", [1, 2])
def test_golden_vectors(tag, gts):
    """Reference euler.step (euler.py:131-313): the py order with the C
    library's pow for squares is bitwise equal; with a*a it is within
    1e-12 (the device plugin's choice); the device order within 1e-15."""
    g = np.load(GOLD)
    q, dx, dt = g[f"{tag}_in"], float(g[f"{tag}_dx"]), float(g[f"{tag}_dt"])
    ref = g[f"{tag}_g{gts}_out"]
    out = euler.step_py(q[1], q[0], gts, GAMMA, dt, dx, dx, sq=euler.sq_pow)
    assert np.array_equal(out, ref)
    out = euler.step_py(q[1], q[0], gts, GAMMA, dt, dx, dx, sq=euler.sq_mul)
    assert np.abs(out - ref).max() <= 1e-12 * np.abs(ref).max()
    out = euler.step_cu(q[1], q[0], gts, GAMMA, dt, dx, dx)
    assert np.abs(out - ref).max() <= 1e-15 * np.abs(ref).max()


def test_vortex_initial_condition_matches_reference():
    """vortex() restates euler.analytical (euler.py:63-105)."""
    g = np.load(GOLD)
    n = g["vortex_in"].shape[-1]
    X, Y = np.meshgrid(euler.grid(n), euler.grid(n), indexing="ij")
    assert np.abs(euler.vortex(X, Y, 0.0, GAMMA) - g["vortex_in"][0]).max() < 1e-14


@pytest.mark.parametrize("fn", [euler.step_py, euler.step_cu])
def test_free_stream_preserved_exactly(fn):
    n = 32
    q = np.zeros((4, n, n))
    q[0], q[1], q[2], q[3] = 1.3, 0.7, -0.4, 3.1
    out = _solve(fn, q, 6, 10.0 / n, 1.0 / n)
    assert np.array_equal(out, q)


@pytest.mark.parametrize("fn", [euler.step_py, euler.step_cu])
def test_conservation(fn):
    """Each face flux enters its two cells with opposite signs from identical
    arguments, so the four integrals are conserved to summation round-off."""
    n = 48
    dx = 10.0 / n
    q0 = euler.initial_condition(n, eps=3.0)
    out = _solve(fn, q0, 10, dx, 0.1 * dx)
    for v in range(4):
        assert abs(out[v].sum() - q0[v].sum()) <= 1e-13 * np.abs(q0[v]).sum()


@pytest.mark.parametrize("fn", [euler.step_py, euler.step_cu])
def test_mirror_symmetry_exact(fn):
    """A state symmetric under x<->y with rho*u<->rho*v stays so bitwise,
    which fails for a swapped axis or spectral component."""
    n = 24
    rng = np.random.default_rng(7)
    a = rng.uniform(-0.1, 0.1, (n, n))
    q = np.empty((4, n, n))
    q[0] = 1.0 + (a + a.T)
    m = rng.uniform(-0.2, 0.2, (n, n))
    q[1], q[2] = m, m.T.copy()
    e = rng.uniform(-0.1, 0.1, (n, n))
    q[3] = 2.5 + (e + e.T)
    out = _solve(fn, q, 4, 10.0 / n, 1.0 / n)
    assert np.array_equal(out[0], out[0].T)
    assert np.array_equal(out[1], out[2].T)
    assert np.array_equal(out[3], out[3].T)


def test_vortex_convergence():
    """Mean error against the advected vortex (euler.py:63-116) falls as dx
    halves (the pressure-ratio limiter makes the scheme roughly first order
    on this flow: measured ratios 1.6-1.9)."""
    errs = []
    t_end = 0.5
    for n in (32, 64):
        dx = 10.0 / n
        dt = 0.1 * dx
        steps = int(round(t_end / dt))
        q0 = euler.initial_condition(n, eps=2.0)
        out = _solve(euler.step_py, q0, 2 * steps, dx, dt)
        X, Y = np.meshgrid(euler.grid(n), euler.grid(n), indexing="ij")
        errs.append(np.abs(out[0] - euler.vortex(X, Y, steps * dt, GAMMA, eps=2.0)[0]).mean())
    assert errs[0] / errs[1] > 1.4


def test_planted_flux_error_fails_golden():
    """Dropping the spectral term (planted on the CPU) breaks the golden match."""
    g = np.load(GOLD)
    q, dx, dt = g["random_in"], float(g["random_dx"]), float(g["random_dt"])
    saved = euler.spectral_py
    try:
        euler.spectral_py = lambda L, R, dim, gamma: 0.0 * L
        out = euler.step_py(q[1], q[0], 1, GAMMA, dt, dx, dx)
    finally:
        euler.spectral_py = saved
    assert np.abs(out - g["random_g1_out"]).max() > 1e-6


def test_round_off_sensitivity_is_a_property_of_the_scheme():
    """Records why multi-stage Euler parity is checked bitwise or stage by
    stage: a 1-ulp change of the vortex initial state is amplified to far
    above 1e-12 by the limiter's zeroing of pressure ratios (euler.py:253-254)
    in the nearly uniform far field."""
    n = 64
    dx = 10.0 / n
    q0 = euler.initial_condition(n)
    rng = np.random.default_rng(3)
    q1 = q0 * (1 + rng.choice([-1, 0, 1], size=q0.shape) * 2.2e-16)
    a = _solve(euler.step_py, q0, 4, dx, 0.1 * dx)
    b = _solve(euler.step_py, q1, 4, dx, 0.1 * dx)
    assert np.abs(a - b).max() > 1e-10
