"""Known-answer plugins (test infrastructure only; see oracle/__init__.py).

Whole-grid periodic restatements of the reference's test plugins; ``q`` is
the level read, ``qb`` the level one back, arrays are ``q[v, x, y]``.
"""
import numpy as np


def example_step(q, qb, gts, scheme_fe):
    """example.py:13-26 / example.cu:10-23: +1 per stage (FE); RK2-shaped:
    q[t-1] + 2 on final stages (gts%2==0), q[t] + 1 otherwise."""
    if scheme_fe:
        return q + 1
    if gts % 2 == 0:
        return qb + 2
    return q + 1


def checker_step(q, qb, gts):
    """checker.py:24-44 / checker.cu:10-32: average of the 4 neighbours,
    summed x+1, x-1, y+1, y-1 as the numpy plugin does."""
    s = np.roll(q, -1, axis=1)
    s = s + np.roll(q, 1, axis=1)
    s = s + np.roll(q, -1, axis=2)
    s = s + np.roll(q, 1, axis=2)
    return s / 4


def half_step(q, qb, gts):
    """half.py:17-21, 44-48: RK2 midpoint heat with the factor 1/100:
    cd = (u_e-2u+u_w)/100 + (u_n-2u+u_s)/100;
    final (gts%2==0): u[t-1] + 1*cd; intermediate: u[t] + 0.5*cd."""
    xp, xm = np.roll(q, -1, axis=1), np.roll(q, 1, axis=1)
    yp, ym = np.roll(q, -1, axis=2), np.roll(q, 1, axis=2)
    cd = (xp - 2 * q + xm) / 100 + (yp - 2 * q + ym) / 100
    if gts % 2 == 0:
        return qb + 1 * cd
    return q + 0.5 * cd


def checker_ic(nv, nx, ny):
    """checker.py:45-57: 1 on (even, even) and (odd, odd) cells."""
    d = np.zeros((nv, nx, ny))
    d[:, 0::2, 0::2] = 1
    d[:, 1::2, 1::2] = 1
    return d


def half_ic(nx, ny):
    """half.py:54-68: cell value = flat index / 100."""
    return (np.arange(nx * ny, dtype=np.float64) / 100).reshape(1, nx, ny)


def half_brute_force(u0, nsteps):
    """The reference's in-test restatement of the half scheme
    (tests.py:243-253): two midpoint stages per step with /100."""
    a0 = u0.copy()
    for _ in range(nsteps):
        xp, xm = np.roll(a0, -1, 0), np.roll(a0, 1, 0)
        yp, ym = np.roll(a0, -1, 1), np.roll(a0, 1, 1)
        a1 = ((xp - 2 * a0 + xm) + (yp - 2 * a0 + ym)) * (0.5 / 100) + a0
        xp, xm = np.roll(a1, -1, 0), np.roll(a1, 1, 0)
        yp, ym = np.roll(a1, -1, 1), np.roll(a1, 1, 1)
        a2 = ((xp - 2 * a1 + xm) + (yp - 2 * a1 + ym)) / 100 + a0
        a0 = a2
    return a0
