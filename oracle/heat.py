"""Heat-equation oracle (test infrastructure only; see oracle/__init__.py).

All steps act on the whole periodic grid ``u[x, y]`` (axis 0 = x, axis 1 = y)
and return the next level.  ``u`` is the level being read (``arrayTimeIndex``
in the reference), ``ub`` the level one back (used by the final RK2 stage).
"""
import numpy as np


def _nbrs(u):
    """x+1, x-1, y+1, y-1 neighbours with periodic wrap."""
    return (np.roll(u, -1, axis=0), np.roll(u, 1, axis=0),
            np.roll(u, -1, axis=1), np.roll(u, 1, axis=1))


def courant_py(alpha, dt, dx):
    """courant = alpha*dt/dx/dx  (pysweep/equations/heat.py:37)."""
    return alpha * dt / dx / dx


def central_difference_py(u, courant):
    """Restates heat.py:59-64: courant*(u_e - 2u + u_w) + courant*(u_n - 2u + u_s),
    evaluated left to right."""
    xp, xm, yp, ym = _nbrs(u)
    d2x = courant * (xp - 2 * u + xm)
    d2y = courant * (yp - 2 * u + ym)
    return d2x + d2y


def step_py(u, ub, gts, courant, scheme_fe):
    """One stage in the order of the numpy plugin.

    forward Euler  heat.py:48-51  -> cd + u
    RK2 (midpoint) heat.py:53-57  -> coeff*cd + u[t - timechange] with
                                     (coeff, timechange) = (1, 1) if gts%2==0
                                     else (0.5, 0)
    """
    cd = central_difference_py(u, courant)
    if scheme_fe:
        return cd + u
    if gts % 2 == 0:
        return 1 * cd + ub
    return 0.5 * cd + u


def central_difference_cu(u, alpha, dt, dx, dy):
    """Restates heat.cu:11-19:
    ALPHA*DT*(u[x+1]-2u+u[x-1])/(DX*DX) + ALPHA*DT*(u[y+1]-2u+u[y-1])/(DY*DY)."""
    xp, xm, yp, ym = _nbrs(u)
    ad = alpha * dt
    s = ad * (xp - 2 * u + xm) / (dx * dx)
    s = s + ad * (yp - 2 * u + ym) / (dy * dy)
    return s


def step_cu(u, ub, gts, alpha, dt, dx, dy, scheme_fe, its):
    """One stage in the order of the device plugin (heat.cu:22-48):
    FE: u + cd;  RK2: u[idx - TIMES*tc] + coeff*cd, final stage when gts%ITS==0."""
    cd = central_difference_cu(u, alpha, dt, dx, dy)
    if scheme_fe:
        return u + cd
    if gts % its == 0:
        return ub + 1.0 * cd
    return u + 0.5 * cd


def grid(n):
    """Periodic unit-square grid, endpoint excluded (heat.py:104-105)."""
    return np.linspace(0.0, 1.0, n, endpoint=False)


def analytical(n, t, alpha, m=2):
    """sin(m*pi*x) sin(m*pi*y) exp(-(m^2+m^2) pi^2 alpha t)  (heat.py:90-113)."""
    x = grid(n)
    X, Y = np.meshgrid(x, x, indexing="ij")
    return np.sin(m * np.pi * X) * np.sin(m * np.pi * Y) * np.exp(-(2 * m * m) * np.pi ** 2 * alpha * t)


def initial_condition(n, noise=1e-3, seed=0, kx=1, ky=1, amp=1.0):
    """u0 = amp*sin(2 pi kx x) sin(2 pi ky y) + noise*U(-1,1) (seeded), shape (1, n, n)."""
    x = grid(n)
    X, Y = np.meshgrid(x, x, indexing="ij")
    u = amp * np.sin(2 * np.pi * kx * X) * np.sin(2 * np.pi * ky * Y)
    if noise:
        rng = np.random.default_rng(seed)
        u = u + noise * rng.uniform(-1.0, 1.0, size=u.shape)
    return u[None, :, :].astype(np.float64)


def fourier_solution(u0, nsteps, courant, scheme_fe):
    """Exact discrete solution of the linear periodic scheme.

    Each Fourier mode (p, q) is multiplied per full step by
    lambda = 1 + z (forward Euler) or 1 + z + z^2/2 (midpoint RK2), where
    z = -4 c [sin^2(pi p / nx) + sin^2(pi q / ny)] and c = alpha dt / dx^2.
    This fixes the scheme independently of any operation order.
    """
    nx, ny = u0.shape
    p = np.arange(nx)[:, None]
    q = np.arange(ny)[None, :]
    z = -4.0 * courant * (np.sin(np.pi * p / nx) ** 2 + np.sin(np.pi * q / ny) ** 2)
    lam = 1.0 + z if scheme_fe else 1.0 + z + 0.5 * z * z
    return np.real(np.fft.ifft2(np.fft.fft2(u0) * lam ** nsteps))
