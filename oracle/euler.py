"""2-D Euler oracle (test infrastructure only; see oracle/__init__.py).

State arrays are ``q[v, x, y]`` with v = (rho, rho*u, rho*v, E); axis 1 is
x, axis 2 is y, both periodic.  Two restatements of the same scheme are
kept, each following one of the reference's statements operation by
operation:

* ``step_py``: pysweep/equations/euler.py:131-313 (the numpy CPU plugin,
  which is the parity target);
* ``step_cu``: pysweep/equations/euler.cu:22-244 (the device plugin, whose
  arithmetic the GPU executes; the GPU is expected to match it bitwise).
"""
import math

import numpy as np

_libm_pow = np.frompyfunc(math.pow, 2, 1)


def sq_mul(a):
    """a*a, correctly rounded."""
    return a * a


def sq_pow(a):
    """a**2 as the reference evaluates it on numpy float64 scalars
    (evaluateFluxInX / evaluateFluxInY, euler.py:194, 239): the C library's
    pow(a, 2.0), which is not always the correctly rounded a*a."""
    return _libm_pow(a, 2.0).astype(np.float64)


def _shift(a, d, axis):
    """a[i + d] along ``axis`` with periodic wrap (axis counted for 3-D arrays;
    2-D arrays such as pressures use axis-1)."""
    ax = axis if a.ndim == 3 else axis - 1
    return np.roll(a, -d, axis=ax)


# --------------------------------------------------------------------------
# numpy-plugin order (euler.py)
# --------------------------------------------------------------------------

def pressure_py(q, gamma):
    """getPressure / getPressureArray: gM1*(q3-(q1*q1+q2*q2)/(2*q0))
    (euler.py:284-293, 295-313)."""
    gm1 = gamma - 1
    return gm1 * (q[3] - (q[1] * q[1] + q[2] * q[2]) / (2 * q[0]))


def limiter_py(q1, q2, num, den):
    """fluxLimiter (euler.py:250-256): Pr=num/den, zeroed when nan, inf,
    >1e6 or <0; returns q1 + min(Pr,1)/2*(q2-q1)."""
    with np.errstate(all="ignore"):
        pr = num / den
        bad = np.isnan(pr) | np.isinf(pr) | (pr > 1e6) | (pr < 0)
        pr = np.where(bad, 0.0, pr)
        m = np.minimum(pr, 1.0)
        return q1 + (m / 2) * (q2 - q1)


def flux_x_py(L, R, gamma, sq=sq_mul):
    """evaluateFluxInX (euler.py:181-197)."""
    PL = pressure_py(L, gamma)
    PR = pressure_py(R, gamma)
    f = np.empty_like(L)
    f[0] = L[1] + R[1]
    f[1] = sq(L[1]) / L[0] + PL + sq(R[1]) / R[0] + PR
    f[2] = L[1] * L[2] / L[0] + R[1] * R[2] / R[0]
    f[3] = (L[3] + PL) * L[1] / L[0] + (R[3] + PR) * R[1] / R[0]
    return f


def flux_y_py(L, R, gamma, sq=sq_mul):
    """evaluateFluxInY (euler.py:225-241)."""
    PL = pressure_py(L, gamma)
    PR = pressure_py(R, gamma)
    f = np.empty_like(L)
    f[0] = L[2] + R[2]
    f[1] = L[1] * L[2] / L[0] + R[1] * R[2] / R[0]
    f[2] = sq(L[2]) / L[0] + PL + sq(R[2]) / R[0] + PR
    f[3] = (L[3] + PL) * L[2] / L[0] + (R[3] + PR) * R[2] / R[0]
    return f


def spectral_py(L, R, dim, gamma):
    """evaluateSpectral (euler.py:258-282): Roe-averaged spectral radius
    times (L - R); the radius is zeroed when it is nan."""
    with np.errstate(all="ignore"):
        rrL = np.sqrt(L[0])
        rrR = np.sqrt(R[0])
        tL = L / L[0]
        tR = R / R[0]
        denom = 1 / (rrL + rrR)
        s0 = rrL * rrR
        s = [s0] + [(rrL * tL[i] + rrR * tR[i]) * denom for i in (1, 2, 3)]
        spvec = np.stack([s[0], s[0] * s[1], s[0] * s[2], s[0] * s[3]])
        P = pressure_py(spvec, gamma)
        sr = np.sqrt(gamma * P / s[0]) + np.abs(s[dim])
        sr = np.where(np.isnan(sr), 0.0, sr)
        return sr * (L - R)


def _dir_flux_py(q, P, axis, gamma, sq=sq_mul):
    """getFluxInX / getFluxInY (euler.py:158-179, 202-223): west face added,
    east face subtracted, each as flux then spectral term, then *0.5."""
    ww, w = _shift(q, -2, axis), _shift(q, -1, axis)
    e, ee = _shift(q, 1, axis), _shift(q, 2, axis)
    Pww, Pw = _shift(P, -2, axis), _shift(P, -1, axis)
    Pe, Pee = _shift(P, 1, axis), _shift(P, 2, axis)
    fl = flux_x_py if axis == 1 else flux_y_py
    dim = 1 if axis == 1 else 2
    c = q
    flux = np.zeros_like(q)
    left = limiter_py(w, c, P - Pw, Pw - Pww)
    right = limiter_py(c, w, P - Pw, Pe - P)
    with np.errstate(all="ignore"):
        flux += fl(left, right, gamma, sq)
        flux += spectral_py(left, right, dim, gamma)
    left = limiter_py(c, e, Pe - P, P - Pw)
    right = limiter_py(e, c, Pe - P, Pee - Pe)
    with np.errstate(all="ignore"):
        flux -= fl(left, right, gamma, sq)
        flux -= spectral_py(left, right, dim, gamma)
    return flux * 0.5


def step_py(q, qb, gts, gamma, dt, dx, dy, sq=sq_mul):
    """step (euler.py:131-151) on the whole periodic grid.

    dtdx = dtdy = dt/dx (euler.py:23).  Final RK2 stage when gts%2==0:
    qb + (fx*dtdx + fy*dtdy); intermediate: q + 0.5*(fx*dtdx + fy*dtdy).
    ``sq`` evaluates the squares of evaluateFluxInX/Y: ``sq_mul`` (a*a, the
    device plugin's choice) or ``sq_pow`` (the C library's pow, as the
    reference's numpy scalars do).
    """
    dtdx = dtdy = dt / dx
    with np.errstate(all="ignore"):
        P = pressure_py(q, gamma)
        fx = _dir_flux_py(q, P, 1, gamma, sq)
        fy = _dir_flux_py(q, P, 2, gamma, sq)
        if gts % 2 == 0:
            return qb + (fx * dtdx + fy * dtdy)
        return q + 0.5 * (fx * dtdx + fy * dtdy)


# --------------------------------------------------------------------------
# device-plugin order (euler.cu)
# --------------------------------------------------------------------------

def pressure_cu(p, gm1):
    """pressure (euler.cu:34-39): vss=(p1*p1+p2*p2)/p0; GAM_M1*(p3-0.5*vss)."""
    vss = (p[1] * p[1] + p[2] * p[2]) / p[0]
    return gm1 * (p[3] - 0.5 * vss)


def spec_press_cu(s, gm1):
    """spec_press (euler.cu:43-49): GAM_M1*(rho*p3-0.5*rho*vss)."""
    rho = s[0]
    vss = s[1] * s[1] + s[2] * s[2]
    return gm1 * (rho * s[3] - 0.5 * rho * vss)


def limiter_cu(p1, p2, p3, num, den):
    """flimiter (euler.cu:127-138): p1 + 0.5*Pr*(p2-p3), Pr=min(Pr,1) after
    zeroing nan/inf/>1e6/<0."""
    with np.errstate(all="ignore"):
        pr = num / den
        bad = np.isnan(pr) | np.isinf(pr) | (pr > 1e6) | (pr < 0.0)
        pr = np.where(bad, 0.0, pr)
        pr = np.minimum(pr, 1.0)
        return p1 + 0.5 * pr * (p2 - p3)


def eflux_cu(L, R, axis, gm1):
    """efluxx / efluxy (euler.cu:83-121) with dir=+1: the face flux vector."""
    rhoL, rhoR = L[0], R[0]
    uL, uR = L[1] / rhoL, R[1] / rhoR
    vL, vR = L[2] / rhoL, R[2] / rhoR
    pL, pR = pressure_cu(L, gm1), pressure_cu(R, gm1)
    f = np.empty_like(L)
    if axis == 1:
        f[0] = L[1] + R[1]
        f[1] = L[1] * uL + pL + R[1] * uR + pR
        f[2] = L[1] * vL + R[1] * vR
        f[3] = (L[3] + pL) * uL + (R[3] + pR) * uR
    else:
        f[0] = L[2] + R[2]
        f[1] = L[2] * uL + R[2] * uR
        f[2] = L[2] * vL + pL + R[2] * vR + pR
        f[3] = (L[3] + pL) * vL + (R[3] + pR) * vR
    return f


def espectral_cu(L, R, dim, gamma, gm1):
    """espectral (euler.cu:54-78) with dir=+1: rs*(L-R)."""
    rrL = np.sqrt(L[0])
    rrR = np.sqrt(R[0])
    s0 = rrL * rrR
    denom = 1 / (rrL + rrR)
    s = [s0]
    for i in (1, 2, 3):
        si = 0.0 + rrL * L[i] / L[0]
        si = si + rrR * R[i] / R[0]
        si = si * denom
        s.append(si)
    rs = np.sqrt(gamma * spec_press_cu(s, gm1) / s[0]) + np.abs(s[dim])
    rs = np.where(np.isnan(rs), 0.0, rs)
    return rs * (L - R)


def _dir_flux_cu(q, axis, gamma, gm1):
    """get_dfdx / get_dfdy (euler.cu:141-227)."""
    ww, w = _shift(q, -2, axis), _shift(q, -1, axis)
    e, ee = _shift(q, 1, axis), _shift(q, 2, axis)
    c = q
    P0, P1, P2 = pressure_cu(ww, gm1), pressure_cu(w, gm1), pressure_cu(c, gm1)
    P3, P4 = pressure_cu(e, gm1), pressure_cu(ee, gm1)
    dim = 1 if axis == 1 else 2
    acc = np.zeros_like(q)
    L = limiter_cu(w, c, w, P2 - P1, P1 - P0)
    R = limiter_cu(c, w, c, P2 - P1, P3 - P2)
    acc = acc + 1.0 * eflux_cu(L, R, axis, gm1)
    acc = acc + 1.0 * espectral_cu(L, R, dim, gamma, gm1)
    L = limiter_cu(c, e, c, P3 - P2, P2 - P1)
    R = limiter_cu(e, c, e, P3 - P2, P4 - P3)
    acc = acc + -1.0 * eflux_cu(L, R, axis, gm1)
    acc = acc + -1.0 * espectral_cu(L, R, dim, gamma, gm1)
    return acc / 2


def step_cu(q, qb, gts, gamma, dt, dx, dy, its=2):
    """step (euler.cu:229-244): final stage when gts%ITS==0:
    q[t-1] + 1*(DTDX*dfdx + DTDY*dfdy); else q[t] + 0.5*(...)."""
    gm1 = gamma - 1
    dtdx, dtdy = dt / dx, dt / dy
    with np.errstate(all="ignore"):
        dfdx = _dir_flux_cu(q, 1, gamma, gm1)
        dfdy = _dir_flux_cu(q, 2, gamma, gm1)
        if gts % its == 0:
            return qb + 1.0 * (dtdx * dfdx + dtdy * dfdy)
        return q + 0.5 * (dtdx * dfdx + dtdy * dfdy)


# --------------------------------------------------------------------------
# initial condition / exact solution
# --------------------------------------------------------------------------

def grid(n, L=5.0):
    """Cell coordinates on [-L, L), endpoint excluded (euler.py:111-112)."""
    return np.linspace(-L, L, n, endpoint=False)


def vortex(x, y, t, gamma=1.4, eps=5.0, xc=0.0, yc=0.0, alpha=np.pi / 4):
    """Isentropic vortex advected by the free stream (euler.py:63-105).

    M_inf = sqrt(2/gamma), beta = M_inf*eps*sqrt(2)/(4 pi)*e^0.5,
    V_inf = M_inf*sqrt(gamma) so (u0, v0) = V_inf (cos a, sin a); the
    pressure carries the factor 1/gamma of euler.py:104.
    """
    minf = np.sqrt(2 / gamma)
    beta = minf * eps * np.sqrt(2) / 4 / np.pi * np.exp(0.5)
    c = np.sqrt(gamma * 1.0 / 1.0)
    vinf = minf * c
    u0, v0 = vinf * np.cos(alpha), vinf * np.sin(alpha)
    dx0 = x - xc - u0 * t
    dy0 = y - yc - v0 * t
    f = -0.5 * (dx0 ** 2 + dy0 ** 2)
    om = beta * np.exp(f)
    du, dv = -dy0 * om, dx0 * om
    dT = -(gamma - 1) * om * om / 2
    rho = (1 + dT) ** (1 / (gamma - 1))
    u, v = u0 + du, v0 + dv
    P = 1 / gamma * (1 + dT) ** (gamma / (gamma - 1))
    return np.stack([rho, rho * u, rho * v, P / (gamma - 1) + 0.5 * rho * (u * u + v * v)])


def initial_condition(n, gamma=1.4, eps=5.0, xc=0.0, yc=0.0, t=0.0, ny=None, L=5.0):
    """Vortex on the periodic grid, shape (4, n, ny) float64."""
    ny = n if ny is None else ny
    X, Y = np.meshgrid(grid(n, L), grid(ny, L), indexing="ij")
    return vortex(X, Y, t, gamma, eps, xc, yc).astype(np.float64)
