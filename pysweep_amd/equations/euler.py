"""2-D Euler plugin (CPU half; the device half is euler.cu).

Same contract as the reference's plugin pair (pysweep/equations/euler.py,
euler.cu): ``set_globals(t0, tf, dt, dx, dy, gamma, source_mod=None)`` and
``step(state, iidx, arrayTimeIndex, globalTimeStep)`` on state[t, 4, X, Y].
Both halves evaluate the numpy plugin's operation order
(pysweep/equations/euler.py:131-313) with squares as a*a.
"""
import numpy

from pysweep_amd.plugins import driver as cuda

gamma = 1.4
gM1 = 0.4
dtdx = dtdy = 0.0


def set_globals(*args, source_mod=None):
    """Set gamma, gamma-1 and dt/dx; upload DT, DX, DY, GAMMA, GAM_M1, DTDX,
    DTDY (float64) to the device module when given.  dtdx = dtdy = dt/dx as
    in the reference's numpy plugin (euler.py:23)."""
    global dtdx, dtdy, gamma, gM1
    t0, tf, dt, dx, dy, gamma = args
    dtdx = dtdy = dt / dx
    gM1 = gamma - 1
    if source_mod is not None:
        vals = (("DT", dt), ("DX", dx), ("DY", dy), ("GAMMA", gamma), ("GAM_M1", gM1),
                ("DTDX", dtdx), ("DTDY", dtdy))
        for key, val in vals:
            ptr, _ = source_mod.get_global(key)
            cuda.memcpy_htod(ptr, numpy.float64(val))


def _pressure(q):
    return gM1 * (q[3] - (q[1] * q[1] + q[2] * q[2]) / (2 * q[0]))


def _limit(q1, q2, num, den):
    with numpy.errstate(all="ignore"):
        pr = num / den
    pr = numpy.where(numpy.isnan(pr) | numpy.isinf(pr) | (pr > 1e6) | (pr < 0), 0.0, pr)
    h = numpy.where(1.0 < pr, 1.0, pr) / 2
    return q1 + h * (q2 - q1)


def _flux(L, R, xdir):
    pl, pr = _pressure(L), _pressure(R)
    a, b = (1, 2) if xdir else (2, 1)
    f = numpy.empty_like(L)
    f[0] = L[a] + R[a]
    f[a] = L[a] * L[a] / L[0] + pl + R[a] * R[a] / R[0] + pr
    f[b] = L[1] * L[2] / L[0] + R[1] * R[2] / R[0]
    f[3] = (L[3] + pl) * L[a] / L[0] + (R[3] + pr) * R[a] / R[0]
    return f


def _dissipation(L, R, xdir):
    sl, sr = numpy.sqrt(L[0]), numpy.sqrt(R[0])
    den = 1 / (sl + sr)
    a0 = sl * sr
    a = [a0] + [(sl * (L[i] / L[0]) + sr * (R[i] / R[0])) * den for i in (1, 2, 3)]
    roe = numpy.stack([a0, a0 * a[1], a0 * a[2], a0 * a[3]])
    rad = numpy.sqrt(gamma * _pressure(roe) / a0) + numpy.abs(a[1] if xdir else a[2])
    rad = numpy.where(numpy.isnan(rad), 0.0, rad)
    return rad * (L - R)


def _divergence(q, P, ii, jj, xdir):
    nx, ny = q.shape[1:]

    def at(d):
        if xdir:
            i = (ii + d) % nx
            return q[:, i, jj], P[i, jj]
        j = (jj + d) % ny
        return q[:, ii, j], P[ii, j]

    (ww, pww), (w, pw), (c, pc), (e, pe), (ee, pee) = at(-2), at(-1), at(0), at(1), at(2)
    acc = numpy.zeros_like(c)
    L, R = _limit(w, c, pc - pw, pw - pww), _limit(c, w, pc - pw, pe - pc)
    acc = acc + _flux(L, R, xdir)
    acc = acc + _dissipation(L, R, xdir)
    L, R = _limit(c, e, pe - pc, pc - pw), _limit(e, c, pe - pc, pee - pe)
    acc = acc - _flux(L, R, xdir)
    acc = acc - _dissipation(L, R, xdir)
    return acc * 0.5


def step(state, iidx, arrayTimeIndex, globalTimeStep):
    """Write state[arrayTimeIndex+1, :] at the points iidx."""
    ij = numpy.asarray(list(iidx), dtype=numpy.intp).reshape(-1, 2)
    ii, jj = ij[:, 0], ij[:, 1]
    t = arrayTimeIndex
    q = state[t]
    with numpy.errstate(all="ignore"):
        P = _pressure(q)
        fx = _divergence(q, P, ii, jj, True)
        fy = _divergence(q, P, ii, jj, False)
        if globalTimeStep % 2 == 0:
            state[t + 1][:, ii, jj] = state[t - 1][:, ii, jj] + (fx * dtdx + fy * dtdy)
        else:
            state[t + 1][:, ii, jj] = state[t][:, ii, jj] + 0.5 * (fx * dtdx + fy * dtdy)


# ------------------------------------------------------------------------
# Initial conditions and the analytic solution (the reference plugin's
# helpers, pysweep/equations/euler.py:33-116): the isentropic vortex of
# strength 5 on [-5, 5]^2, free stream at Mach sqrt(2/gamma) along 45
# degrees, advected without change of shape.

def analytical(x, y, t, gamma=1.4, alpha=numpy.pi / 4):
    """(rho, rho u, rho v, E) of the advected vortex at (x, y, t)
    (euler.py:58-105); scalars or broadcastable arrays.  No periodic wrap
    of the centre, as in the reference."""
    mach = numpy.sqrt(2 / gamma)
    beta = mach * 5 * numpy.sqrt(2) / 4 / numpy.pi * numpy.exp(0.5)
    speed = mach * numpy.sqrt(gamma)  # free-stream speed, rho = p = 1
    ddx = x - speed * numpy.cos(alpha) * t
    ddy = y - speed * numpy.sin(alpha) * t
    om = beta * numpy.exp(-0.5 * (ddx * ddx + ddy * ddy))
    dT = -(gamma - 1) * om * om / 2
    rho = (1 + dT) ** (1 / (gamma - 1))
    u = speed * numpy.cos(alpha) - ddy * om
    v = speed * numpy.sin(alpha) + ddx * om
    p = (1 + dT) ** (gamma / (gamma - 1)) / gamma
    return rho, rho * u, rho * v, p / (gamma - 1) + 0.5 * rho * (u * u + v * v)


def createInitialConditions(npx, npy, t=0, gamma=1.4, filename="eulerConditions.hdf5"):
    """Write the vortex at time t to an HDF5 file with dataset ``data`` of
    shape (4, npx, npy) and return its name (euler.py:33-56).  The grid
    includes both end points (linspace(-5, 5, n)), and the dataset is stored
    as float32, as the reference's h5py default dataset is."""
    from pysweep_amd import hdf5io
    x = numpy.linspace(-5, 5, npx)
    y = numpy.linspace(-5, 5, npy)
    q = numpy.stack(numpy.broadcast_arrays(*analytical(x[:, None], y[None, :], t, gamma=gamma)))
    return hdf5io.write_datasets(filename, {"data": q.astype(numpy.float32)})


def getAnalyticalArray(npx, npy, t):
    """The vortex at time t on the periodic grid -5 + 10 i/npx (endpoints
    excluded), shape (4, npx, npy), with the module's gamma (euler.py:107-116)."""
    x = numpy.linspace(-5, 5, npx, endpoint=False)
    y = numpy.linspace(-5, 5, npy, endpoint=False)
    return numpy.stack(numpy.broadcast_arrays(*analytical(x[:, None], y[None, :], t, gamma=gamma)))


def getFinalTime(npx, timesteps, d=0.1, gamma=1.4):
    """(final time, dt) for dt = d dx, dx = 10/npx (euler.py:58-61)."""
    dx = 10 / npx
    dt = d * dx
    return timesteps * dt, dt
