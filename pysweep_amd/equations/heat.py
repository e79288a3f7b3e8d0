"""Heat equation plugin (CPU half; the device half is heat.cu).

Same contract as the reference's plugin pair (pysweep/equations/heat.py,
heat.cu): ``set_globals(t0, tf, dt, dx, dy, alpha, scheme, source_mod=None)``
and ``step(state, iidx, arrayTimeIndex, globalTimeStep)``; ``scheme`` True
selects forward Euler, False midpoint RK2.
"""
import numpy

from pysweep_amd.plugins import driver as cuda

scheme = True
dt = dx = dy = alpha = courant = 0.0


def set_globals(*args, source_mod=None):
    """Set module globals and, given a source module, the device constants
    DT, DX, DY, ALPHA, COURANT (float64) and SCHEME (int)."""
    global dt, dx, dy, alpha, scheme, courant
    t0, tf, dt, dx, dy, alpha, scheme = args
    courant = alpha * dt / dx / dx
    if source_mod is not None:
        for key, val in (("DT", dt), ("DX", dx), ("DY", dy), ("ALPHA", alpha), ("COURANT", courant)):
            ptr, _ = source_mod.get_global(key)
            cuda.memcpy_htod(ptr, numpy.float64(val))
        ptr, _ = source_mod.get_global("SCHEME")
        cuda.memcpy_htod(ptr, numpy.intc(bool(scheme)))


def _laplace(u, ii, jj):
    nx, ny = u.shape
    c = u[ii, jj]
    lx = courant * (u[(ii + 1) % nx, jj] - 2 * c + u[ii - 1, jj])
    ly = courant * (u[ii, (jj + 1) % ny] - 2 * c + u[ii, jj - 1])
    return lx + ly


def step(state, iidx, arrayTimeIndex, globalTimeStep):
    """Write state[arrayTimeIndex+1, 0] at the points iidx."""
    ij = numpy.asarray(list(iidx), dtype=numpy.intp).reshape(-1, 2)
    ii, jj = ij[:, 0], ij[:, 1]
    t = arrayTimeIndex
    d = _laplace(state[t, 0], ii, jj)
    if scheme:
        state[t + 1, 0, ii, jj] = d + state[t, 0, ii, jj]
    elif globalTimeStep % 2 == 0:
        state[t + 1, 0, ii, jj] = 1 * d + state[t - 1, 0, ii, jj]
    else:
        state[t + 1, 0, ii, jj] = 0.5 * d + state[t, 0, ii, jj]
