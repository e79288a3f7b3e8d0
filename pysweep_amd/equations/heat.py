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


# ------------------------------------------------------------------------
# Initial conditions and the analytic solution (the reference plugin's
# helpers, pysweep/equations/heat.py:66-120): the mode sin(2 pi x) sin(2 pi y)
# of the periodic unit square decays as exp(-8 pi^2 alpha t).

def analyticalEquation(x, y, t, alpha=0.1):
    """u(x, y, t) of the decaying (2, 2) half-wave mode (heat.py:95-98);
    scalars or broadcastable arrays."""
    k = 2 * numpy.pi
    return numpy.sin(k * x) * numpy.sin(k * y) * numpy.exp(-2 * k * k * alpha * t)


def createInitialConditions(npx, npy, alpha=0.1, t=0, filename="heatConditions.hdf5"):
    """Write the mode at time t to an HDF5 file with dataset ``data`` of
    shape (1, npx, npy) and return its name (heat.py:66-93).  The grid
    includes both end points (linspace(0, 1, n)), and the dataset is stored
    as float32, as the reference's h5py default dataset is."""
    from pysweep_amd import hdf5io
    x = numpy.linspace(0, 1, npx)
    y = numpy.linspace(0, 1, npy)
    u = analyticalEquation(x[:, None], y[None, :], t, alpha)
    return hdf5io.write_datasets(filename, {"data": u[None].astype(numpy.float32)})


def analytical(npx, npy, t, alpha=0.1):
    """(u, X, Y): the mode at time t on the periodic grid X = i/npx,
    Y = j/npy (heat.py:100-113); u has shape (1, npx, npy)."""
    X = numpy.linspace(0, 1, npx, endpoint=False)
    Y = numpy.linspace(0, 1, npy, endpoint=False)
    u = analyticalEquation(X[:, None], Y[None, :], t, alpha)
    return u[None], X, Y


def getFinalTime(npx, tsteps, alpha):
    """(final time, dt) for tsteps steps of dt = 0.1 dx^2 / alpha with
    dx = 1/npx.  As in the reference (heat.py:115-120) the alpha argument is
    not used: the time step is computed with alpha = 0.01."""
    dx = 1 / npx
    dt = float(0.1 * dx ** 2 / 0.01)
    return dt * tsteps, dt
