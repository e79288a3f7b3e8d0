"""Known-answer plugin (CPU half of half.cu): midpoint RK2 heat with the
coefficient 1/100.  Contract as the reference's half plugin
(pysweep/equations/half.py:17-52)."""
import numpy

from pysweep_amd.plugins import driver as cuda

dt = dx = dy = 0.0


def set_globals(*args, source_mod=None):
    global dt, dx, dy
    t0, tf, dt, dx, dy = args
    if source_mod is not None:
        for key, val in (("DT", dt), ("DX", dx), ("DY", dy)):
            ptr, _ = source_mod.get_global(key)
            cuda.memcpy_htod(ptr, numpy.float64(val))


def step(state, iidx, arrayTimeIndex, globalTimeStep):
    ij = numpy.asarray(list(iidx), dtype=numpy.intp).reshape(-1, 2)
    ii, jj = ij[:, 0], ij[:, 1]
    t = arrayTimeIndex
    u = state[t, 0]
    nx, ny = u.shape
    c = u[ii, jj]
    d = (u[(ii + 1) % nx, jj] - 2 * c + u[ii - 1, jj]) / 100
    d = d + (u[ii, (jj + 1) % ny] - 2 * c + u[ii, jj - 1]) / 100
    if globalTimeStep % 2 == 0:
        state[t + 1, 0, ii, jj] = state[t - 1, 0, ii, jj] + 1 * d
    else:
        state[t + 1, 0, ii, jj] = state[t, 0, ii, jj] + 0.5 * d


def createInitialConditions(npx, npy, t=0, filename="halfConditions.hdf5"):
    """u = (flat index of the cell)/100 on an (npx, npy) grid, shape
    (1, npx, npy), in dataset ``data`` of an HDF5 file (half.py:54-69);
    returns the file name."""
    from pysweep_amd import hdf5io
    u = numpy.arange(npx * npy, dtype=numpy.float64).reshape(1, npx, npy) / 100
    return hdf5io.write_datasets(filename, {"data": u})
