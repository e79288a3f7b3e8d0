"""Known-answer plugin (CPU half of checker.cu): mean of the four
neighbours.  Contract as the reference's checker plugin
(pysweep/equations/checker.py:11-71)."""
import numpy

from pysweep_amd.plugins import driver as cuda

scheme = True
dt = dx = dy = 0.0


def set_globals(*args, source_mod=None):
    global dt, dx, dy, scheme
    t0, tf, dt, dx, dy, scheme = args
    if source_mod is not None:
        for key, val in (("DT", dt), ("DX", dx), ("DY", dy)):
            ptr, _ = source_mod.get_global(key)
            cuda.memcpy_htod(ptr, numpy.float64(val))
        ptr, _ = source_mod.get_global("SCHEME")
        cuda.memcpy_htod(ptr, numpy.intc(bool(scheme)))


def step(state, iidx, arrayTimeIndex, globalTimeStep):
    ij = numpy.asarray(list(iidx), dtype=numpy.intp).reshape(-1, 2)
    ii, jj = ij[:, 0], ij[:, 1]
    t = arrayTimeIndex
    q = state[t]
    nx, ny = q.shape[1:]
    s = q[:, (ii + 1) % nx, jj]
    s = s + q[:, ii - 1, jj]
    s = s + q[:, ii, (jj + 1) % ny]
    s = s + q[:, ii, jj - 1]
    state[t + 1][:, ii, jj] = s / 4


def createInitialConditions(nv, nx, ny, filename="checkerConditions.hdf5"):
    """A checkerboard of ones on the cells with i + j even, zeros elsewhere,
    shape (nv, nx, ny), in dataset ``data`` of an HDF5 file
    (checker.py:45-57); returns the file name."""
    from pysweep_amd import hdf5io
    i, j = numpy.indices((nx, ny))
    board = ((i + j) % 2 == 0).astype(numpy.float64)
    return hdf5io.write_datasets(filename, {"data": numpy.repeat(board[None], nv, axis=0)})
