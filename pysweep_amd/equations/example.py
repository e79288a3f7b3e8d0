"""Known-answer plugin (CPU half of example.cu): +1 per stage, or +1/+2 in
an RK2 shape.  Contract as the reference's example plugin
(pysweep/equations/example.py:13-47)."""
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
    if scheme:
        state[t + 1][:, ii, jj] = state[t][:, ii, jj] + 1
    elif globalTimeStep % 2 == 0:
        state[t + 1][:, ii, jj] = state[t - 1][:, ii, jj] + 2
    else:
        state[t + 1][:, ii, jj] = state[t][:, ii, jj] + 1


def createInitialConditions(nv, nx, ny, filename="exampleConditions.hdf5"):
    """Zeros of shape (nv, nx, ny) in dataset ``data`` of an HDF5 file
    (example.py:29-34); returns the file name."""
    from pysweep_amd import hdf5io
    return hdf5io.write_datasets(filename, {"data": numpy.zeros((nv, nx, ny))})
