"""HDF5 input/output in the reference's layout (core/io.py:78-90).

Output datasets: blocksize (1,), share (1,), exid (len), globals (len),
clocktime (1,), data = (nt, nv, nx, ny) of the solver dtype.  Initial
conditions are read from dataset ``data`` (core/solver.py:89-101).  h5py
is optional: without it a .npz file with the same dataset names is written
next to the requested path.
"""
import os

import numpy

try:
    import h5py
except ImportError:  # not installed in every environment
    h5py = None


def read_initial_conditions(path):
    if path.endswith(".npz"):
        with numpy.load(path) as f:
            return numpy.array(f["data"])
    if h5py is None:
        raise RuntimeError("reading HDF5 initial conditions needs h5py")
    with h5py.File(path, "r") as f:
        return numpy.array(f["data"])


def output_datasets(solver):
    return {
        "blocksize": numpy.array([solver.blocksize[0]]),
        "share": numpy.array([solver.share], dtype=numpy.float64),
        "exid": numpy.array(solver.exid),
        "globals": numpy.array(solver.globals, dtype=numpy.float64),
        "clocktime": numpy.array([getattr(solver, "runtime", 0.0)], dtype=numpy.float64),
        "data": solver.data.astype(solver.dtype, copy=False),
    }


def write_output(solver, path):
    ds = output_datasets(solver)
    if h5py is None:
        npz = os.path.splitext(path)[0] + ".npz"
        numpy.savez(npz, **ds)
        return npz
    with h5py.File(path, "w") as f:
        for k, v in ds.items():
            f.create_dataset(k, data=v)
    return path
