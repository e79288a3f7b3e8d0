"""HDF5 input/output in the reference's layout (core/io.py:78-90).

Output datasets: blocksize (1,) int64, share (1,) float64, exid (len)
float64, globals (len) float64, clocktime (1,) float64, data = (nt, nv, nx,
ny) of the solver dtype.  Initial conditions are read from dataset ``data``
(core/solver.py:89-101).

Writers, in order of preference: the HDF5 C library through
libpysweep_h5.so (built by pysweep_amd.build where the library is
installed), h5py, and -- when neither is available -- a .npz file with the
same dataset names next to the requested path.
"""
import ctypes
import os

import numpy

try:
    import h5py
except ImportError:  # not installed in every environment
    h5py = None

_H5LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libpysweep_h5.so")
_lib = None


def native():
    """The native HDF5 helper, or None when it was not built."""
    global _lib
    if _lib is None and os.path.exists(_H5LIB):
        lib = ctypes.CDLL(_H5LIB)
        lib.ps_h5_error.restype = ctypes.c_char_p
        lib.ps_h5_write.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p),
                                    ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                    ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_void_p)]
        lib.ps_h5_dims.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(ctypes.c_longlong)]
        lib.ps_h5_read_f64.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_void_p]
        _lib = lib
    return _lib


_TYPES = {numpy.dtype(numpy.float64): 0, numpy.dtype(numpy.float32): 1, numpy.dtype(numpy.int64): 2}


def write_datasets(path, datasets):
    """Write {name: array} (float64, float32 or int64) to a new HDF5 file."""
    lib = native()
    arrays = [(k, numpy.ascontiguousarray(v)) for k, v in datasets.items()]
    if lib is not None:
        n = len(arrays)
        names = (ctypes.c_char_p * n)(*[k.encode() for k, _ in arrays])
        types = (ctypes.c_int * n)(*[_TYPES[a.dtype] for _, a in arrays])
        ranks = (ctypes.c_int * n)(*[a.ndim for _, a in arrays])
        dims = [d for _, a in arrays for d in a.shape]
        cdims = (ctypes.c_longlong * max(1, len(dims)))(*dims)
        data = (ctypes.c_void_p * n)(*[a.ctypes.data for _, a in arrays])
        if lib.ps_h5_write(path.encode(), n, names, types, ranks, cdims, data) != 0:
            raise OSError(lib.ps_h5_error().decode())
        return path
    if h5py is not None:
        with h5py.File(path, "w") as f:
            for k, a in arrays:
                f.create_dataset(k, data=a)
        return path
    npz = os.path.splitext(path)[0] + ".npz"
    numpy.savez(npz, **dict(arrays))
    return npz


def read_dataset(path, name="data"):
    """Dataset ``name`` of an HDF5 (or .npz) file as float64."""
    if path.endswith(".npz"):
        with numpy.load(path) as f:
            return numpy.array(f[name], dtype=numpy.float64)
    lib = native()
    if lib is not None:
        rank = ctypes.c_int(0)
        dims = (ctypes.c_longlong * 8)()
        if lib.ps_h5_dims(path.encode(), name.encode(), ctypes.byref(rank), dims) != 0:
            raise OSError(lib.ps_h5_error().decode())
        out = numpy.empty(tuple(dims[i] for i in range(rank.value)), dtype=numpy.float64)
        if lib.ps_h5_read_f64(path.encode(), name.encode(), out.ctypes.data) != 0:
            raise OSError(lib.ps_h5_error().decode())
        return out
    if h5py is None:
        raise RuntimeError("reading HDF5 needs the HDF5 C library (pysweep_amd.build) or h5py")
    with h5py.File(path, "r") as f:
        return numpy.array(f[name], dtype=numpy.float64)


def read_initial_conditions(path):
    return read_dataset(path, "data")


def output_datasets(solver):
    return {
        "blocksize": numpy.array([solver.blocksize[0]], dtype=numpy.int64),
        "share": numpy.array([solver.share], dtype=numpy.float64),
        "exid": numpy.array(solver.exid, dtype=numpy.float64),
        "globals": numpy.array(solver.globals, dtype=numpy.float64),
        "clocktime": numpy.array([getattr(solver, "runtime", 0.0)], dtype=numpy.float64),
        "data": numpy.ascontiguousarray(solver.data, dtype=solver.dtype),
    }


def write_output(solver, path):
    return write_datasets(path, output_datasets(solver))
