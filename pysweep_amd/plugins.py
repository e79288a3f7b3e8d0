"""Plugin loading and the device facade handed to plugins.

A plugin pair is a Python module with ``set_globals(*args, source_mod=None)``
and ``step(state, iidx, arrayTimeIndex, globalTimeStep)`` plus a device source
with ``__device__ void step(double*, int, int)`` and its own ``__constant__``
globals (the reference's contract, README.md:45-112).  The reference's
``set_globals`` writes device constants through PyCUDA:
``ptr, _ = source_mod.get_global(name); cuda.memcpy_htod(ptr, value)``
(pysweep/equations/heat.py:38-46).  ``SourceModuleFacade`` and ``driver``
provide those two calls on top of libpysweep_hip; ``load_cpu_module`` injects
``driver`` as the module-level name ``cuda`` when the module did not bind one
(the reference's plugins import pycuda inside try/except and run without it).
"""
import importlib.util
import os
import sys

import numpy


class GlobalHandle:
    """A device global that may exist in several compiled modules (the swept
    and standard kernels are separate compilations of the same plugin
    source); writes go to every copy."""

    def __init__(self, name, copies):
        self.name = name
        self.copies = copies  # list of (device pointer, size in bytes)

    @property
    def nbytes(self):
        return min(n for _, n in self.copies)

    def __int__(self):
        return int(self.copies[0][0])

    def __repr__(self):
        return f"GlobalHandle({self.name!r}, {len(self.copies)} copies)"


class HostField:
    """A constant of one problem's host-side constant block (instance-table
    modules); ``memcpy_htod`` writes into the block, which is uploaded when
    the problem is admitted."""

    def __init__(self, buf, offset, size, name):
        self.buf, self.offset, self.size, self.name = buf, offset, size, name

    def __repr__(self):
        return f"HostField({self.name!r}, offset={self.offset}, size={self.size})"


def _as_bytes(src):
    if isinstance(src, (bytes, bytearray, memoryview)):
        return bytes(src)
    arr = numpy.asarray(src)
    if arr.dtype == object:
        raise TypeError("memcpy_htod needs a numpy scalar/array or bytes")
    return numpy.ascontiguousarray(arr).tobytes()


def _hip():
    from pysweep_amd import hip
    return hip


class _Driver:
    """The subset of ``pycuda.driver`` the reference's plugins call."""

    def memcpy_htod(self, dst, src):
        data = _as_bytes(src)
        if isinstance(dst, HostField):
            if len(data) > dst.size:
                raise ValueError(f"{len(data)} bytes do not fit device constant {dst.name!r} ({dst.size} bytes)")
            dst.buf[dst.offset:dst.offset + len(data)] = data
            return
        if isinstance(dst, GlobalHandle):
            for ptr, n in dst.copies:
                if len(data) > n:
                    raise ValueError(f"{len(data)} bytes do not fit device global {dst.name!r} ({n} bytes)")
                _hip().htod_bytes(ptr, data)
        else:
            _hip().htod_bytes(int(dst), data)

    def memcpy_dtoh(self, dst, src):
        from pysweep_amd import hip
        out = numpy.asarray(dst)
        ptr = int(src) if not isinstance(src, GlobalHandle) else src.copies[0][0]
        import ctypes
        hip.check(hip._dtoh(out.ctypes.data, ptr, out.nbytes), "ps_memcpy_dtoh")
        return out

    class Context:
        @staticmethod
        def synchronize():
            from pysweep_amd import hip
            hip.synchronize()


driver = _Driver()


class SourceModuleFacade:
    """What a plugin's ``set_globals(..., source_mod=...)`` receives: the
    compiled device modules of one plugin source."""

    def __init__(self, modules):
        self.modules = list(modules)

    def get_global(self, name):
        copies = [m.get_global(name) for m in self.modules]
        for m in self.modules:
            # a plugin writes its constants through this handle: whoever
            # recorded the module's constants as theirs no longer owns them
            m.const_owner = None
        h = GlobalHandle(name, copies)
        return h, h.nbytes

    def get_function(self, name):
        for m in self.modules:
            try:
                return m.get_function(name)
            except Exception:
                continue
        raise KeyError(name)


def load_cpu_module(path, name=None):
    """Import a plugin's Python half from a file path (the reference's
    Solver.loadCPUModule, pysweep/core/solver.py:83-87) and provide ``cuda``."""
    path = os.path.abspath(path)
    name = name or f"pysweep_plugin_{os.path.splitext(os.path.basename(path))[0]}_{abs(hash(path))}"
    spec = importlib.util.spec_from_file_location(name, path)
    if spec is None or spec.loader is None:
        raise ImportError(f"cannot load plugin module {path}")
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    for fn in ("set_globals", "step"):
        if not callable(getattr(mod, fn, None)):
            raise AttributeError(f"plugin {path} does not define {fn}()")
    if not hasattr(mod, "cuda"):
        mod.cuda = driver
    return mod
