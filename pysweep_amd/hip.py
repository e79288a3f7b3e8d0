"""ctypes binding of libpysweep_hip.so (C ABI in include/pysweep_hip.h).

This is the only module that talks to the native library.  It raises
``ImportError`` at import time when the library has not been built, and
``HipError`` whenever an entry point reports a failure; there is no CPU
fallback.
"""
import ctypes
import os

import numpy

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "_lib")
LIBPATH = os.path.join(LIBDIR, "libpysweep_hip.so")


class HipError(RuntimeError):
    pass


if not os.path.exists(LIBPATH):
    raise ImportError(f"{LIBPATH} is missing: build it with __graft_entry__.build() "
                      "or python -m pysweep_amd.build")

_lib = ctypes.CDLL(LIBPATH)

c_int, c_size_t, c_void_p, c_char_p = ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_char_p


class Grid(ctypes.Structure):
    _fields_ = [("nx", c_int), ("ny", c_int), ("T", c_int), ("xoff", c_int),
                ("tiles_x", c_int), ("tiles_y", c_int), ("keep_from", c_int)]


class SweptPlan(ctypes.Structure):
    _fields_ = [("state", c_void_p), ("nx", c_int), ("ny", c_int), ("b", c_int), ("ops", c_int),
                ("its", c_int), ("T", c_int), ("threads", c_int), ("store_all", c_int),
                ("tile_rows", c_int), ("up", c_void_p), ("ybridge", c_void_p), ("xbridge", c_void_p),
                ("octahedron", c_void_p), ("down", c_void_p)]


class StandardPlan(ctypes.Structure):
    _fields_ = [("state", c_void_p), ("nx", c_int), ("ny", c_int), ("its", c_int), ("T", c_int),
                ("bx", c_int), ("by", c_int), ("threads", c_int), ("standard", c_void_p)]


class BandPlan(ctypes.Structure):
    _fields_ = [("state", c_void_p), ("nx", c_int), ("ny", c_int), ("its", c_int), ("T", c_int),
                ("levels", c_int), ("wout", c_int), ("rows", c_int), ("store_all", c_int), ("band", c_void_p)]


class Inst(ctypes.Structure):
    """ps_inst: one problem of an instance-table launch."""
    _fields_ = [("ring", c_void_p), ("consts", c_void_p), ("nx", c_int), ("ny", c_int), ("T", c_int),
                ("tiles_x", c_int), ("tiles_y", c_int), ("stiles_x", c_int), ("stiles_y", c_int),
                ("G", c_int), ("tail", c_int), ("c", c_int), ("active", c_int),
                ("brows", c_int), ("bands", c_int), ("bsegs", c_int)]


class MultiPlan(ctypes.Structure):
    _fields_ = [("insts", c_void_p), ("prefix", c_void_p), ("sprefix", c_void_p), ("owner", c_void_p),
                ("sowner", c_void_p), ("ninst", c_int),
                ("total_tiles", c_int), ("total_stiles", c_int),
                ("threads", c_int), ("sthreads", c_int),
                ("up", c_void_p), ("ybridge", c_void_p), ("xbridge", c_void_p), ("octahedron", c_void_p),
                ("down", c_void_p), ("standard", c_void_p), ("advance", c_void_p),
                ("bprefix", c_void_p), ("bowner", c_void_p), ("total_bwaves", c_int), ("band", c_void_p)]


class RunStats(ctypes.Structure):
    """Per-kernel launch time totals (index: KERNEL_KINDS) of the timed
    launches: every launch by default, or one in ``every`` launches of the
    kernels named in ``kinds``."""
    _fields_ = [("ms", ctypes.c_double * 7), ("launches", ctypes.c_longlong * 7), ("kinds", c_int),
                ("every", c_int)]

    @classmethod
    def sampled(cls, kinds=(), every=1):
        st = cls()
        st.kinds = sum(1 << KERNEL_KINDS.index(k) for k in kinds)
        st.every = every
        return st

    def as_dict(self):
        return {k: {"ms": self.ms[i], "launches": self.launches[i],
                    "avg_ms": self.ms[i] / self.launches[i] if self.launches[i] else None}
                for i, k in enumerate(KERNEL_KINDS)}


KERNEL_KINDS = ("UpPyramid", "YBridge", "XBridge", "Octahedron", "DownPyramid", "Standard", "Band")


def _sig(name, *argtypes, restype=c_int):
    f = getattr(_lib, name)
    f.argtypes = list(argtypes)
    f.restype = restype
    return f


_lib.ps_last_error.restype = c_char_p
_lib.ps_last_error.argtypes = []
_version = _sig("ps_version")
_device_count = _sig("ps_device_count", ctypes.POINTER(c_int))
_init = _sig("ps_init", c_int)
_shutdown = _sig("ps_shutdown")
_ctx_create = _sig("ps_context_create", c_int, ctypes.POINTER(c_void_p))
_ctx_destroy = _sig("ps_context_destroy", c_void_p)
_ctx_push = _sig("ps_context_push", c_void_p)
_ctx_pop = _sig("ps_context_pop")
_ctx_current = _sig("ps_context_current", ctypes.POINTER(c_void_p))
_ctx_device = _sig("ps_context_device", ctypes.POINTER(c_int))
_compile = _sig("ps_compile", c_char_p, c_char_p, ctypes.POINTER(c_char_p), c_int,
                ctypes.POINTER(c_void_p), ctypes.POINTER(c_size_t), ctypes.POINTER(c_void_p))
_free_host = _sig("ps_free_host", c_void_p, restype=None)
_compiler_version = _sig("ps_compiler_version", ctypes.POINTER(c_int), ctypes.POINTER(c_int))
_module_load = _sig("ps_module_load", c_void_p, c_size_t, ctypes.POINTER(c_void_p))
_module_unload = _sig("ps_module_unload", c_void_p)
_get_global = _sig("ps_module_get_global", c_void_p, c_char_p, ctypes.POINTER(c_void_p), ctypes.POINTER(c_size_t))
_get_function = _sig("ps_module_get_function", c_void_p, c_char_p, ctypes.POINTER(c_void_p))
_function_info = _sig("ps_function_info", c_void_p, ctypes.POINTER(c_int), ctypes.POINTER(c_int),
                      ctypes.POINTER(c_int))
_malloc = _sig("ps_malloc", ctypes.POINTER(c_void_p), c_size_t)
_free = _sig("ps_free", c_void_p)
_htod = _sig("ps_memcpy_htod", c_void_p, c_void_p, c_size_t)
_dtoh = _sig("ps_memcpy_dtoh", c_void_p, c_void_p, c_size_t)
_dtod = _sig("ps_memcpy_dtod", c_void_p, c_void_p, c_size_t)
_memset = _sig("ps_memset", c_void_p, c_int, c_size_t)
_host_malloc = _sig("ps_host_malloc", ctypes.POINTER(c_void_p), c_size_t)
_host_free = _sig("ps_host_free", c_void_p)
_htod_async = _sig("ps_memcpy_htod_async", c_void_p, c_void_p, c_size_t)
_dtoh_async = _sig("ps_memcpy_dtoh_async", c_void_p, c_void_p, c_size_t)
_dtod_async = _sig("ps_memcpy_dtod_async", c_void_p, c_void_p, c_size_t)
_launch = _sig("ps_launch", c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
               ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.POINTER(c_void_p))
_sync = _sig("ps_synchronize")
_event_create = _sig("ps_event_create", ctypes.POINTER(c_void_p))
_event_record = _sig("ps_event_record", c_void_p)
_event_elapsed = _sig("ps_event_elapsed_ms", c_void_p, c_void_p, ctypes.POINTER(ctypes.c_float))
_event_destroy = _sig("ps_event_destroy", c_void_p)
_stream_wait = _sig("ps_stream_wait_event", c_void_p)
_event_sync = _sig("ps_event_synchronize", c_void_p)
_event_query = _sig("ps_event_query", c_void_p, ctypes.POINTER(c_int))
_swept_run = _sig("ps_swept_run", ctypes.POINTER(SweptPlan), c_int, c_int, c_int)
_standard_run = _sig("ps_standard_run", ctypes.POINTER(StandardPlan), c_int, c_int)
_swept_run_timed = _sig("ps_swept_run_timed", ctypes.POINTER(SweptPlan), c_int, c_int, c_int,
                        ctypes.POINTER(RunStats))
_standard_run_timed = _sig("ps_standard_run_timed", ctypes.POINTER(StandardPlan), c_int, c_int,
                           ctypes.POINTER(RunStats))
_band_run = _sig("ps_band_run", ctypes.POINTER(BandPlan), c_int, c_int)
_band_run_timed = _sig("ps_band_run_timed", ctypes.POINTER(BandPlan), c_int, c_int, ctypes.POINTER(RunStats))
_copy_f64 = _sig("ps_copy_f64", c_void_p, c_void_p, c_size_t)
_level_sums = _sig("ps_level_sums", c_void_p, c_int, c_int, c_size_t, c_void_p)
_level_sums_noalloc = _sig("ps_level_sums_noalloc", c_void_p, c_int, c_int, c_size_t, c_void_p)
_multi_run = _sig("ps_multi_run", ctypes.POINTER(MultiPlan), c_int)
_multi_index = _sig("ps_multi_index", ctypes.POINTER(MultiPlan))
_multi_run_timed = _sig("ps_multi_run_timed", ctypes.POINTER(MultiPlan), c_int, ctypes.POINTER(RunStats))
_graph_begin = _sig("ps_graph_begin")
_graph_end = _sig("ps_graph_end", ctypes.POINTER(c_void_p))
_graph_launch = _sig("ps_graph_launch", c_void_p)
_graph_destroy = _sig("ps_graph_destroy", c_void_p)

#: every symbol the header declares (checked by the CPU test suite)
EXPORTS = ("ps_version", "ps_last_error", "ps_device_count", "ps_init", "ps_shutdown", "ps_context_create",
           "ps_context_destroy", "ps_context_push", "ps_context_pop", "ps_context_current", "ps_context_device",
           "ps_compile", "ps_compiler_version",
           "ps_free_host", "ps_module_load", "ps_module_unload", "ps_module_get_global",
           "ps_module_get_function", "ps_function_info", "ps_malloc", "ps_free", "ps_memcpy_htod",
           "ps_memcpy_dtoh", "ps_memcpy_dtod", "ps_memset", "ps_host_malloc", "ps_host_free",
           "ps_memcpy_htod_async", "ps_memcpy_dtoh_async", "ps_memcpy_dtod_async", "ps_launch", "ps_synchronize",
           "ps_event_create", "ps_event_record", "ps_event_elapsed_ms", "ps_event_synchronize", "ps_event_query",
           "ps_event_destroy", "ps_stream_wait_event",
           "ps_swept_run", "ps_standard_run", "ps_swept_run_timed", "ps_standard_run_timed",
           "ps_band_run", "ps_band_run_timed", "ps_copy_f64",
           "ps_level_sums", "ps_level_sums_noalloc",
           "ps_multi_index", "ps_multi_run", "ps_multi_run_timed", "ps_graph_begin", "ps_graph_end", "ps_graph_launch",
           "ps_graph_destroy")


def last_error():
    return _lib.ps_last_error().decode(errors="replace")


def check(rc, what=""):
    if rc != 0:
        raise HipError(f"{what}: {last_error()}" if what else last_error())


def version():
    return _version()


def device_count():
    n = c_int(0)
    check(_device_count(ctypes.byref(n)), "ps_device_count")
    return n.value


_initialised = False


def init(device=0):
    global _initialised
    check(_init(device), "ps_init")
    _initialised = True


def ensure_init(device=0):
    if not _initialised:
        init(device)


def shutdown():
    global _initialised
    check(_shutdown(), "ps_shutdown")
    _initialised = False


def current_context():
    """Handle of the context current on the calling thread (``Context.handle``),
    or None when the thread uses the process context of ``init``."""
    h = c_void_p()
    check(_ctx_current(ctypes.byref(h)), "ps_context_current")
    return h.value


def current_device():
    """Device of the calling thread's current context."""
    d = c_int()
    check(_ctx_device(ctypes.byref(d)), "ps_context_device")
    return d.value


class Context:
    """A device context of its own: its device, work stream and launch timer
    (ps_context_create).  ``with ctx:`` makes it current on the calling thread
    (the reference's ``make_context()`` / ``push()`` / ``pop()``,
    core/block.py:36-38); every copy, launch, run and synchronize issued by
    that thread inside the block is ordered on this context's stream, and the
    modules loaded inside it (``kernels.load``) are this context's own, with
    their own ``__constant__`` memory.  One context must not be used by two
    threads at once; distinct contexts may run on distinct threads."""

    def __init__(self, device=0):
        h = c_void_p()
        check(_ctx_create(device, ctypes.byref(h)), "ps_context_create")
        self.handle = h.value
        self.device = device

    def push(self):
        check(_ctx_push(self.handle), "ps_context_push")

    def pop(self):
        check(_ctx_pop(), "ps_context_pop")

    def __enter__(self):
        self.push()
        return self

    def __exit__(self, exc_type, exc, tb):
        self.pop()
        return False

    def close(self):
        """Drain the stream and release the context (not while current).
        Caches keyed by the context (``kernels.load``) forget it first, so a
        later context that happens to get the same handle starts afresh."""
        if self.handle:
            check(_ctx_destroy(self.handle), "ps_context_destroy")
            for hook in CONTEXT_CLOSE_HOOKS:
                hook(self.handle)
            self.handle = None


#: functions called with a context's handle when it is closed
CONTEXT_CLOSE_HOOKS = []


_compiler_version_value = None


def compiler_version():
    """(major, minor) of the hiprtc library behind ps_compile."""
    global _compiler_version_value
    if _compiler_version_value is None:
        a, b = c_int(), c_int()
        check(_compiler_version(ctypes.byref(a), ctypes.byref(b)), "ps_compiler_version")
        _compiler_version_value = (a.value, b.value)
    return _compiler_version_value


def compile_source(source, name, options):
    """hiprtc-compile source text; returns (code bytes, log text)."""
    opts = (c_char_p * len(options))(*[o.encode() for o in options])
    code, size, log = c_void_p(), c_size_t(), c_void_p()
    rc = _compile(source.encode(), name.encode(), opts, len(options), ctypes.byref(code),
                  ctypes.byref(size), ctypes.byref(log))
    logtxt = ""
    if log.value:
        logtxt = ctypes.string_at(log.value).decode(errors="replace")
        _free_host(log)
    if rc != 0:
        raise HipError(last_error())
    data = ctypes.string_at(code.value, size.value)
    _free_host(code)
    return data, logtxt


class Module:
    """A loaded code object."""

    def __init__(self, code):
        ensure_init()
        self._code = code  # kept alive for the runtime
        # the problem whose constants the module's __constant__ globals hold,
        # as recorded by a caller that re-uses them (None: unknown)
        self.const_owner = None
        h = c_void_p()
        check(_module_load(code, len(code), ctypes.byref(h)), "ps_module_load")
        self.handle = h.value

    def get_function(self, name):
        f = c_void_p()
        check(_get_function(self.handle, name.encode(), ctypes.byref(f)), "ps_module_get_function")
        return f.value

    def get_global(self, name):
        p, n = c_void_p(), c_size_t()
        check(_get_global(self.handle, name.encode(), ctypes.byref(p), ctypes.byref(n)), "ps_module_get_global")
        return p.value, n.value

    def unload(self):
        if self.handle:
            _module_unload(self.handle)
            self.handle = None


def function_info(fn):
    r, l, m = c_int(), c_int(), c_int()
    check(_function_info(fn, ctypes.byref(r), ctypes.byref(l), ctypes.byref(m)), "ps_function_info")
    return {"regs": r.value, "lds_bytes": l.value, "max_threads": m.value}


class DeviceArray:
    """Owned device allocation (freed on close / garbage collection)."""

    def __init__(self, nbytes):
        ensure_init()
        p = c_void_p()
        check(_malloc(ctypes.byref(p), max(int(nbytes), 1)), "ps_malloc")
        self.ptr = p.value
        self.nbytes = int(nbytes)

    def upload(self, arr, offset=0):
        a = numpy.ascontiguousarray(arr)
        assert offset + a.nbytes <= self.nbytes
        check(_htod(self.ptr + offset, a.ctypes.data, a.nbytes), "ps_memcpy_htod")

    def download(self, out, offset=0):
        assert out.flags.c_contiguous and offset + out.nbytes <= self.nbytes
        check(_dtoh(out.ctypes.data, self.ptr + offset, out.nbytes), "ps_memcpy_dtoh")
        return out

    def fill_zero(self):
        check(_memset(self.ptr, 0, self.nbytes), "ps_memset")

    def fill_bytes(self, value, nbytes=None):
        """Set the first ``nbytes`` (default all) bytes to ``value``
        (stream ordered)."""
        check(_memset(self.ptr, int(value) & 0xFF, self.nbytes if nbytes is None else int(nbytes)), "ps_memset")

    def close(self):
        if self.ptr:
            _free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PinnedArray:
    """Page-locked host buffer (ps_host_malloc) with a numpy view; the host
    side of the asynchronous copies, which return before the data moved."""

    def __init__(self, nbytes):
        ensure_init()
        p = c_void_p()
        check(_host_malloc(ctypes.byref(p), max(int(nbytes), 1)), "ps_host_malloc")
        self.ptr = p.value
        self.nbytes = int(nbytes)

    def view(self, dtype, shape):
        """numpy array over the first bytes of the buffer (valid while it lives)."""
        dtype = numpy.dtype(dtype)
        n = int(numpy.prod(shape))
        assert n * dtype.itemsize <= self.nbytes
        raw = (ctypes.c_char * (n * dtype.itemsize)).from_address(self.ptr)
        return numpy.frombuffer(raw, dtype=dtype, count=n).reshape(shape)

    def upload_async(self, dptr, nbytes, offset=0):
        """Copy nbytes from offset of this buffer to device address dptr, in
        stream order, without waiting."""
        assert offset + nbytes <= self.nbytes
        check(_htod_async(dptr, self.ptr + offset, nbytes), "ps_memcpy_htod_async")

    def download_async(self, dptr, nbytes, offset=0):
        """Copy nbytes from device address dptr to offset of this buffer, in
        stream order, without waiting."""
        assert offset + nbytes <= self.nbytes
        check(_dtoh_async(self.ptr + offset, dptr, nbytes), "ps_memcpy_dtoh_async")

    def close(self):
        if self.ptr:
            _host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dtod_async(dst, src, nbytes):
    """Stream-ordered device copy that returns without waiting."""
    check(_dtod_async(dst, src, nbytes), "ps_memcpy_dtod_async")


def htod(dptr, arr):
    a = numpy.ascontiguousarray(arr)
    check(_htod(dptr, a.ctypes.data, a.nbytes), "ps_memcpy_htod")


def dtoh(out, dptr):
    """Copy out.nbytes bytes from device address dptr into the contiguous
    numpy array ``out`` (waits for the work stream)."""
    assert out.flags.c_contiguous
    check(_dtoh(out.ctypes.data, dptr, out.nbytes), "ps_memcpy_dtoh")
    return out


def htod_bytes(dptr, data):
    buf = ctypes.create_string_buffer(bytes(data), len(data))
    check(_htod(dptr, ctypes.addressof(buf), len(data)), "ps_memcpy_htod")


def dtod(dst, src, nbytes):
    check(_dtod(dst, src, nbytes), "ps_memcpy_dtod")


def synchronize():
    check(_sync(), "ps_synchronize")


def launch(fn, grid, block, args):
    """args: list of ctypes objects (kernel arguments in order)."""
    arr = (c_void_p * len(args))(*[ctypes.cast(ctypes.pointer(a), c_void_p) for a in args])
    g = tuple(grid) + (1,) * (3 - len(grid))
    b = tuple(block) + (1,) * (3 - len(block))
    check(_launch(fn, g[0], g[1], g[2], b[0], b[1], b[2], 0, arr), "ps_launch")


def swept_run(plan, G, c_begin, c_end):
    check(_swept_run(ctypes.byref(plan), G, c_begin, c_end), "ps_swept_run")


def standard_run(plan, g_begin, g_end):
    check(_standard_run(ctypes.byref(plan), g_begin, g_end), "ps_standard_run")


def swept_run_timed(plan, G, c_begin, c_end, stats):
    check(_swept_run_timed(ctypes.byref(plan), G, c_begin, c_end, ctypes.byref(stats)), "ps_swept_run_timed")


def standard_run_timed(plan, g_begin, g_end, stats):
    check(_standard_run_timed(ctypes.byref(plan), g_begin, g_end, ctypes.byref(stats)), "ps_standard_run_timed")


def band_run(plan, A, stages, stats=None):
    if stats is None:
        check(_band_run(ctypes.byref(plan), A, stages), "ps_band_run")
    else:
        check(_band_run_timed(ctypes.byref(plan), A, stages, ctypes.byref(stats)), "ps_band_run_timed")


def multi_index(plan):
    """Fill the plan's owner tables from its prefix sums (on the work stream)."""
    check(_multi_index(ctypes.byref(plan)), "ps_multi_index")


def multi_run(plan, rounds, stats=None):
    if stats is None:
        check(_multi_run(ctypes.byref(plan), rounds), "ps_multi_run")
    else:
        check(_multi_run_timed(ctypes.byref(plan), rounds, ctypes.byref(stats)), "ps_multi_run_timed")


def copy_f64(dst, src, n):
    """Stream-ordered device copy of n doubles by the library's copy kernel."""
    check(_copy_f64(dst, src, n), "ps_copy_f64")


def level_sums_async(level_ptr, fp32, nv, cells, out_ptr):
    """Stream-ordered FP64 sums of a level's nv variables into device memory
    ``out_ptr`` (2*nv doubles: the sums, then the sums of magnitudes)."""
    check(_level_sums(level_ptr, int(bool(fp32)), nv, cells, out_ptr), "ps_level_sums")


def level_sums_noalloc_async(level_ptr, fp32, nv, cells, out_ptr):
    """``level_sums_async`` through the entry point that allocates nothing
    (its partial sums in the library's fixed per-GPU array)."""
    check(_level_sums_noalloc(level_ptr, int(bool(fp32)), nv, cells, out_ptr), "ps_level_sums_noalloc")


def level_sums(level_ptr, fp32, nv, cells):
    """FP64 (sums, sums of magnitudes) of a device level [nv][cells], each an
    array of nv values (waits for the work stream)."""
    out = DeviceArray(2 * nv * 8)
    level_sums_async(level_ptr, fp32, nv, cells, out.ptr)
    host = numpy.empty(2 * nv)
    out.download(host)
    out.close()
    return host[:nv].copy(), host[nv:].copy()


class Event:
    def __init__(self):
        h = c_void_p()
        check(_event_create(ctypes.byref(h)), "ps_event_create")
        self.h = h.value

    def record(self):
        check(_event_record(self.h), "ps_event_record")

    def synchronize(self):
        check(_event_sync(self.h), "ps_event_synchronize")

    def query(self):
        """True once the work recorded before the event has completed."""
        done = c_int()
        check(_event_query(self.h, ctypes.byref(done)), "ps_event_query")
        return bool(done.value)

    def wait(self):
        """The current context's work stream waits (on the device) for the
        work recorded before this event, on whichever stream it was
        recorded; the host does not wait."""
        check(_stream_wait(self.h), "ps_stream_wait_event")

    def elapsed_ms(self, stop):
        ms = ctypes.c_float()
        check(_event_elapsed(self.h, stop.h, ctypes.byref(ms)), "ps_event_elapsed_ms")
        return ms.value

    def __del__(self):
        try:
            _event_destroy(self.h)
        except Exception:
            pass


class Graph:
    """Capture of the work stream: ``with Graph() as g: <launches>``; then
    ``g.launch()`` replays them."""

    def __init__(self):
        self.h = None

    def __enter__(self):
        check(_graph_begin(), "ps_graph_begin")
        return self

    def __exit__(self, exc_type, exc, tb):
        h = c_void_p()
        rc = _graph_end(ctypes.byref(h))
        if exc_type is None:
            check(rc, "ps_graph_end")
            self.h = h.value
        return False

    def launch(self):
        check(_graph_launch(self.h), "ps_graph_launch")

    def __del__(self):
        try:
            if self.h:
                _graph_destroy(self.h)
        except Exception:
            pass
