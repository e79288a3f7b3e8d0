"""Many problems on one GPU: batched launches, continuous admission, graphs.

Problems of the same plugin, state precision and block size form a class.
A class owns one instance-table module of each kernel kind (the swept
phases and the standard stage, compiled with per-instance constant blocks)
and a table ``ps_inst[]`` on the device.  One *round* (``ps_multi_run``)
launches every phase kernel once over all tiles of all problems in the
table; each problem executes the phases of its own cycle (FirstPrism,
UpPrism, LastPrism or an exact-step tail stage), so problems admitted at
different times, with different grids and step counts, share launches.  A
plugin with a band sweep of one step per launch (Euler) adds the band
method: such a problem's rounds are its stages, and the class's BandMulti
launch takes one whole launch's stages (one RK2 step for Euler, up to 12
levels for heat) of every band-method problem whose stage begins one, each
problem in row segments of its own length (``kernels.band_multi_rows``).

``Engine`` keeps the state of every admitted problem resident in HBM, admits
queued problems while the HBM budget allows, retires finished ones (their
final level is copied to the host and their memory reused), and replays a
captured hipGraph of the round while the set of running problems does not
change.  Results do not depend on what else is running: each tile reads only
its own problem's ring and constants.

The host never waits for the GPU to drain between rounds: initial
conditions, constants and instance tables reach the device through
page-locked staging buffers and stream-ordered copies, a retiring problem's
final level comes back the same way (collected once an event behind the
copy has completed), and freed rings go to a pool instead of ``hipFree``
(which waits for the whole device).
"""
import collections
import ctypes
import dataclasses
import math
import time
import warnings

import numpy

from pysweep_amd import device, hip, kernels, plugins, schedule
from pysweep_amd.equations import path as eq_path

PLUGINS = {
    # name: (python half, device source fp64, device source fp32, nv, ops, its)
    "euler": ("euler.py", "euler.cu", "euler_fp32.cu", 4, 2, 2),
    "heat": ("heat.py", "heat.cu", "heat_fp32.cu", 1, 1, None),
}


#: method of a Problem with method="auto" when no tuner record says
#: otherwise: the faster one measured on MI355X for grids of 64..2048 cells a
#: side: the band method.  Euler: 512-problem mix of 64..1024 sides, the
#: tuner's 24-problem 40-step sample: band 9.6 GCUPS, standard 8.4, swept
#: b=32 6.4; the saturated served stream of 128..2048 sides under the "all"
#: policy: 29.8 solves/s against 23.8, its latencies unchanged.  Heat: the
#: mix's 256 heat problems alone 0.107 s against 0.157 s on the standard
#: kernel, for rings of 13 levels instead of 2 (the whole mix's peak HBM 14.4
#: GB against 8.4).  "auto" falls back to standard for a class without the
#: band method (Engine._method_for).
DEFAULT_METHOD = {
    ("euler", "fp64"): "band",
    ("euler", "fp32"): "band",
    ("heat", "fp64"): "band",
    ("heat", "fp32"): "band",
}


@dataclasses.dataclass(eq=False)
class Problem:
    """One solve: plugin name, initial condition (nv, nx, ny), the plugin's
    globals [t0, tf, dt, dx, dy, ...], intermediate steps ``its`` and state
    precision ("fp64" or "fp32")."""
    plugin: str
    ic: numpy.ndarray
    globals: list
    its: int
    precision: str = "fp64"
    b: int = None
    exact: bool = True
    tag: object = None
    method: str = "swept"  # "swept"; "standard": every stage on the standard kernel; "band": every stage on the
    #                        band kernel (kernels.band_multi_capable); "auto": the tuner's record for the class,
    #                        else DEFAULT_METHOD


class _HostConsts:
    """The plugin's device constants of one problem, written by its
    set_globals through ``source_mod.get_global`` / ``cuda.memcpy_htod``."""

    def __init__(self, layout):
        self.layout = layout
        self.buf = bytearray(layout.size)

    def get_global(self, name):
        f = self.layout.field(name)
        return plugins.HostField(self.buf, f.offset, f.size, name), f.size

    def get_function(self, name):
        raise KeyError(name)


def class_specs(plugin, precision, its, b, slots=None, threads=None, stile=None, sthreads=None):
    """(swept, standard) instance-table module specs of a class: ``slots``
    and ``threads`` shape the swept tile's workgroup (LDS window levels,
    lanes), ``stile`` = (rows, columns) and ``sthreads`` the standard one;
    None takes the measured defaults (the standard tile of a 4b x 4b grid)."""
    py, cu64, cu32, nv, ops, _ = PLUGINS[plugin]
    real = "double" if precision == "fp64" else "float"
    cu = eq_path(cu64 if precision == "fp64" else cu32)
    P = b // (2 * ops) - 1
    spec = kernels.swept_spec(cu, nv, ops, its, b, slots=slots, threads=threads, real=real, multi=True)
    bx, by = stile if stile else (None, None)
    sspec = kernels.standard_spec(cu, nv, ops, its, max(b * 4, bx or 0), max(b * 4, by or 0), bx=bx, by=by,
                                  threads=sthreads, real=real, multi=True, mpss=P)
    return spec, sspec


def solo_standard_spec(plugin, precision, its, nx, ny):
    """The plain standard module a problem of this grid uses when it runs
    alone (its tile fits the grid)."""
    py, cu64, cu32, nv, ops, _ = PLUGINS[plugin]
    real = "double" if precision == "fp64" else "float"
    cu = eq_path(cu64 if precision == "fp64" else cu32)
    return kernels.standard_spec(cu, nv, ops, its, nx, ny, real=real)


def solo_specs(plugin, precision, its, b):
    """The plain modules a class uses for a problem running alone: the
    swept phases and the standard modules of the tiles its grids can take
    (a tile no larger than the grid; the tile depends on the grid only
    through that bound)."""
    py, cu64, cu32, nv, ops, _ = PLUGINS[plugin]
    real = "double" if precision == "fp64" else "float"
    cu = eq_path(cu64 if precision == "fp64" else cu32)
    out = [kernels.swept_spec(cu, nv, ops, its, b, real=real)]
    sides = [s for s in (16, 32, 64, 128, 1 << 16) if s >= b]
    for nx in sides:
        for ny in sides:
            spec = solo_standard_spec(plugin, precision, its, nx, ny)
            if spec not in out:
                out.append(spec)
    return out


def shipped_class_specs():
    """Instance-table modules (and their one-problem modules) used by the
    tests and benchmarks."""
    from pysweep_amd import autotune
    out = []
    for plugin, precision, its, b in (("euler", "fp64", 2, 8), ("euler", "fp64", 2, 16), ("euler", "fp64", 2, 32),
                                      ("euler", "fp32", 2, 16), ("euler", "fp32", 2, 32),
                                      ("heat", "fp64", 1, 16), ("heat", "fp64", 1, 32), ("heat", "fp64", 2, 16),
                                      ("heat", "fp64", 1, 64), ("heat", "fp32", 1, 16), ("heat", "fp32", 1, 32)):
        out.extend(class_specs(plugin, precision, its, b))
        out.extend(solo_specs(plugin, precision, its, b))
    # the autotuner's candidates of the batched workload's classes
    for plugin, precision, its in (("euler", "fp64", 2), ("heat", "fp64", 1), ("heat", "fp64", 2)):
        for cand in autotune.CANDIDATES[plugin]:
            out.extend(class_specs(plugin, precision, its, cand["b"], **autotune.shape_of(cand)))
    # the instance-table band modules of the classes that have one
    for plugin, precision, its in (("euler", "fp64", 2), ("euler", "fp32", 2), ("heat", "fp64", 1),
                                   ("heat", "fp64", 2), ("heat", "fp32", 1)):
        py, cu64, cu32, nv, ops, _ = PLUGINS[plugin]
        out.append(kernels.band_spec(eq_path(cu64 if precision == "fp64" else cu32), nv, ops, its,
                                     real="double" if precision == "fp64" else "float", multi=True))
    seen, uniq = set(), []
    for sp in out:
        if sp not in seen:
            seen.add(sp)
            uniq.append(sp)
    return uniq


class PluginClass:
    """Compiled instance-table modules of one (plugin, precision, b)."""

    def __init__(self, plugin, precision, its, b, slots=None, threads=None, stile=None, sthreads=None):
        py, cu64, cu32, nv, ops, default_its = PLUGINS[plugin]
        self.plugin, self.precision, self.nv, self.ops, self.its, self.b = plugin, precision, nv, ops, its, b
        self.real = "double" if precision == "fp64" else "float"
        self.dtype = numpy.float64 if precision == "fp64" else numpy.float32
        self.rb = numpy.dtype(self.dtype).itemsize
        self.P = b // (2 * ops) - 1
        self.spec, self.sspec = class_specs(plugin, precision, its, b, slots, threads, stile, sthreads)
        self.layout = self.spec.const_layout()
        m = kernels.load(self.spec)
        sm = kernels.load(self.sspec)
        self.modules = [m, sm]  # every module the class has run (division-range flags)
        device.range_fault(self.modules, clear=True)
        self.fn = {k: m.get_function(k + "Multi") for k in kernels.SWEPT_KERNELS}
        self.fn["Standard"] = sm.get_function("StandardMulti")
        self.fn["advance"] = m.get_function("ps_advance")
        self.cpu = plugins.load_cpu_module(eq_path(py), name=f"pysweep_batch_{plugin}")
        self.key = (plugin, precision, its, b)
        self._solo = None
        self._band = None
        self.cu = eq_path(cu64 if precision == "fp64" else cu32)
        self.band_multi_ok = kernels.band_multi_capable(self.cu, nv, ops, its)
        # levels per band-method launch (a band-method problem's rounds come
        # in whole launches of this many stages)
        self.band_h = kernels.band_multi_levels(self.cu, its, self.real) if self.band_multi_ok else 0
        self.bspec = None  # the instance-table band module's spec, once loaded

    def band_module(self, nx, ny):
        """The class's band module, loaded on first use, when the band kernel
        outruns the standard kernel on an nx x ny grid (a problem running
        alone then takes its stages on it); else None."""
        if not kernels.band_beats_standard(self.cu, self.nv, self.ops, self.its, nx, ny):
            return None
        if self._band is None:
            spec = kernels.band_spec(self.cu, self.nv, self.ops, self.its, real=self.real)
            m = kernels.load(spec)
            self.modules.append(m)
            self._band = (spec, m, m.get_function("Band"))
        return self._band

    def band_multi(self):
        """The class's BandMulti kernel (its instance-table band module,
        loaded on first use)."""
        if not self.band_multi_ok:
            raise ValueError(f"class {self.key} has no instance-table band form (kernels.band_multi_capable)")
        if "BandMulti" not in self.fn:
            self.bspec = kernels.band_spec(self.cu, self.nv, self.ops, self.its, real=self.real, multi=True)
            m = kernels.load(self.bspec)
            self.modules.append(m)
            self.fn["BandMulti"] = m.get_function("BandMulti")
        return self.fn["BandMulti"]

    def levels(self, method):
        """Ring levels of a problem: the swept window 2P+s (core/solver.py:123),
        s+1 for a problem that runs every stage on the standard kernel, and
        band_h+1 on the band method (a launch reads its first level and
        writes its last, band_h later: s+1 for Euler's one-step launches)."""
        if method == "swept":
            return 2 * self.P + self.its
        return self.band_h + 1 if method == "band" else self.its + 1

    def ring_bytes(self, nx, ny, method="swept"):
        return self.levels(method) * self.nv * nx * ny * self.rb

    def solo_modules(self, nx, ny):
        """The class's one-problem modules (the plain swept phases and the
        standard kernel with a tile that fits an nx x ny grid), for a
        problem that runs alone."""
        if self._solo is None:
            sw = solo_specs(self.plugin, self.precision, self.its, self.b)[0]
            msw = kernels.load(sw)
            self.modules.append(msw)
            self._solo = (sw, msw, {n: msw.get_function(n) for n in kernels.SWEPT_KERNELS})
        sw, msw, f = self._solo
        st = solo_standard_spec(self.plugin, self.precision, self.its, nx, ny)
        mst = kernels.load(st)
        if mst not in self.modules:
            self.modules.append(mst)
        return sw, st, msw, mst, f, mst.get_function("Standard")


class _Pool:
    """Free buffers kept for reuse: ``take`` returns the smallest free one
    of at least ``nbytes`` and at most ``slack`` times that, else a new one
    from ``make``, after releasing free buffers (largest first) until the
    free ones and the new one fit in ``limit`` bytes."""

    def __init__(self, make, slack=1.25):
        self.make = make
        self.slack = slack
        self.free = []
        self.bytes = 0  # held by free buffers

    def take(self, nbytes, limit=None):
        best = None
        for b in self.free:
            if nbytes <= b.nbytes <= nbytes * self.slack and (best is None or b.nbytes < best.nbytes):
                best = b
        if best is not None:
            self.free.remove(best)
            self.bytes -= best.nbytes
            return best
        if limit is not None:
            self.trim(limit - nbytes)
        return self.make(nbytes)

    def give(self, b):
        self.free.append(b)
        self.bytes += b.nbytes

    def trim(self, keep):
        """Release free buffers, largest first, until they hold at most
        ``keep`` bytes."""
        while self.free and self.bytes > keep:
            b = max(self.free, key=lambda x: x.nbytes)
            self.free.remove(b)
            self.bytes -= b.nbytes
            b.close()


class _Slice:
    """``nbytes`` bytes at ``offset`` of a page-locked buffer."""

    def __init__(self, buf, offset, nbytes):
        self.buf, self.offset, self.nbytes = buf, offset, nbytes
        self.ptr = buf.ptr + offset

    def view(self, dtype, shape):
        dtype = numpy.dtype(dtype)
        n = int(numpy.prod(shape))
        assert n * dtype.itemsize <= self.nbytes
        return self.buf.view(numpy.uint8, (self.offset + n * dtype.itemsize,))[self.offset:].view(dtype).reshape(shape)

    def upload_async(self, dptr, nbytes, offset=0):
        assert offset + nbytes <= self.nbytes
        self.buf.upload_async(dptr, nbytes, self.offset + offset)

    def download_async(self, dptr, nbytes, offset=0):
        assert offset + nbytes <= self.nbytes
        self.buf.download_async(dptr, nbytes, self.offset + offset)


class _StagingRing:
    """One page-locked buffer handed out in first-in first-out order: the
    engine's copies to and from the host go through it, and each piece is
    given back when the event recorded behind its copy has completed, which
    happens in allocation order (one stream).  Every page is touched when
    the buffer is made: first touches of page-locked memory cost about
    0.36 ms per MiB on MI355X hosts, twenty times a later copy into it."""

    ALIGN = 256

    def __init__(self, nbytes, make=None):
        self.size = int(nbytes)
        self.buf = (make or hip.PinnedArray)(self.size)
        self.buf.view(numpy.uint8, (self.size,))[:] = 0
        self.live = collections.deque()  # (offset, bytes) in allocation order

    def alloc(self, nbytes):
        """Offset of ``nbytes`` free bytes, or None if they do not fit now."""
        n = -(-max(int(nbytes), 1) // self.ALIGN) * self.ALIGN
        if n > self.size:
            return None
        if not self.live:
            off = 0
        else:
            first, (loff, ln) = self.live[0][0], self.live[-1]
            head = loff + ln
            if loff >= first:  # free: [head, size) and [0, first)
                off = head if head + n <= self.size else (0 if n <= first else None)
            else:  # wrapped: free [head, first)
                off = head if head + n <= first else None
        if off is not None:
            self.live.append((off, n))
        return off

    def release(self, count):
        """Give back the ``count`` oldest pieces."""
        for _ in range(count):
            self.live.popleft()


def _fetch(item):
    """A finished ticket's result out of its staging piece (FP64 copy), and
    its conserved-sum record from the device sums of its first and final
    levels."""
    t, (piece, dtype, shape) = item
    t.result = piece.view(dtype, shape).astype(numpy.float64)
    nv = shape[0]
    lvl = t.result.size * numpy.dtype(dtype).itemsize
    s = _Slice(piece.buf, piece.offset + lvl, 256).view(numpy.float64, (2, 16))
    t.conservation = device.conservation((s[0, :nv], s[0, nv:2 * nv]), (s[1, :nv], s[1, nv:2 * nv]))


_COPY_POOL = []


def _copy_pool():
    if not _COPY_POOL:
        import concurrent.futures
        _COPY_POOL.append(concurrent.futures.ThreadPoolExecutor(max_workers=8))
    return _COPY_POOL[0]


def _copy_into(dst, src, piece=4 << 20):
    """dst[...] = src (with the assignment's casting), in slabs of rows on
    the copy threads when it is large (numpy drops the GIL while copying):
    one thread copies about 25 GB/s into page-locked memory on MI355X hosts,
    which left the GPU waiting while a batch's initial conditions went in."""
    n = src.shape[1]
    parts = min(8, n, src.nbytes // piece)
    if parts < 2:
        dst[...] = src
        return
    edges = [n * k // parts for k in range(parts + 1)]

    def one(k):
        dst[:, edges[k]:edges[k + 1]] = src[:, edges[k]:edges[k + 1]]
    list(_copy_pool().map(one, range(parts)))


_FREE_RINGS = []  # staging rings of finished engines, for the next engine
STAGING_BYTES = 1 << 30


def _take_ring(nbytes=0):
    """A free staging ring of at least ``nbytes`` (and STAGING_BYTES)."""
    nbytes = max(nbytes, STAGING_BYTES)
    for r in sorted(_FREE_RINGS, key=lambda r: r.size):
        if r.size >= nbytes:
            _FREE_RINGS.remove(r)
            return r
    return _StagingRing(nbytes)


_DEVICE_RINGS = {}  # (context handle, copies on the copy stream) -> free problem rings


def _ring_pool(copy_stream=True):
    """The free problem rings of the calling thread's context, shared by its
    engines that queue their copies the same way: a finished engine's rings
    serve the next engine instead of being freed (hipFree waits for the
    whole device, so an engine collected while another runs would stall it
    once per ring).  A ring given back after its last copy was queued is
    safe to reuse for later copies on the same stream (the copy stream, or
    the work stream), so engines that differ there keep separate rings."""
    key = (hip.current_context(), bool(copy_stream))
    pool = _DEVICE_RINGS.get(key)
    if pool is None:
        if _forget_context not in hip.CONTEXT_CLOSE_HOOKS:
            hip.CONTEXT_CLOSE_HOOKS.append(_forget_context)
        pool = _DEVICE_RINGS[key] = _Pool(hip.DeviceArray)
    return pool


_COPY_CONTEXTS = {}  # context handle -> the context whose stream carries its engines' copies


def _copy_context():
    """The context whose stream carries the copies of the calling thread's
    context's engines: one per work context, so that every copy into or out
    of a pooled ring is ordered on one stream (a download out of a retired
    problem's ring before the next problem's upload into it)."""
    key = hip.current_context()
    ctx = _COPY_CONTEXTS.get(key)
    if ctx is None:
        if _forget_context not in hip.CONTEXT_CLOSE_HOOKS:
            hip.CONTEXT_CLOSE_HOOKS.append(_forget_context)
        ctx = _COPY_CONTEXTS[key] = hip.Context(hip.current_device())
    return ctx


def _forget_context(handle):
    """A closed work context's pooled rings and copy context go with it."""
    for copy_stream in (True, False):
        pool = _DEVICE_RINGS.pop((handle, copy_stream), None)
        if pool is not None:
            pool.trim(0)
    ctx = _COPY_CONTEXTS.pop(handle, None)
    if ctx is not None:
        ctx.close()


def release_rings():
    """Free the pooled problem rings of the calling thread's context (waits
    for the device); engines made later allocate afresh."""
    for copy_stream in (True, False):
        pool = _DEVICE_RINGS.pop((hip.current_context(), copy_stream), None)
        if pool is not None:
            pool.trim(0)


#: bytes after a problem's ring: its constant block (256) and the FP64 sums
#: of its first and final levels (128 each: 2 * nv <= 16 doubles)
_TAIL_BYTES = 512


@dataclasses.dataclass(eq=False)
class Ticket:
    problem: Problem
    cls: PluginClass = None
    nx: int = 0
    ny: int = 0
    times: object = None
    ring: object = None  # device buffer: the ring, the constant block, the level sums
    consts: int = 0      # device address of the constant block
    sums: int = 0        # device address of the FP64 level sums: first level, then final
    conservation: dict = None  # conserved-sum record of the result (device.conservation)
    division_exact: bool = True  # False: its class met a state outside the plugin's division range
    c: int = 0
    c_end: int = 0
    final_level: int = 0
    result: numpy.ndarray = None
    arrival: float = 0.0
    admitted: float = None
    finished: float = None
    nbytes: int = 0
    method: str = None
    T: int = 0           # ring levels
    cells: int = 0
    last_active: float = 0.0
    boosted: bool = False

    @property
    def remaining_work(self):
        """Cell-stages still to compute: a problem's swept rounds together
        cover (G+1)P stages (UpPrism rounds P each, FirstPrism and LastPrism
        P/2 each), then its tail rounds one stage each."""
        if self.times is None:
            return 0
        G, P = self.times.G, self.times.P
        swept_end = G + 2 if G >= 0 else 1
        n = max(0, swept_end - max(self.c, 0))  # swept rounds left
        # UpPrism rounds advance P stages; FirstPrism and LastPrism half of that each
        left = P * n - (0.5 * P if n and self.c <= 0 else 0) - (0.5 * P if n else 0)
        left += max(0, self.c_end - max(self.c, swept_end))  # tail stages
        return left * self.cells

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival

    @property
    def done(self):
        return self.result is not None


class _ClassTable:
    """Device instance table of one class."""

    def __init__(self, cls):
        self.cls = cls
        self.running = []
        self.dirty = True
        self.insts = self.prefix = self.sprefix = self.bprefix = None
        self.cap = 0
        self.plan = None
        self.graph = None
        self.owner = self.sowner = self.bowner = None
        self.flags = None
        # arrays replaced by larger ones: freed once the engine is idle
        # (hipFree waits for the whole device; queued rounds may still read them)
        self.old = []

    def needs(self, active):
        """True when the table's run flags differ from ``active``."""
        return self.flags != tuple(t in active for t in self.running)

    def rebuild(self, active=None, stage=None):
        """Write the table of the running problems (those in ``active``, or
        all, take part in the rounds).  ``stage(nbytes)`` returns a page-locked
        buffer that stays untouched until the copies queued now have run; the
        table goes to the device from it without waiting.  Without it the
        copies wait."""
        cls = self.cls
        n = len(self.running)
        if n == 0:
            self.plan = None
            self.graph = None
            self.dirty = False
            self.flags = ()
            return
        if n > self.cap:
            self.cap = max(8, 2 * n)
            self.old.extend(a for a in (self.insts, self.prefix, self.sprefix, self.bprefix) if a is not None)
            self.insts = hip.DeviceArray(self.cap * ctypes.sizeof(hip.Inst))
            self.prefix = hip.DeviceArray((self.cap + 1) * 4)
            self.sprefix = hip.DeviceArray((self.cap + 1) * 4)
            self.bprefix = hip.DeviceArray((self.cap + 1) * 4)
        isz, psz = n * ctypes.sizeof(hip.Inst), 4 * (n + 1)
        host = stage(isz + 3 * psz) if stage else None
        if host is not None:
            arr = (hip.Inst * n).from_address(host.ptr)
            words = host.view(numpy.int32, ((isz + 3 * psz) // 4,))
            o = isz // 4
            pre, spre, bpre = words[o:o + n + 1], words[o + n + 1:o + 2 * n + 2], words[o + 2 * n + 2:]
            pre[:] = 0
            spre[:] = 0
            bpre[:] = 0
        else:
            arr = (hip.Inst * n)()
            pre = numpy.zeros(n + 1, dtype=numpy.int32)
            spre = numpy.zeros(n + 1, dtype=numpy.int32)
            bpre = numpy.zeros(n + 1, dtype=numpy.int32)
        sbx, sby = cls.sspec.bx, cls.sspec.by
        # band-method problems taking part: row segments of one launch
        banded = [t for t in self.running if t.method == "band" and (active is None or t in active)]
        geo, wout = {}, 1
        if banded:
            cls.band_multi()
            geo = dict(zip(banded, kernels.band_multi_rows(cls.bspec, [(t.nx, t.ny) for t in banded])))
            wout = kernels.band_out_columns(cls.bspec)
        for i, t in enumerate(self.running):
            st = arr[i]
            st.ring, st.consts = t.ring.ptr, t.consts
            st.nx, st.ny, st.T = t.nx, t.ny, t.T
            st.tiles_x, st.tiles_y = t.nx // cls.b, t.ny // cls.b
            st.stiles_x, st.stiles_y = -(-t.nx // sbx), -(-t.ny // sby)
            st.G, st.tail, st.c = t.times.G, t.times.tail, t.c
            st.active = int(active is None or t in active)
            band = t.method == "band"
            rows = geo.get(t, 0) if band else 0
            st.brows, st.bands = rows, (-(-t.ny // wout) if rows else 0)
            st.bsegs = -(-t.nx // rows) if rows else 0
            # problems without swept cycles (G = -1) take no swept work items,
            # problems sitting out of the rounds none at all; band-method
            # problems take band work items instead of standard tiles
            run = bool(st.active)
            pre[i + 1] = pre[i] + (st.tiles_x * st.tiles_y if t.times.G >= 0 and run else 0)
            spre[i + 1] = spre[i] + (st.stiles_x * st.stiles_y if t.times.tail and run and not band else 0)
            bpre[i + 1] = bpre[i] + (st.bands * st.bsegs if t.times.tail and run and band else 0)
        if host is not None:
            host.upload_async(self.insts.ptr, isz)
            host.upload_async(self.prefix.ptr, psz, isz)
            host.upload_async(self.sprefix.ptr, psz, isz + psz)
            host.upload_async(self.bprefix.ptr, psz, isz + 2 * psz)
        else:
            hip.check(hip._htod(self.insts.ptr, ctypes.addressof(arr), ctypes.sizeof(arr)), "ps_memcpy_htod")
            self.prefix.upload(pre)
            self.sprefix.upload(spre)
            self.bprefix.upload(bpre)
        self.flags = tuple(bool(arr[i].active) for i in range(n))
        # owner tables (problem of each work item), grown geometrically
        for attr, total in (("owner", int(pre[-1])), ("sowner", int(spre[-1])), ("bowner", int(bpre[-1]))):
            buf = getattr(self, attr, None)
            if total and (buf is None or buf.nbytes < 4 * total):
                if buf is not None:
                    self.old.append(buf)
                setattr(self, attr, hip.DeviceArray(8 * total))
        f = cls.fn
        self.plan = hip.MultiPlan(insts=self.insts.ptr, prefix=self.prefix.ptr, sprefix=self.sprefix.ptr,
                                  owner=self.owner.ptr if getattr(self, "owner", None) else None,
                                  sowner=self.sowner.ptr if getattr(self, "sowner", None) else None,
                                  ninst=n, total_tiles=int(pre[-1]), total_stiles=int(spre[-1]),
                                  threads=cls.spec.threads, sthreads=cls.sspec.threads,
                                  up=f["UpPyramid"], ybridge=f["YBridge"], xbridge=f["XBridge"],
                                  octahedron=f["Octahedron"], down=f["DownPyramid"], standard=f["Standard"],
                                  advance=f["advance"],
                                  bprefix=self.bprefix.ptr,
                                  bowner=self.bowner.ptr if getattr(self, "bowner", None) else None,
                                  total_bwaves=int(bpre[-1]), band=f.get("BandMulti"))
        hip.multi_index(self.plan)
        self.graph = None
        self.dirty = False


def short_first(running, now, capacity=1 << 20, max_wait=2.0):
    """Problems that take part in the next rounds (shortest remaining work
    first): the shortest always; others in order of remaining work while
    their cells fit in ``capacity`` (about what fills the GPU); and any
    problem that has been left out for ``max_wait`` seconds, from then on
    in every round until it finishes (so nothing starves).  A short request
    arriving while a long one runs is served at nearly its solo speed, while
    many short ones still share rounds."""
    order = sorted(running, key=lambda t: (not t.boosted, t.remaining_work))
    out, used = set(), 0
    for t in order:
        if not t.boosted and now - t.last_active >= max_wait:
            t.boosted = True
        if t.boosted or not out or used + t.cells <= capacity:
            out.add(t)
            if not t.boosted:
                used += t.cells
    for t in out:
        t.last_active = now
    return out


def oldest_first(running, capacity=1 << 20):
    """Problems that take part in the next rounds: in arrival order while
    their cells fit in ``capacity`` (about what fills the GPU), the oldest
    always.  A large problem then runs nearly alone, as a first-come,
    first-served server would run it, while small ones share its rounds only
    as far as they fill the GPU; nothing starves, since the oldest problem
    is always in."""
    out, used = set(), 0
    for t in sorted(running, key=lambda t: (t.arrival, t.admitted)):
        if out and used + t.cells > capacity:
            break
        out.add(t)
        used += t.cells
    return out


class Engine:
    """Continuous batching of swept solves on one GPU.

    ``policy``: "all" (default: every running problem in every round),
    "oldest_first" (arrival order up to ``capacity`` cells; see
    ``oldest_first``) or "short_first" (least remaining work first up to
    ``capacity`` cells, with aging; see ``short_first``).  Measured on
    MI355X on a Poisson stream of Euler solves: "oldest_first" matched the
    one-at-a-time server's p50 and p99; "short_first" lowered the mean
    latency below it but raised the p99; "all" had the best saturated
    throughput and the worst p99.

    ``use_graphs``: replay each table's rounds from a captured hipGraph of
    ``graph_rounds`` rounds (the remainder, and runs shorter than that,
    launched from the library's C loop).  Measured on MI355X: one-round
    graphs replayed from Python were 1-2 % slower than the C loop on the
    saturated served stream; graphs of 8 or 32 rounds run level with it
    (17.05-17.12 solves/s against 16.97-17.10; 512-problem mix 193-195
    against 191-195).

    ``poison``: a debug mode that fills each problem's ring with NaN before
    its initial condition goes in (a read of a cell that no phase wrote then
    turns the result into NaN).

    ``band_solo``: a problem running alone takes its standard stages on the
    band kernel where that outruns the standard kernel (heat at every size,
    Euler from 1024^2 cells: ``kernels.band_beats_standard``), in whole RK2
    steps, the stages before the first full step on the standard kernel;
    results are bitwise unchanged.

    ``copy_stream``: initial conditions go up and results come down on the
    stream of a second context (one per work context, shared by its
    engines), beside the rounds: the work stream waits on the device for
    each upload before using the problem, each download for the problem's
    last round.  False queues them on the work stream, in order with the
    rounds.

    ``tuner``: the ``autotune.Tuner`` whose records give each class its
    block size, the method of problems submitted with method="auto" and the
    workgroup shapes of its kernels; "default" loads the stored records
    (falling back to the measured defaults), None takes the first block
    size that divides the grid and the default shapes."""

    def __init__(self, hbm_budget=None, max_active=None, use_graphs=True, tuner="default", clock=None,
                 poll_rounds=4, max_rounds_ahead=1 << 20, policy="all", max_inflight=2, solo=True,
                 capacity=1 << 20, graph_rounds=8, admit_bytes=256 << 20, admit_rounds=16, poison=False,
                 band_solo=True, copy_stream=True):
        hip.ensure_init(0)
        self.policy = policy
        self.admit_bytes = admit_bytes  # initial-condition bytes staged per step at most
        # rounds run between two such steps (measured on MI355X, 512-problem
        # mix, band method for Euler, 256 MiB chunks: 4 rounds 1.010 s, 8
        # 0.944, 12 0.896, 16 0.893, 24 0.921, 32 0.928)
        self.admit_rounds = admit_rounds
        self._throttled = False
        self.capacity = capacity  # cells per round for "oldest_first" and "short_first"
        self.solo = solo  # a problem running alone uses the plain kernels
        # ... and the band kernel for its standard stages where that outruns
        # the standard kernel (kernels.band_beats_standard)
        self.band_solo = band_solo
        self.band_stages = 0  # stages run on the band kernel (solo problems)
        # host steps whose rounds may be queued on the GPU at once: admitting
        # a request waits for the queued rounds, so this bounds its delay
        self.max_inflight = max_inflight
        self.inflight = []
        self._last_done = None
        self.round_s = None  # measured GPU seconds per round (moving average)
        self.hbm_budget = hbm_budget if hbm_budget is not None else 64 << 30
        self.max_active = max_active
        self.use_graphs = use_graphs
        self.graph_rounds = graph_rounds
        self.poll_rounds = poll_rounds
        self.max_rounds_ahead = max_rounds_ahead
        if tuner == "default":
            from pysweep_amd import autotune
            tuner = autotune.Tuner()  # the stored records, else the measured defaults
        self.tuner = tuner
        self.poison = poison  # debug: rings filled with NaN before each problem's initial condition
        self.division_faults = set()  # class keys whose modules flagged a division-range fault
        self.classes = {}
        self.tables = {}
        self.queue = []
        self.running = []
        self.finished = []
        self.in_use = 0  # device bytes of the running problems' rings
        self.peak_in_use = 0  # peak of in_use plus the pooled free rings
        # the context's free rings, within this engine's budget
        self.rings = _ring_pool(copy_stream)
        self.rings.trim(self.hbm_budget)
        # uploads and downloads on the stream of the context's copy context,
        # beside the rounds (or, copy_stream=False, on the work stream)
        self._copy = _copy_context() if copy_stream else None
        # page-locked staging, made (and touched) here rather than in the
        # first timed rounds; given back when the engine falls idle
        self._ring = _take_ring()
        self._staged = 0  # staging pieces handed out since the last fence
        self._pending = collections.deque()  # (event, pieces, ticket or None, result), in stream order
        self.rounds = 0
        self.clock = clock or time.perf_counter
        self.t0 = self.clock()

    # ---------------------------------------------------------------- API
    def now(self):
        return self.clock() - self.t0

    def submit(self, problem, arrival=None):
        t = Ticket(problem=problem, arrival=self.now() if arrival is None else arrival)
        self.queue.append(t)
        return t

    def idle(self):
        return not self.queue and not self.running and not self._pending

    def run(self, max_rounds=None):
        """Admit and run until every submitted problem has finished (its
        result on the host)."""
        r = 0
        while not self.idle():
            self.step()
            r += 1
            if max_rounds and r >= max_rounds:
                break
        if self.idle():
            self._release_old()
        if self.idle() and self._ring is not None:
            _FREE_RINGS.append(self._ring)
            self._ring = None
        if self.idle():
            self._check_division_range()
        return self.finished

    def _release_old(self):
        """Free the tables' replaced arrays (the engine is idle: no queued
        work reads them)."""
        for tab in self.tables.values():
            while tab.old:
                tab.old.pop().close()

    def _check_division_range(self):
        """Classes whose modules met a state outside the plugin's division
        range (device.range_fault) during the run: their tickets are marked
        ``division_exact = False`` (their last bits may differ from IEEE
        division; solve them again with ``Solver(tuning={"exact_division":
        True})``), a warning names them, and the flags are cleared."""
        for key, cls in self.classes.items():
            if device.range_fault(cls.modules):
                self.division_faults.add(key)
                for t in self.finished:
                    if t.cls is cls:
                        t.division_exact = False
                warnings.warn(f"class {key} met a state outside the plugin's division range: its results may "
                              "differ from IEEE division in the last bits")
                device.range_fault(cls.modules, clear=True)

    def _stage(self, nbytes):
        """Page-locked piece for copies queued before the next fence; waits
        for the oldest fences while the staging ring is full."""
        if self._ring is None:
            self._ring = _take_ring(nbytes)
        while True:
            off = self._ring.alloc(nbytes)
            if off is not None:
                self._staged += 1
                return _Slice(self._ring.buf, off, nbytes)
            if self._pending:
                self._collect(wait=True, most=1)
            elif self._staged:
                self._fence()
            else:  # larger than the ring: a bigger one
                _FREE_RINGS.append(self._ring)
                self._ring = _take_ring(2 * nbytes)

    def _fence(self, ticket=None, result=None):
        """Record an event behind the copies queued so far; at its
        completion their staging pieces are given back and ``ticket`` (whose
        final level is in ``result`` = (piece, dtype, shape)) is finished."""
        if not self._staged and ticket is None:
            return
        ev = hip.Event()
        ev.record()
        self._pending.append((ev, self._staged, ticket, result))
        self._staged = 0

    def _collect(self, wait=False, most=None):
        """Finish what the completed events cover (events complete in
        stream order); with ``wait``, the ``most`` oldest (default all)
        whether complete or not."""
        done = []
        while self._pending and (most is None or len(done) < most):
            ev, pieces, t, res = self._pending[0]
            if wait:
                ev.synchronize()
            elif not ev.query():
                break
            done.append(self._pending.popleft())
        tickets = [(t, res) for _, _, t, res in done if t is not None]
        if len(tickets) > 1:  # the copies out of page-locked memory in parallel (numpy drops the GIL)
            list(_copy_pool().map(_fetch, tickets))
        elif tickets:
            _fetch(tickets[0])
        now = self.now()
        for t, _ in tickets:
            t.finished = now
            self.finished.append(t)
        for _, pieces, _, _ in done:
            self._ring.release(pieces)

    def _rounds_ahead(self, active):
        """Rounds to launch before the host next needs to act: up to the
        next retirement; while requests that have not arrived yet are queued,
        as many as the measured round time fits before the next arrival (at
        least ``poll_rounds``), so they are admitted promptly; while some
        running problems sit out under "short_first", ``poll_rounds`` (its
        selection changes with time; "oldest_first" changes only at
        retirements and arrivals)."""
        r = min(t.c_end - t.c for t in active)
        if self._throttled:  # more requests to admit after a few rounds
            r = min(r, self.admit_rounds)
        elif len(active) < len(self.running) and self.policy == "short_first":
            r = min(r, self.poll_rounds)  # the selection changes with time
        elif self.queue:
            now = self.now()
            if any(t.arrival <= now for t in self.queue):  # waiting for memory: until a retirement
                pass
            else:
                gap = min(t.arrival for t in self.queue) - now
                fit = int(gap / self.round_s) if self.round_s else 0
                r = min(r, max(self.poll_rounds, fit))
        return max(1, min(r, self.max_rounds_ahead))

    @staticmethod
    def _band_rounds(active, R):
        """R rounded up to whole launches while a band-method problem takes
        part: its BandMulti launch on the round that begins one computes
        band_h stages and writes only the last, so every host step leaves
        such a problem at the start of a launch (where the one-problem
        kernels can take it over).  Other problems are unaffected by the
        extra rounds: their kernels skip stages past a problem's last one,
        and the one-problem path stops at it."""
        s = 1
        for t in active:
            if t.method == "band":
                s = math.lcm(s, t.cls.band_h)
        return -(-R // s) * s

    def _run_solo(self, t, R):
        """Rounds c .. c+R-1 of a problem that runs alone, with the plain
        one-problem kernels (the same phases and stages as the batched
        rounds: swept cycles through ps_swept_run, tail stages through
        ps_standard_run), without the batched kernels' per-work-item
        lookups."""
        cls = t.cls
        sw, st, msw, mst, f, fstd = cls.solo_modules(t.nx, t.ny)
        band = cls.band_module(t.nx, t.ny) if self.band_solo else None
        mods = [msw, mst] + ([band[1]] if band else [])
        # Modules are shared by every class, engine and Solver that load the
        # same spec (kernels.load), so the owner of their constants is kept
        # on the modules themselves; any constant write clears it.
        if any(m.const_owner is not t for m in mods):
            cls.cpu.set_globals(*t.problem.globals, source_mod=plugins.SourceModuleFacade(mods))
            for m in mods:
                m.const_owner = t
        G, P, T = t.times.G, cls.P, t.T
        c0, c1 = t.c, min(t.c + R, t.c_end)  # R may run past the last stage (whole band launches)
        if G >= 0 and c0 <= G + 1:
            plan = hip.SweptPlan(state=t.ring.ptr, nx=t.nx, ny=t.ny, b=cls.b, ops=cls.ops, its=cls.its, T=T,
                                 threads=sw.threads, store_all=0, up=f["UpPyramid"], ybridge=f["YBridge"],
                                 xbridge=f["XBridge"], octahedron=f["Octahedron"], down=f["DownPyramid"])
            hip.swept_run(plan, G, c0, min(c1, G + 2))
        lo = max(c0, G + 2 if G >= 0 else 1)
        if c1 > lo:
            done = G + 1
            base = done * P if G >= 0 else 0
            g0, g1 = base + lo - done, base + c1 - done
            plan = hip.StandardPlan(state=t.ring.ptr, nx=t.nx, ny=t.ny, its=cls.its, T=T, bx=st.bx, by=st.by,
                                    threads=st.threads, standard=fstd)
            if band is not None and T >= band[0].levels + 1:
                # the band kernel starts on a full step: stage g reads level
                # s-2+g, a full step when (g - 1) % s == 0
                while (g0 - 1) % cls.its and g0 < g1:
                    hip.standard_run(plan, g0, g0 + 1)
                    g0 += 1
                if g0 < g1:
                    spec = band[0]
                    bplan = hip.BandPlan(state=t.ring.ptr, nx=t.nx, ny=t.ny, its=cls.its, T=T, levels=spec.levels,
                                         wout=kernels.band_out_columns(spec),
                                         rows=kernels.band_rows(spec, t.nx, t.ny), store_all=0, band=band[2])
                    hip.band_run(bplan, cls.its - 2 + g0, g1 - g0)
                    self.band_stages += g1 - g0
            else:
                hip.standard_run(plan, g0, g1)

    def _track(self, R):
        """Bound the queued GPU work (see ``max_inflight``) and keep the
        measured time per round."""
        if not self.max_inflight:
            return
        ev = hip.Event()
        ev.record()
        self.inflight.append((ev, R))
        while len(self.inflight) > self.max_inflight:
            done, rounds = self.inflight.pop(0)
            done.synchronize()
            if self._last_done is not None:
                per = self._last_done.elapsed_ms(done) / 1e3 / rounds
                self.round_s = per if self.round_s is None else 0.7 * self.round_s + 0.3 * per
            self._last_done = done

    def step(self):
        """Admit what fits, run rounds of every class up to the next
        retirement (or admission check), retire finished problems."""
        self._collect()
        self._admit()
        if not self.running:
            self._fence()
            self._collect(wait=True)
            self._release_old()  # every queued round finished before the last download
            if self.queue:
                time.sleep(max(0.0, min(t.arrival for t in self.queue) - self.now()))
            return
        if self.policy == "short_first":
            active = short_first(self.running, self.now(), capacity=self.capacity)
        elif self.policy == "oldest_first":
            active = oldest_first(self.running, self.capacity)
        else:
            active = set(self.running)
        if self.solo and len(active) == 1:
            # one problem: the plain kernels, as a one-at-a-time solve runs it
            t = next(iter(active))
            R = self._band_rounds(active, self._rounds_ahead(active))
            self._run_solo(t, R)
            self.tables[t.cls.key].dirty = True  # its round counter moved on the host only
            self.rounds += R
            self._fence()
            self._track(R)
            t.c += R
            if t.c >= t.c_end:
                self._retire(t)
            self._collect()
            return
        R = self._band_rounds(active, self._rounds_ahead(active))
        for key, tab in self.tables.items():
            if tab.dirty or tab.needs(active):
                tab.rebuild(active, stage=self._stage)
            if tab.plan is None or not any(tab.flags):
                continue
            K = self.graph_rounds
            if self.use_graphs and (tab.graph is not None or R >= K):
                # a graph of K rounds is captured once per table layout and
                # replayed; the remainder (and a layout about to change
                # again within K rounds) runs eagerly
                if tab.graph is None:
                    g = hip.Graph()
                    with g:
                        hip.multi_run(tab.plan, K)
                    tab.graph = g
                for _ in range(R // K):
                    tab.graph.launch()
                if R % K:
                    hip.multi_run(tab.plan, R % K)
            else:
                hip.multi_run(tab.plan, R)
        self.rounds += R
        self._fence()
        self._track(R)
        for t in list(self.running):
            if t not in active:
                continue
            t.c += R
            if t.c >= t.c_end:
                self._retire(t)
        self._collect()

    # ---------------------------------------------------------- internals
    def _class_for(self, p, b):
        key = (p.plugin, p.precision, p.its, b)
        cls = self.classes.get(key)
        if cls is None:
            cfg = self.tuner.config(p.plugin, p.precision, p.its, b) if self.tuner else {}
            cls = PluginClass(p.plugin, p.precision, p.its, b, **cfg)
            self.classes[key] = cls
            self.tables[key] = _ClassTable(cls)
        return cls

    def _block_size(self, p):
        nv, nx, ny = p.ic.shape
        if p.b:
            return p.b
        if self.tuner:
            return self.tuner.block_size(p.plugin, p.precision, p.its, nx, ny)
        for b in (16, 32, 8):
            if nx % b == 0 and ny % b == 0:
                return b
        raise ValueError(f"grid {nx}x{ny} has no supported block size")

    def _admit(self):
        """Start the arrived requests that fit the HBM budget, in arrival
        order; after ``admit_bytes`` of initial conditions, the rest wait
        for the next step, so that the GPU computes the admitted problems'
        rounds while the host stages the next ones."""
        now = self.now()
        pending = sorted((t for t in self.queue if t.arrival <= now), key=lambda t: t.arrival)
        self._throttled = False
        staged = 0
        for t in pending:
            if self.max_active is not None and len(self.running) >= self.max_active:
                break
            if staged >= self.admit_bytes:
                self._throttled = True
                break
            p = t.problem
            b = self._block_size(p)
            cls = self._class_for(p, b)
            nv, nx, ny = p.ic.shape
            method = self._method_for(p, cls)
            need = cls.ring_bytes(nx, ny, method) + _TAIL_BYTES
            if self.in_use + need > self.hbm_budget:
                if not self.running and need > self.hbm_budget:
                    raise MemoryError(f"problem needs {need} B, more than the HBM budget {self.hbm_budget} B")
                break
            self.queue.remove(t)
            self._start(t, cls, b, need, method)
            staged += p.ic.size * cls.rb

    def _method_for(self, p, cls):
        """The kernel family a problem runs on ("auto": the tuner's record,
        else the default table; the band method only where the class has
        it, else the standard one)."""
        if p.method != "auto":
            return p.method
        m = (self.tuner.method(p.plugin, p.precision, p.its) if self.tuner
             else DEFAULT_METHOD.get((p.plugin, p.precision), "swept"))
        return "standard" if m == "band" and not cls.band_multi_ok else m

    def _start(self, t, cls, b, need, method):
        p = t.problem
        nv, nx, ny = p.ic.shape
        if nv != cls.nv or nx % b or ny % b:
            raise ValueError(f"problem shape {p.ic.shape} does not fit class {cls.key}")
        N = schedule.requested_steps(p.globals[0], p.globals[1], p.globals[2])
        if method not in ("swept", "standard", "band"):
            raise ValueError(f"unknown method {method!r} (swept, standard, band or auto)")
        if method == "band" and not cls.band_multi_ok:
            raise ValueError(f"method='band' needs a point update or a plugin band sweep of one step per "
                             f"launch (kernels.band_multi_capable); class {cls.key} has neither")
        t.method = method
        if method in ("standard", "band"):
            t.times = schedule.standard_times(N, b, cls.ops, cls.its)
        else:
            t.times = schedule.swept_times(N, b, cls.ops, cls.its, exact=p.exact)
        t.nx, t.ny = nx, ny
        T = t.T = cls.levels(method)
        lvl = nv * nx * ny * cls.rb
        hc = _HostConsts(cls.layout)
        cls.cpu.set_globals(*p.globals, source_mod=hc)
        assert len(hc.buf) <= 256
        # one buffer: the ring, then the constant block (lvl is a multiple
        # of 256 bytes: nx and ny are multiples of b >= 8)
        t.ring = self.rings.take(need, limit=self.hbm_budget - self.in_use)
        t.consts = t.ring.ptr + T * lvl
        t.sums = t.consts + 256
        # levels 0 .. its-1 hold the initial condition: one upload from a
        # staging buffer, then device copies
        host = self._stage(lvl + 256)
        _copy_into(host.view(cls.dtype, (nv, nx, ny)), p.ic)
        host.view(numpy.uint8, (lvl + 256,))[lvl:lvl + len(hc.buf)] = numpy.frombuffer(bytes(hc.buf), numpy.uint8)
        if self._copy is not None:
            self._copy.push()
        try:
            if self.poison:  # every level NaN until written (stream ordered, before the uploads)
                t.ring.fill_bytes(0xFF, T * lvl)
            host.upload_async(t.ring.ptr, lvl)
            host.upload_async(t.consts, len(hc.buf), lvl)
            if self._copy is not None:
                up = hip.Event()
                up.record()
        finally:
            if self._copy is not None:
                self._copy.pop()
        if self._copy is not None:
            # everything the work stream does with the problem (and every
            # fence recorded there later) follows its upload
            up.wait()
        for A in range(1, cls.its):
            hip.dtod_async(t.ring.ptr + (A % T) * lvl, t.ring.ptr, lvl)
        hip.level_sums_async(t.ring.ptr, cls.rb == 4, nv, nx * ny, t.sums)
        G = t.times.G
        t.c = 0 if G >= 0 else 1
        t.c_end = G + 2 + t.times.tail
        t.final_level = cls.its - 1 + t.times.stages + t.times.tail
        t.cls = cls
        t.nbytes = t.ring.nbytes
        t.cells = nx * ny
        t.admitted = t.last_active = self.now()
        self.in_use += t.nbytes
        self.peak_in_use = max(self.peak_in_use, self.in_use + self.rings.bytes)
        self.running.append(t)
        tab = self.tables[cls.key]
        tab.running.append(t)
        tab.dirty = True
        if t.c >= t.c_end:  # nothing to run
            self._retire(t)

    def _retire(self, t):
        """Queue the copy of the final level to the host (the ticket is
        finished when it has arrived) and return the ring to the pool: later
        work that reuses it runs after the copy (on the work stream, or on
        the copy stream, where the next problem's upload into it follows
        the copy)."""
        cls = t.cls
        nv = cls.nv
        T = t.T
        lvl = nv * t.nx * t.ny * cls.rb
        self._fence()  # what is staged so far completes on its own
        fin = t.ring.ptr + (t.final_level % T) * lvl
        hip.level_sums_async(fin, cls.rb == 4, nv, t.nx * t.ny, t.sums + 128)
        piece = self._stage(lvl + 256)  # the level, then its sums (lvl is a multiple of 256)
        if self._copy is None:
            piece.download_async(fin, lvl)
            piece.download_async(t.sums, 256, lvl)
            self._fence(t, (piece, cls.dtype, (nv, t.nx, t.ny)))
        else:
            # on the copy stream, after the problem's last round and sums
            last = hip.Event()
            last.record()
            with self._copy:
                last.wait()
                piece.download_async(fin, lvl)
                piece.download_async(t.sums, 256, lvl)
                self._fence(t, (piece, cls.dtype, (nv, t.nx, t.ny)))  # its event on the copy stream
        self.rings.give(t.ring)
        t.ring, t.consts, t.sums = None, 0, 0
        self.in_use -= t.nbytes
        self.running.remove(t)
        tab = self.tables[cls.key]
        tab.running.remove(t)
        tab.dirty = True


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * q / 100.0
    f, c = math.floor(k), math.ceil(k)
    return xs[f] if f == c else xs[f] + (xs[c] - xs[f]) * (k - f)
