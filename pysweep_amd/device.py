"""Device-resident swept and standard solves for one plugin on one GPU.

The state lives in HBM as a ring ``state[T][nv][nx][ny]`` (float64, y
contiguous); absolute level A is slot A % T.  Nothing is copied between
phases: the reference's host gather / H2D / D2H / scatter per phase group
(core/functions.py:21-33) and its window roll and shift (functions.py:90-123,
io.py:134-144) are replaced by ring indexing and a per-cycle lattice offset
inside the kernels.
"""
import ctypes

import numpy

from pysweep_amd import hip, kernels, plugins, schedule


def range_fault(modules, clear=False):
    """Whether a plugin that checks its division range (the Euler plugin's
    EU_RANGE_FAULT) met a state outside it in any of ``modules`` since the
    last clear; ``clear`` resets the flags instead (both wait for the work
    stream).  Modules without the flag never report one."""
    hit = False
    for m in modules:
        try:
            ptr, _ = m.get_global("EU_RANGE_FAULT")
        except hip.HipError:
            continue
        if clear:
            hip.htod(ptr, numpy.zeros(1, numpy.int32))
        else:
            v = numpy.zeros(1, numpy.int32)
            hip.dtoh(v, ptr)
            hit = hit or bool(v[0])
    return hit


class _Run:
    def __init__(self, nv, its, nx, ny, T, real="double"):
        self.nv, self.its, self.nx, self.ny, self.T = nv, its, nx, ny, T
        self.dtype = numpy.float64 if real == "double" else numpy.float32
        self.level_bytes = nv * nx * ny * numpy.dtype(self.dtype).itemsize
        self.ring = hip.DeviceArray(T * self.level_bytes)
        self.modules = []

    def facade(self):
        return plugins.SourceModuleFacade(self.modules)

    def set_globals(self, cpu_module, globals_):
        """The plugin writes its device constants into every module."""
        cpu_module.set_globals(*globals_, source_mod=self.facade())

    def poison(self):
        """Fill every level of the ring with NaN (all bytes 0xFF, a NaN in
        FP64 and in FP32): a debug mode in which a read of a cell that no
        phase wrote -- a skipped private cell, a face outside the sets --
        turns the result into NaN instead of passing unnoticed."""
        self.ring.fill_bytes(0xFF)

    def upload_ic(self, ic):
        ic = numpy.ascontiguousarray(ic, dtype=self.dtype)
        if ic.shape != (self.nv, self.nx, self.ny):
            raise ValueError(f"initial condition shape {ic.shape} != {(self.nv, self.nx, self.ny)}")
        for A in range(self.its):
            self.ring.upload(ic, (A % self.T) * self.level_bytes)

    def keep_ic_on_device(self, ic):
        """Device copy of the initial condition for stream-ordered resets."""
        ic = numpy.ascontiguousarray(ic, dtype=self.dtype)
        self.ic_dev = hip.DeviceArray(ic.nbytes)
        self.ic_dev.upload(ic)

    def reset(self):
        """Levels 0..s-1 <- initial condition, on the work stream (no host copy)."""
        n = self.level_bytes // 8  # copied as 8-byte words
        for A in range(self.its):
            hip.copy_f64(self.ring.ptr + (A % self.T) * self.level_bytes, self.ic_dev.ptr, n)

    def level(self, A, out=None):
        """Level A as float64 (written into ``out`` when given)."""
        buf = numpy.empty((self.nv, self.nx, self.ny), dtype=self.dtype)
        self.ring.download(buf, (A % self.T) * self.level_bytes)
        if out is None:
            return buf.astype(numpy.float64, copy=False)
        out[...] = buf
        return out

    def range_fault(self, clear=False):
        """Whether the plugin met a state outside its division range in any
        of the run's modules since the last clear (``range_fault``)."""
        return range_fault(self.modules, clear)

    def sums(self, A):
        """FP64 (sums, sums of magnitudes) of level A's variables, summed on
        the device (for FP32 state too)."""
        return hip.level_sums(self.ring.ptr + (A % self.T) * self.level_bytes, self.dtype == numpy.float32,
                              self.nv, self.nx * self.ny)

    def close(self):
        self.ring.close()
        ic = getattr(self, "ic_dev", None)
        if ic is not None:
            ic.close()


def conservation(initial, final):
    """Conserved-sum record of a solve from the device sums of its first and
    final levels (each a (sums, sums of magnitudes) pair): per variable the
    change of the sum relative to the initial sum of magnitudes, and the
    largest of those as ``drift``.  On the periodic grid both plugins
    conserve every sum (heat content; mass, momentum and energy) up to
    rounding, so the drift measures the rounding of the stored state."""
    s0, m0 = initial
    s1 = final[0]
    rel = [float(abs(b - a) / m) if m > 0 else float(abs(b - a)) for a, b, m in zip(s0, s1, m0)]
    return {"initial": [float(x) for x in s0], "final": [float(x) for x in s1], "relative_change": rel,
            "drift": max(rel)}


class StandardRun(_Run):
    """Standard decomposition: one launch of ``Standard`` per stage."""

    def __init__(self, cu_path, nv, ops, its, nx, ny, bx=None, by=None, threads=None, T=None, real="double",
                 generic=False, defines=()):
        super().__init__(nv, its, nx, ny, T or its + 1, real)
        self.spec = kernels.standard_spec(cu_path, nv, ops, its, nx, ny, bx=bx, by=by, threads=threads, real=real,
                                          generic=generic, defines=defines)
        mod = kernels.load(self.spec)
        self.modules = [mod]
        self.plan = hip.StandardPlan(state=self.ring.ptr, nx=nx, ny=ny, its=its, T=self.T, bx=self.spec.bx,
                                     by=self.spec.by, threads=self.spec.threads,
                                     standard=mod.get_function("Standard"))

    def run(self, g_begin, g_end, on_level=None, stats=None):
        """Stages g_begin..g_end-1; ``on_level(A)`` after each full step;
        ``stats`` (hip.RunStats) accumulates per-launch event timings."""
        if stats is not None:
            hip.standard_run_timed(self.plan, g_begin, g_end, stats)
            return
        if on_level is None:
            hip.standard_run(self.plan, g_begin, g_end)
            return
        for g in range(g_begin, g_end):
            hip.standard_run(self.plan, g, g + 1)
            A = self.its - 1 + g
            if (A - (self.its - 1)) % self.its == 0:
                on_level(A)


class SweptRun(_Run):
    """Swept decomposition: UpPyramid, bridges, Octahedron, DownPyramid."""

    def __init__(self, cu_path, nv, ops, its, nx, ny, b, slots=None, threads=None, tail_standard=True,
                 real="double", generic=False, tile_rows=0, pad=None, pfd=None, defines=()):
        P = b // (2 * ops) - 1
        if b % (2 * ops) or P < 1:
            raise ValueError(f"invalid blocksize {b} for operating_points {ops}")
        if nx % b or ny % b:
            raise ValueError(f"grid {nx}x{ny} is not divisible by the blocksize {b}")
        super().__init__(nv, its, nx, ny, 2 * P + its, real)
        self.P, self.b, self.ops = P, b, ops
        self.spec = kernels.swept_spec(cu_path, nv, ops, its, b, slots=slots, threads=threads, real=real,
                                       generic=generic, pad=pad, pfd=pfd, defines=defines)
        mod = kernels.load(self.spec)
        self.modules = [mod]
        f = {n: mod.get_function(n) for n in kernels.SWEPT_KERNELS}
        self.plan = hip.SweptPlan(state=self.ring.ptr, nx=nx, ny=ny, b=b, ops=ops, its=its, T=self.T,
                                  threads=self.spec.threads, tile_rows=tile_rows, up=f["UpPyramid"],
                                  ybridge=f["YBridge"],
                                  xbridge=f["XBridge"], octahedron=f["Octahedron"], down=f["DownPyramid"])
        self.tail = None
        if tail_standard:
            sspec = kernels.standard_spec(cu_path, nv, ops, its, nx, ny, real=real, generic=generic, defines=defines)
            smod = kernels.load(sspec)
            self.modules.append(smod)
            self.tail = hip.StandardPlan(state=self.ring.ptr, nx=nx, ny=ny, its=its, T=self.T, bx=sspec.bx,
                                         by=sspec.by, threads=sspec.threads, standard=smod.get_function("Standard"))

    def run(self, times, on_level=None, stats=None):
        """Run the swept cycles of ``times`` (schedule.SweptTimes) and then
        its tail stages; ``on_level(A)`` for every full-step level as soon as
        it is complete; ``stats`` (hip.RunStats) accumulates per-launch event
        timings.  Returns the final absolute level."""
        G, P, s = times.G, self.P, self.its
        done = s - 1
        # levels handed to on_level must be complete in the ring; otherwise
        # the phases skip the cells that no later phase reads
        self.plan.store_all = 1 if on_level is not None else 0
        if G >= 0:
            if stats is not None and on_level is None:
                hip.swept_run_timed(self.plan, G, 0, G + 2, stats)
                done = schedule.complete_after_cycle(G + 1, G, P, s)
            elif on_level is None:
                hip.swept_run(self.plan, G, 0, G + 2)
                done = schedule.complete_after_cycle(G + 1, G, P, s)
            else:
                for c in range(G + 2):
                    hip.swept_run(self.plan, G, c, c + 1)
                    hi = schedule.complete_after_cycle(c, G, P, s)
                    for A in schedule.full_step_levels(s, done, hi):
                        on_level(A)
                    done = hi
        if times.tail:
            if self.tail is None:
                raise RuntimeError("exact-step tail needs the standard kernel")
            g0 = done - (s - 1) + 1
            if stats is not None and on_level is None:
                hip.standard_run_timed(self.tail, g0, g0 + times.tail, stats)
            for g in (range(g0, g0 + times.tail) if on_level is not None or stats is None else ()):
                hip.standard_run(self.tail, g, g + 1)
                A = s - 1 + g
                if on_level is not None and (A - (s - 1)) % s == 0:
                    on_level(A)
            done += times.tail
        return done


class BandRun(_Run):
    """Swept solve through the band kernel, for a plugin with a point update
    (heat) or a band sweep of its own (Euler): every cell advances
    ``levels`` levels per launch with the intermediate levels in registers,
    the blocks' sloped edges recomputed from a halo instead of exchanged
    through bridge phases (band.inc).  The same stages as ``SweptRun`` for
    the same schedule, on a ring of levels + s levels; results bitwise
    equal."""

    def __init__(self, cu_path, nv, ops, its, nx, ny, levels=None, rows=None, real="double", defines=()):
        self.spec = kernels.band_spec(cu_path, nv, ops, its, real=real, levels=levels, defines=defines)
        super().__init__(nv, its, nx, ny, self.spec.levels + its, real)
        mod = kernels.load(self.spec)
        self.modules = [mod]
        self.rows = int(rows or kernels.band_rows(self.spec, nx, ny))
        if self.rows < 1:
            raise ValueError("band rows must be positive")
        self.plan = hip.BandPlan(state=self.ring.ptr, nx=nx, ny=ny, its=its, T=self.T, levels=self.spec.levels,
                                 wout=kernels.band_out_columns(self.spec), rows=self.rows, store_all=0,
                                 band=mod.get_function("Band"))

    def run(self, times, on_level=None, stats=None):
        """The ``times.stages + times.tail`` stages of a swept schedule
        (schedule.SweptTimes) from level s-1; ``on_level(A)`` for every
        full-step level as soon as it is complete; ``stats`` (hip.RunStats)
        accumulates per-launch event timings.  Returns the final level."""
        s = self.its
        A, total = s - 1, times.stages + times.tail
        if on_level is None:
            self.plan.store_all = 0
            hip.band_run(self.plan, A, total, stats)
            return A + total
        self.plan.store_all = 1
        done = 0
        while done < total:
            hh = min(self.spec.levels, total - done)
            hip.band_run(self.plan, A, hh, stats)
            for lv in schedule.full_step_levels(s, A, A + hh):
                on_level(lv)
            A += hh
            done += hh
        return A


def launch_info(run):
    """Registers / LDS of the loaded kernels (for reports)."""
    out = {}
    for m in run.modules:
        for name in ("UpPyramid", "Octahedron", "Standard", "Band"):
            try:
                out[name] = hip.function_info(m.get_function(name))
            except Exception:
                pass
    return out
