"""The CPU/GPU share split of a swept solve (core/process.py:81-176;
core/functions.py:12-88).

The grid's block rows are divided as the reference divides them: the GPU
takes the first ceil(share * nx/b) rows of blocks, the host the rest
(share is updated to the effective fraction, process.py:145-148).  Both
sides keep a ring of all levels of the whole grid -- the GPU's in HBM, the
host's in memory -- and compute the tiles of their own block rows:

* a cycle's tiles start at rows xoff + I*b with xoff = 0 or b/2, so with
  xoff = b/2 the GPU's tiles cover rows [b/2, gx + b/2) and the host's
  [gx + b/2, nx + b/2) (periodic).  A tile reads only its own rows, so the
  two sides never read each other's rows within a cycle, except in the two
  bands [gx, gx + b/2) and [0, b/2) whose owner alternates with xoff;
* each cycle runs on both sides at once (the GPU's phases are queued
  asynchronously while the host computes its own), then the two bands are
  copied from whichever side computed them to the other.  This replaces the
  reference's window roll between ranks (functions.py:90-123);
* the host computes a phase level by level with the plugin's own Python
  ``step(state, iidx, arrayTimeIndex, globalTimeStep)`` over the cells of
  its tiles' active sets, the reference's CPU path (core/geometry.py:26-36);
* stages after the swept cycles (exact-step mode) run on the GPU after the
  host's rows are copied over.

With share = 0 there is no GPU part at all (a pure host solve).
"""
import numpy

from pysweep_amd import schedule


class SharedSweptRun:
    """Swept solve of one plugin over a GPU part and a host part."""

    def __init__(self, cpu_module, gpu_path, nv, ops, its, nx, ny, b, share, ring_dtype=numpy.float64,
                 slots=None, threads=None, generic=False):
        rows = nx // b
        self.gpu_rows = int(numpy.ceil(rows * share))  # block rows on the GPU
        self.share = self.gpu_rows / rows
        self.nv, self.ops, self.its, self.nx, self.ny, self.b = nv, ops, its, nx, ny, b
        self.P = b // (2 * ops) - 1
        self.T = 2 * self.P + its
        self.cpu = cpu_module
        self.host = numpy.zeros((self.T, nv, nx, ny), dtype=ring_dtype)
        self.gx = self.gpu_rows * b
        self.gpu = None
        if self.gpu_rows:
            from pysweep_amd import device
            self.gpu = device.SweptRun(gpu_path, nv, ops, its, nx, ny, b, slots=slots, threads=threads,
                                       tail_standard=True, generic=generic, tile_rows=self.gpu_rows)
        self.modules = self.gpu.modules if self.gpu else []

    # ------------------------------------------------------------ set-up
    def set_globals(self, cpu_module, globals_):
        if self.gpu:
            self.gpu.set_globals(cpu_module, globals_)
        else:
            cpu_module.set_globals(*globals_)

    def upload_ic(self, ic):
        ic = numpy.asarray(ic, dtype=self.host.dtype)
        for A in range(self.its):
            self.host[A % self.T] = ic
        if self.gpu:
            self.gpu.upload_ic(ic)

    # ------------------------------------------------------------ host side
    def _cells(self, kind, r, xoff, yoff):
        """Global (i, j) of the host's tiles' active set of level r."""
        b, nx, ny = self.b, self.nx, self.ny
        x0, x1, y0, y1 = schedule.phase_set(kind, r, b, self.ops, self.P)
        rows = numpy.arange(self.gpu_rows, nx // b)
        cols = numpy.arange(ny // b)
        li, lj = numpy.meshgrid(numpy.arange(x0, x1), numpy.arange(y0, y1), indexing="ij")
        gi = (xoff + rows[:, None, None, None] * b + li[None, None]) % nx
        gj = (yoff + cols[None, :, None, None] * b + lj[None, None]) % ny
        gi, gj = numpy.broadcast_arrays(gi, gj)
        return numpy.stack([gi.ravel(), gj.ravel()], axis=1)

    def _host_phase(self, kind, A0, gts0, xoff, yoff):
        L = 2 * self.P if kind == "Octahedron" else self.P
        T = self.T
        for r in range(1, L + 1):
            ij = self._cells(kind, r, xoff, yoff)
            if not len(ij):
                continue
            state = numpy.stack([self.host[(A0 + r - 2) % T], self.host[(A0 + r - 1) % T],
                                 self.host[(A0 + r) % T]])
            self.cpu.step(state, ij, 1, gts0 + r - 1)
            ii, jj = ij[:, 0], ij[:, 1]
            self.host[(A0 + r) % T][:, ii, jj] = state[2][:, ii, jj]

    # ------------------------------------------------------------ exchange
    def _band_rows(self, xoff):
        """(rows computed by the GPU, rows computed by the host) among the
        two shared bands for a cycle with lattice offset xoff."""
        h = self.b // 2
        s1 = slice(self.gx, self.gx + h)   # [gx, gx + b/2)
        s2 = slice(0, h)                   # [0, b/2)
        return ([s1], [s2]) if xoff else ([s2], [s1])

    def _exchange(self, xoff):
        if not self.gpu or self.gx == self.nx:
            return
        g_rows, h_rows = self._band_rows(xoff)
        lvl = self.gpu.level_bytes
        rb = self.host.dtype.itemsize
        for A in range(self.T):
            for sl in g_rows:
                band = numpy.empty((self.nv, sl.stop - sl.start, self.ny), dtype=self.host.dtype)
                for v in range(self.nv):
                    off = A * lvl + (v * self.nx * self.ny + sl.start * self.ny) * rb
                    self.gpu.ring.download(band[v], off)
                self.host[A][:, sl] = band
            for sl in h_rows:
                band = numpy.ascontiguousarray(self.host[A][:, sl])
                for v in range(self.nv):
                    off = A * lvl + (v * self.nx * self.ny + sl.start * self.ny) * rb
                    self.gpu.ring.upload(band[v], off)

    # ------------------------------------------------------------ run
    def run(self, times, on_level=None):
        """Swept cycles on both sides, then the exact-step tail on the GPU
        (or the host with share = 0).  Returns the final absolute level."""
        G, P, s = times.G, self.P, self.its
        done = s - 1
        if self.gpu:
            # recorded levels must be complete on the GPU side too
            self.gpu.plan.store_all = 1 if on_level is not None else 0
        if self.gpu:
            from pysweep_amd import hip
        for c in range(G + 2 if G >= 0 else 0):
            if self.gpu:
                hip.swept_run(self.gpu.plan, G, c, c + 1)
            for kind, A0, gts0, xoff, yoff in schedule.cycle_phases(c, G, P, s, self.b):
                self._host_phase(kind, A0, gts0, xoff, yoff)
            if self.gpu:
                hip.synchronize()
                self._exchange((c % 2) * (self.b // 2))
            hi = schedule.complete_after_cycle(c, G, P, s)
            if on_level is not None:
                for A in schedule.full_step_levels(s, done, hi):
                    on_level(A)
            done = hi
        if times.tail:
            done = self._tail(times, done, on_level)
        return done

    def _tail(self, times, done, on_level):
        s = self.its
        g0 = done - (s - 1) + 1
        if self.gpu:
            from pysweep_amd import hip
            # the host's rows of every ring level go to the GPU, which finishes
            rb = self.host.dtype.itemsize
            lvl = self.gpu.level_bytes
            for A in range(self.T):
                for v in range(self.nv):
                    rows = numpy.ascontiguousarray(self.host[A][v, self.gx:])
                    self.gpu.ring.upload(rows, A * lvl + (v * self.nx * self.ny + self.gx * self.ny) * rb)
            self.tail_on_gpu = True  # from here on the GPU ring holds every row
            for g in range(g0, g0 + times.tail):
                hip.standard_run(self.gpu.tail, g, g + 1)
                A = s - 1 + g
                if on_level is not None and (A - (s - 1)) % s == 0:
                    on_level(A)
        else:
            ij = numpy.stack(numpy.meshgrid(numpy.arange(self.nx), numpy.arange(self.ny), indexing="ij"),
                             axis=-1).reshape(-1, 2)
            T = self.T
            for g in range(g0, g0 + times.tail):
                A = s - 2 + g
                state = numpy.stack([self.host[(A - 1) % T], self.host[A % T], self.host[(A + 1) % T]])
                self.cpu.step(state, ij, 1, g)
                self.host[(A + 1) % T] = state[2]
                if on_level is not None and (A + 1 - (s - 1)) % s == 0:
                    on_level(A + 1)
        return done + times.tail

    def level(self, A, out=None):
        """Level A as float64: the GPU's rows from HBM, the host's from its
        ring (after a GPU tail, everything from HBM)."""
        if self.gpu and (getattr(self, "tail_on_gpu", False) or self.gx == self.nx):
            return self.gpu.level(A, out)
        lvl = self.host[A % self.T].astype(numpy.float64)
        if self.gpu:
            g = self.gpu.level(A)
            lvl[:, :self.gx] = g[:, :self.gx]
        if out is None:
            return lvl
        out[...] = lvl
        return out

    def close(self):
        if self.gpu:
            self.gpu.close()
