"""``Solver``: the reference's user-facing object (core/solver.py:9-189) on
one MI355X.

Same construction and attribute API as the reference:

    solver = Solver(initialConditions, "input.yaml")   # or set attributes
    solver()                                           # run

Attributes (core/io.py:35-75): simulation (True = swept), verbose,
intermediate, operating, blocksize (b, b, 1), share, exid, globals
[t0, tf, dt, dx, dy, *equation args], cpu / gpu plugin paths, output,
dtypeStr / dtype.  The MPI ranks (core/process.py) collapse to one process:
with share = 1 the whole grid is on the GPU; with share < 1 (swept only) the
GPU takes the first ceil(share * nx/b) block rows and the host the rest,
computed with the plugin's Python step (pysweep_amd.share).

Additions: ``exact`` (take exactly N steps instead of the reference's
quantization to multiples of the pyramid height), ``history`` (keep every
full-step level, the reference's output dataset ``data``), ``result`` (final
full-step state), ``solve_seconds`` and ``gcups``; ``tuning`` (launch shape:
slots / threads / bx / by / pfd; band_levels / band_rows for the band
kernel, which runs the swept solve of a plugin with a point update (heat),
or with a band sweep of its own (Euler) on grids of at least 2^17 cells;
``band=True`` or ``phases=True`` to run it on the band kernel or the five
phase kernels regardless; ``generic=True`` to call the plugin's
per-cell step even when it provides a cooperative level update, and
``exact_division=True`` to build the plugin with EU_EXACT_DIVISION, and
``poison=True`` to fill the device ring with NaN before the initial
condition goes in, so that a read of a cell no phase wrote shows up as NaN);
``conservation`` (the FP64 sums of the first and final levels).

Division precondition.  The Euler plugin forms its quotients and square
roots from refined hardware reciprocals; they are bitwise the IEEE ones for
states with rho, E and p in [2^-120, 2^120] and each momentum 0 or of
magnitude in [2^-400, 2^120].  The plugin flags any state outside that range
it meets on the device (EU_RANGE_FAULT); the Solver then solves again with
IEEE division and square root (``division_rerun`` is True afterwards and a
warning says so), so results always follow IEEE arithmetic.
"""
import os
import time
import warnings

import numpy

from pysweep_amd import config, plugins, schedule


class Solver(object):
    """Swept / standard solver of a plugin pair on one GPU."""

    def __init__(self, initialConditions=None, yamlFileName=None, sendWarning=True):
        super().__init__()
        self.moments = [time.time()]
        # single process: rank 0 is node master and cluster master
        self.rank = 0
        self.nodeMasterBool = self.clusterMasterBool = True
        self.simulation = True
        self.verbose = False
        self.intermediate = 1
        self.operating = 1
        self.blocksize = (32, 32, 1)
        self.share = 1.0
        self.exid = []
        self.globals = []
        self.cpu = self.cpuStr = None
        self.gpu = self.gpuStr = None
        self.output = "output.hdf5"
        self.dtypeStr = "float64"
        self.dtype = numpy.dtype(self.dtypeStr)
        self.exact = False
        self.history = False
        self.write_output = False
        self.logfile = None
        self.tuning = {}
        self.result = None
        self.data = None
        self.solve_seconds = None
        self.assignInitialConditions(initialConditions, sendWarning=sendWarning)
        if yamlFileName is not None:
            self.yamlFileName = yamlFileName
            config.yamlManager(self)
        elif sendWarning:
            warnings.warn("yaml not specified, requires manual input.")

    # ------------------------------------------------------------------ API
    def __call__(self, start=0, stop=-1, libname=None, recompile=False):
        """Run the solve (core/solver.py:25-70)."""
        self.moments.append(time.time())
        self._verbose("-----------------------------PySweep-------------------------\n")
        self._check_inputs()
        self.createTimeStepData()
        self.moments.append(time.time())
        self._setup_device()
        self.moments.append(time.time())
        self._verbose(str(self))
        self._solve()
        self.division_rerun = False
        if self._division_range_fault():
            # a state outside the range where the Euler plugin's fast
            # quotients are the IEEE ones: the same solve with IEEE division
            warnings.warn("the state left the range in which the plugin's fast division is exact "
                          "(EU_RANGE_FAULT); solving again with IEEE division and square root")
            saved = dict(self.tuning)
            self.tuning["exact_division"] = True
            self._run.close()  # the fast build's ring; the IEEE build gets its own
            try:
                self._setup_device()
                self._solve()
            finally:
                self.tuning = saved
            self.division_rerun = True
        self.moments.append(time.time())
        self._finish(self.moments[start], self.moments[stop])

    def _solve(self):
        if self.simulation:
            self.sweptSolve()
        else:
            self.standardSolve()

    def _division_range_fault(self):
        """Whether the GPU solve met a state outside the plugin's division
        range (never for plugins without the check, nor for host shares)."""
        if self.share < 1 or self.tuning.get("exact_division"):
            return False
        return self._run.range_fault()

    def __str__(self):
        from pysweep_amd import reporting
        return reporting.solver_print(self)

    def compactPrint(self):
        """core/solver.py:76-81."""
        short = "swept: " if self.simulation else "standard: "
        moment = self.moments[-1] - self.moments[0] if self.moments else None
        short += "{}, {}, {}, {}, {}, {}".format(self.blocksize[0], self.share, self.intermediate, self.operating,
                                                  getattr(self, "timeSteps", None), moment)
        print(short)

    def loadCPUModule(self):
        """Import the plugin's Python half (core/solver.py:83-87)."""
        self.cpu = plugins.load_cpu_module(self.cpu)

    def assignInitialConditions(self, initialConditions, sendWarning=True):
        """Array (nv, nx, ny) or HDF5 file with dataset ``data`` (core/solver.py:89-101)."""
        if isinstance(initialConditions, str):
            from pysweep_amd import hdf5io
            self.initialConditions = hdf5io.read_initial_conditions(initialConditions)
            self.arrayShape = self.initialConditions.shape
        elif initialConditions is not None:
            self.initialConditions = numpy.asarray(initialConditions)
            self.arrayShape = numpy.shape(initialConditions)
        elif sendWarning:
            warnings.warn("Initial condition not provided, requires manual assignment with "
                          "assignInitialConditionsFunction.")

    def setGPU(self, gpu):
        self.gpu = gpu
        self.gpuStr = gpu

    def setCPU(self, cpu):
        self.cpu = cpu
        self.cpuStr = cpu

    def createTimeStepData(self):
        """core/solver.py:113-128, plus the exact-step option."""
        t0, tf, dt = self.globals[0], self.globals[1], self.globals[2]
        self.timeSteps = schedule.requested_steps(t0, tf, dt)
        nv, nx, ny = self.initialConditions.shape
        self.globalShape = (nv, nx, ny)
        s = self.intermediate
        if self.simulation:
            b = self.blocksize[0]
            self.splitx = self.splity = b // 2
            self.times = schedule.swept_times(self.timeSteps, b, self.operating, s, exact=self.exact)
            self.maxPyramidSize = self.times.P
            self.maxGlobalSweptStep = self.times.G
            self.maxOctSize = 2 * self.times.P
            self.subtraction = self.times.subtraction
            self.sharedShape = (self.times.window, nv, nx, ny)
            self.arrayShape = (self.times.nt, nv, nx, ny)
            self.stages = self.times.stages + self.times.tail
        else:
            self.sharedShape = (s + 1, nv, nx, ny)
            self.arrayShape = (self.timeSteps + 1, nv, nx, ny)
            self.stages = s * self.timeSteps

    def sweptSolve(self):
        """FirstPrism, G UpPrism cycles, LastPrism (core/solver.py:158-176)."""
        run = self._run
        collect = self._collector()
        if self.share < 1:
            t0 = time.perf_counter()
            self.final_level = run.run(self.times, on_level=collect)
            self.solve_seconds = time.perf_counter() - t0
            self.globalTimeStep = self.final_level - (self.intermediate - 1) + 1
            return
        from pysweep_amd import hip
        start, stop = hip.Event(), hip.Event()
        start.record()
        self.final_level = run.run(self.times, on_level=collect)
        stop.record()
        self.solve_seconds = start.elapsed_ms(stop) / 1e3
        self.globalTimeStep = self.final_level - (self.intermediate - 1) + 1

    def standardSolve(self):
        """One stage per launch (core/solver.py:178-189)."""
        from pysweep_amd import hip
        run = self._run
        collect = self._collector()
        start, stop = hip.Event(), hip.Event()
        start.record()
        run.run(1, self.stages + 1, on_level=collect)
        stop.record()
        self.solve_seconds = start.elapsed_ms(stop) / 1e3
        self.final_level = self.intermediate - 1 + self.stages
        self.globalTimeStep = self.stages + 1

    # ------------------------------------------------------------ internals
    def _verbose(self, msg):
        if self.verbose:
            print(msg)

    def _check_inputs(self):
        if self.initialConditions is None:
            raise ValueError("no initial conditions")
        if len(self.initialConditions.shape) != 3:
            raise ValueError("initial conditions must be an array (nv, nx, ny)")
        if not 0 <= self.share <= 1:
            raise ValueError("share must be in [0, 1]")
        if self.share < 1 and not self.simulation:
            raise NotImplementedError("the CPU/GPU share split runs the swept decomposition; "
                                      "use share=1 for the standard one")
        if self.dtype != numpy.float64:
            raise NotImplementedError("device state is float64 (set dtype: float64)")
        b = self.blocksize[0]
        assert b % (2 * self.operating) == 0, ("Invalid blocksize, blocksize must satisfy BS = 2*ops*k and the "
                                               "architectural limit where k is any integer factor.")
        nv, nx, ny = self.initialConditions.shape
        if self.simulation:
            assert nx % b == 0 and ny % b == 0, \
                "Provided array dimensions is not divisible by the specified block size."
        if isinstance(self.cpu, str):
            self.loadCPUModule()
        if self.cpu is None or self.gpu is None:
            raise ValueError("cpu and gpu plugin files must be set")

    def _setup_device(self):
        nv, nx, ny = self.initialConditions.shape
        self._sums0 = None
        t = dict(self.tuning)
        if self.simulation and self.share < 1:
            # CPU/GPU split (core/process.py:81-176): block rows divided by share
            from pysweep_amd import share
            if self.share > 0:
                from pysweep_amd import hip
                hip.ensure_init(0)
            self._run = share.SharedSweptRun(self.cpu, self.gpu, nv, self.operating, self.intermediate, nx, ny,
                                             self.blocksize[0], self.share, slots=t.get("slots"),
                                             threads=t.get("threads"), generic=t.get("generic", False))
            self.share = self._run.share  # effective share (process.py:148)
            self.cpu.set_globals(*self.globals)
            self._run.set_globals(self.cpu, self.globals)
            self._run.upload_ic(self.initialConditions)
            return
        from pysweep_amd import device, hip
        hip.ensure_init(0)
        defines = ("EU_EXACT_DIVISION",) if t.get("exact_division") else ()
        from pysweep_amd import kernels
        capable = self.simulation and not t.get("generic") and \
            kernels.band_capable(self.gpu, nv, self.operating, self.intermediate)
        if t.get("band") and not t.get("phases") and self.simulation and not capable:
            warnings.warn("tuning band=True: this plugin (or the per-cell step contract) has no band form; "
                          "the swept solve runs on the phase kernels")
        if t.get("phases") or not capable:
            band = False
        elif t.get("band"):
            band = True
        else:
            band = kernels.band_default(self.gpu, nv, self.operating, self.intermediate, nx, ny)
        if band:
            self._run = device.BandRun(self.gpu, nv, self.operating, self.intermediate, nx, ny,
                                       levels=t.get("band_levels"), rows=t.get("band_rows"), defines=defines)
        elif self.simulation:
            self._run = device.SweptRun(self.gpu, nv, self.operating, self.intermediate, nx, ny,
                                        self.blocksize[0], slots=t.get("slots"), threads=t.get("threads"),
                                        tail_standard=True, generic=t.get("generic", False), pfd=t.get("pfd"),
                                        defines=defines)
        else:
            self._run = device.StandardRun(self.gpu, nv, self.operating, self.intermediate, nx, ny,
                                           bx=t.get("bx"), by=t.get("by"), threads=t.get("threads"),
                                           generic=t.get("generic", False), defines=defines)
        # CPU globals, then the plugin's device constants (core/block.py:50, 90)
        self.cpu.set_globals(*self.globals)
        self._run.set_globals(self.cpu, self.globals)
        if t.get("poison"):
            self._run.poison()
        self._run.upload_ic(self.initialConditions)
        self._run.range_fault(clear=True)
        # FP64 sums of the conserved quantities at the start (device-side)
        self._sums0 = self._run.sums(self.intermediate - 1)

    def _collector(self):
        s = self.intermediate
        if not (self.history or self.write_output):
            return None
        nt = self.arrayShape[0]
        self.data = numpy.empty(self.arrayShape)
        self.data[0] = self.initialConditions
        run = self._run

        def on_level(A):
            m = (A - (s - 1)) // s
            if m < nt:
                run.level(A, out=self.data[m])

        return on_level

    def _finish(self, t_start, t_stop):
        s = self.intermediate
        last_full = self.final_level - ((self.final_level - (s - 1)) % s)
        self.result = self._run.level(last_full)
        self.result_step = (last_full - (s - 1)) // s
        self.result_level = self.final_level
        self.final_state = self._run.level(self.final_level)
        # the final level's conserved sums against the initial ones (GPU-only runs)
        self.conservation = None
        if getattr(self, "_sums0", None) is not None:
            from pysweep_amd import device
            self.conservation = device.conservation(self._sums0, self._run.sums(self.final_level))
        full_steps = self.stages / s
        nv, nx, ny = self.globalShape
        self.gcups = nx * ny * full_steps / self.solve_seconds / 1e9 if self.solve_seconds else None
        self.runtime = t_stop - t_start
        if self.write_output and self.data is not None:
            from pysweep_amd import hdf5io
            hdf5io.write_output(self, self.output)
        if self.logfile:
            from pysweep_amd import reporting
            reporting.update_log_file(self, self.runtime, self.logfile)
        self._verbose("Done in {} seconds...\n".format(self.moments[-1] - self.moments[0]))
