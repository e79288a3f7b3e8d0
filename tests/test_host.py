"""CPU tests of the host side: the C-ABI library exports, run-time
compilation (no GPU needed), plugin loading and the device facade, the YAML
input, the schedule arithmetic and the per-instance constant blocks."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import swept as o_swept
from pysweep_amd import compiler, kernels, plugins, schedule
from pysweep_amd.equations import path as eq_path

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/pysweep/equations"


def _header_symbols():
    with open(os.path.join(ROOT, "include", "pysweep_hip.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(ps_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from pysweep_amd import hip
    declared = _header_symbols()
    assert declared, "no declarations found"
    lib = ctypes.CDLL(hip.LIBPATH)
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    assert sorted(hip.EXPORTS) == declared


def test_device_count_without_gpu_is_not_an_error():
    from pysweep_amd import hip
    assert hip.device_count() >= 0


@pytest.mark.parametrize("name,nv,k,s", [("heat.cu", 1, 1, 1), ("euler.cu", 4, 2, 2), ("euler_fp32.cu", 4, 2, 2)])
def test_shipped_plugins_compile_for_gfx950(name, nv, k, s):
    real = "float" if "fp32" in name else "double"
    for spec in (kernels.swept_spec(eq_path(name), nv, k, s, 16, real=real),
                 kernels.swept_spec(eq_path(name), nv, k, s, 32, real=real),
                 kernels.standard_spec(eq_path(name), nv, k, s, 64, 64, real=real),
                 kernels.swept_spec(eq_path(name), nv, k, s, 16, real=real, multi=True)):
        code = kernels.build_module_code(spec)
        assert code[:4] == b"\x7fELF"
        for kernel, size in _private_segment_sizes(code).items():
            assert size == 0, f"{spec.describe()}: {kernel} uses {size} B of scratch memory"


@pytest.mark.parametrize("name,s,real", [("heat.cu", 1, "double"), ("heat.cu", 2, "double"),
                                         ("heat_fp32.cu", 1, "float")])
def test_band_kernels_compile_without_scratch(name, s, real):
    """The band kernel keeps every level's history in registers: it must
    compile to a code object with no scratch memory (a spill would put the
    history in per-lane memory)."""
    code = kernels.build_module_code(kernels.band_spec(eq_path(name), 1, 1, s, real=real))
    assert code[:4] == b"\x7fELF"
    sizes = _private_segment_sizes(code)
    assert not os.path.exists(READELF) or sizes.get("Band") == 0, sizes


READELF = "/opt/rocm/llvm/bin/llvm-readelf"


def _private_segment_sizes(code):
    """Scratch bytes per kernel from the code object's metadata note."""
    import subprocess
    import tempfile
    if not os.path.exists(READELF):
        return {}
    with tempfile.NamedTemporaryFile(suffix=".hsaco") as f:
        f.write(code)
        f.flush()
        out = subprocess.run([READELF, "--notes", f.name], capture_output=True, text=True).stdout
    return {m.group(1): int(m.group(2)) for m in
            re.finditer(r"\.name:\s+(\w+).*?\.private_segment_fixed_size:\s+(\d+)", out, re.S)}


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not present")
@pytest.mark.parametrize("name,nv,k,s", [("heat.cu", 1, 1, 1), ("euler.cu", 4, 2, 2), ("example.cu", 1, 1, 1),
                                         ("checker.cu", 1, 1, 1), ("half.cu", 1, 1, 2)])
def test_reference_plugin_sources_compile_unchanged(name, nv, k, s, tmp_path):
    """The reference's device plugins build unchanged behind the harness
    (compile only; hiprtc needs no GPU)."""
    spec = kernels.swept_spec(os.path.join(REF, name), nv, k, s, 16)
    code = compiler.compile_cached(spec.source(), name, cache_dir=str(tmp_path))
    assert code[:4] == b"\x7fELF"


def test_compile_error_is_reported():
    from pysweep_amd import hip
    with pytest.raises(hip.HipError, match="error"):
        hip.compile_source("this is not C++;", "bad.hip", list(compiler.OPTIONS))


def test_constant_layout_and_block_macros():
    src = open(eq_path("heat.cu")).read()
    lay = compiler.constant_layout(src)
    names = [f.name for f in lay.fields]
    assert names == ["DX", "DY", "DT", "ALPHA", "COURANT", "SCHEME"]
    assert [f.offset for f in lay.fields] == [0, 8, 16, 24, 32, 40]
    assert lay.size == 48
    pre, body, _ = compiler.multi_prelude_and_source(src)
    assert "__constant__" not in "".join(l for l in body.splitlines() if not l.strip().startswith("//"))
    assert "#define COURANT (ps_pc.COURANT)" in pre


def test_host_constant_block_facade():
    from pysweep_amd.batch import _HostConsts
    lay = compiler.constant_layout(open(eq_path("heat.cu")).read())
    hc = _HostConsts(lay)
    mod = plugins.load_cpu_module(eq_path("heat.py"))
    mod.set_globals(0.0, 1.0, 0.01, 0.1, 0.2, 0.5, False, source_mod=hc)
    vals = np.frombuffer(bytes(hc.buf[:40]), dtype=np.float64)
    assert list(vals) == [0.1, 0.2, 0.01, 0.5, 0.5 * 0.01 / 0.1 / 0.1]
    assert np.frombuffer(bytes(hc.buf[40:44]), dtype=np.intc)[0] == 0


def test_plugin_loader_requires_contract(tmp_path):
    bad = tmp_path / "bad.py"
    bad.write_text("def set_globals(*a, source_mod=None):\n    pass\n")
    with pytest.raises(AttributeError, match="step"):
        plugins.load_cpu_module(str(bad))


def test_plugin_loader_injects_driver(tmp_path):
    p = tmp_path / "p.py"
    p.write_text("try:\n    import pycuda.driver as cuda\nexcept Exception:\n    pass\n"
                 "def set_globals(*a, source_mod=None):\n    pass\n"
                 "def step(state, iidx, t, g):\n    pass\n")
    mod = plugins.load_cpu_module(str(p))
    assert mod.cuda is plugins.driver


@pytest.mark.parametrize("N,b,k,s", [(100, 8, 1, 1), (500, 16, 2, 2), (1000, 32, 1, 1), (1000, 32, 2, 2),
                                     (7, 16, 2, 2), (1, 32, 2, 2)])
def test_schedule_matches_reference_arithmetic(N, b, k, s):
    ref = o_swept.time_step_data(0.0, float(N), 1.0, b, k, s)
    got = schedule.swept_times(N, b, k, s, exact=False) if ref["G"] >= 0 else None
    if got is not None:
        assert (got.P, got.G, got.stages, got.nt, got.subtraction) == \
            (ref["P"], ref["G"], ref["stages"], ref["nt"], ref["subtraction"])
    ex = schedule.swept_times(N, b, k, s, exact=True)
    assert ex.stages + ex.tail == s * N and ex.stages % ex.P == 0 and ex.tail < ex.P + s * N * (ex.G < 0)


def test_phase_sets_match_oracle():
    for b, k in ((8, 1), (16, 2), (32, 2), (24, 1)):
        P = b // (2 * k) - 1
        for kind in schedule.PHASES:
            L = 2 * P if kind == "Octahedron" else P
            for r in range(1, L + 1):
                assert schedule.phase_set(kind, r, b, k, P) == o_swept.phase_rect(kind, r, b, k, P)
        per_cycle = sum(schedule.phase_cell_stages(kd, b, k, P) for kd in ("XBridge", "Octahedron", "YBridge"))
        assert per_cycle == P * b * b


def test_yaml_input(tmp_path):
    from pysweep_amd import Solver
    y = tmp_path / "in.yaml"
    y.write_text(f"""
swept: True
blocksize: 16
share: 1
dtype: float64
globals: [0, 0.5, 0.1, 0.1, 0.1, 1.4]
intermediate_steps: 2
operating_points: 2
cpu: {eq_path('euler.py')}
gpu: {eq_path('euler.cu')}
""")
    s = Solver(np.zeros((4, 32, 32)), str(y))
    assert s.blocksize == (16, 16, 1) and s.intermediate == 2 and s.operating == 2
    assert callable(s.cpu.step) and s.gpu == eq_path("euler.cu")
    s.createTimeStepData()
    assert s.timeSteps == 5 and s.maxPyramidSize == 3


def test_yaml_rejects_bad_blocksize(tmp_path):
    from pysweep_amd import Solver
    y = tmp_path / "in.yaml"
    y.write_text("blocksize: 10\noperating_points: 2\n")
    with pytest.raises(AssertionError, match="blocksize"):
        Solver(np.zeros((1, 8, 8)), str(y))


def test_solver_rejects_out_of_range_share():
    from pysweep_amd import Solver
    s = Solver(np.zeros((1, 16, 16)), sendWarning=False)
    s.share = 1.5
    s.blocksize = (8, 8, 1)
    s.globals = [0, 1, 0.1, 0.1, 0.1, True]
    s.setCPU(eq_path("example.py"))
    s.setGPU(eq_path("example.cu"))
    with pytest.raises(ValueError, match="share"):
        s()


@pytest.mark.parametrize("name", ["heat", "euler", "example", "checker", "half"])
def test_cpu_plugin_halves_follow_oracle(name):
    """The plugins' Python step (the share=0 contract) on a periodic block."""
    from oracle import euler as o_euler
    from oracle import heat as o_heat
    from oracle import plugins as o_pl
    mod = plugins.load_cpu_module(eq_path(name + ".py"))
    n = 16
    rng = np.random.default_rng(5)
    iidx = [(i, j) for i in range(n) for j in range(n)]
    if name == "euler":
        q = o_euler.initial_condition(n, eps=3.0)
        st = np.stack([q, q, q])
        mod.set_globals(0.0, 1.0, 0.05, 10 / n, 10 / n, 1.4)
        mod.step(st, iidx, 1, 1)
        want = o_euler.step_py(q, q, 1, 1.4, 0.05, 10 / n, 10 / n)
    elif name == "heat":
        u = rng.uniform(-1, 1, (3, 1, n, n))
        mod.set_globals(0.0, 1.0, 1e-4, 1 / n, 1 / n, 1.0, True)
        st = u.copy()
        mod.step(st, iidx, 1, 1)
        want = o_heat.step_py(u[1, 0], None, 1, o_heat.courant_py(1.0, 1e-4, 1 / n), True)[None]
    elif name == "example":
        u = rng.uniform(-1, 1, (3, 1, n, n))
        mod.set_globals(0, 1, 0.1, 0.1, 0.1, False)
        st = u.copy()
        mod.step(st, iidx, 1, 2)
        want = o_pl.example_step(u[1], u[0], 2, False)
    elif name == "checker":
        u = rng.uniform(-1, 1, (3, 1, n, n))
        mod.set_globals(0, 1, 0.1, 0.1, 0.1, True)
        st = u.copy()
        mod.step(st, iidx, 1, 1)
        want = o_pl.checker_step(u[1], u[0], 1)
    else:
        u = rng.uniform(-1, 1, (3, 1, n, n))
        mod.set_globals(0, 1, 0.1, 0.1, 0.1)
        st = u.copy()
        mod.step(st, iidx, 1, 2)
        want = o_pl.half_step(u[1], u[0], 2)
    assert np.array_equal(st[2], want)


def test_float_reciprocal_division_is_exact():
    """The kernels' ps_div / eu_div / heat_div: q = int(float(t) * rw)
    corrected once is t // w for 0 <= t < 2^22, for any rw within one ulp of
    1/w (the hardware reciprocal the kernels use is accurate to one ulp);
    restated in float32 arithmetic with the correctly rounded reciprocal and
    its two float neighbours."""
    rng = np.random.default_rng(3)
    for w in list(range(1, 300)) + [1023, 1024, 1025, 4096, 65535]:
        t = np.concatenate([np.arange(0, 4 * w + 3), (1 << 22) - 1 - np.arange(4 * w),
                            rng.integers(0, 1 << 22, 2000)]).astype(np.int64)
        r0 = np.float32(1.0) / np.float32(w)
        for rw in (r0, np.nextafter(r0, np.float32(0)), np.nextafter(r0, np.float32(1))):
            q = (t.astype(np.float32) * rw).astype(np.int64)
            r = t - q * w
            q = np.where(r < 0, q - 1, np.where(r >= w, q + 1, q))
            assert np.array_equal(q, t // w), (w, rw)


def test_prefetch_register_bound_covers_every_face_load():
    """PS_PF * threads >= the most face cells any swept phase loads after a
    level (counted by brute force over the oracle's phase sets)."""
    from oracle import swept as sw
    from pysweep_amd import compiler
    for b, k, s in [(8, 1, 1), (16, 2, 2), (32, 1, 1), (32, 2, 2), (16, 1, 2), (24, 3, 2), (12, 1, 3)]:
        P = b // (2 * k) - 1
        most = 0
        for kind in (sw.UP, sw.YBRIDGE, sw.XBRIDGE, sw.OCT, sw.DOWN):
            L = sw.phase_levels(kind, P)
            for l in range(1, L):
                need = np.zeros((b, b), dtype=bool)
                x0, x1, y0, y1 = sw.phase_rect(kind, l + 1, b, k, P)
                need[max(x0 - k, 0):x1 + k, max(y0 - k, 0):y1 + k] = True
                for r in range(l + 2, min(l + s, L) + 1):
                    x0, x1, y0, y1 = sw.phase_rect(kind, r, b, k, P)
                    lo = np.argwhere(need).min(0)
                    hi = np.argwhere(need).max(0) + 1
                    need[min(lo[0], x0):max(hi[0], x1), min(lo[1], y0):max(hi[1], y1)] = True
                x0, x1, y0, y1 = sw.phase_rect(kind, l, b, k, P)
                need[x0:x1, y0:y1] = False
                most = max(most, int(need.sum()))
        assert compiler._face_cells(b, k, s) == most, (b, k, s)


def test_oldest_first_selection():
    """The engine's "oldest_first" round selection: arrival order while the
    cells fit the capacity, the oldest problem always in."""
    from pysweep_amd import batch

    def ticket(n, arrival, admitted=0.0):
        return batch.Ticket(problem=None, cells=n * n, arrival=arrival, admitted=admitted)

    big, s1, s2, s3 = ticket(2048, 0.0), ticket(256, 1.0), ticket(256, 2.0), ticket(512, 3.0)
    cap = 1 << 19
    # the oldest is in even though it alone exceeds the capacity; nothing else fits beside it
    assert batch.oldest_first([s2, big, s1], cap) == {big}
    # without it: s1, s2 (2 x 65536 cells) fit, s3 (262144) too, in arrival order
    assert batch.oldest_first([s3, s2, s1], cap) == {s1, s2, s3}
    assert batch.oldest_first([s3, s2, s1], 3 * 65536) == {s1, s2}
    # selection stops at the first that does not fit (no skipping ahead of older problems)
    huge = ticket(1024, 1.5)
    assert batch.oldest_first([s3, huge, s2, s1], cap) == {s1}
    # ties in arrival: admission order
    a, b = ticket(128, 5.0, admitted=5.1), ticket(128, 5.0, admitted=5.0)
    assert batch.oldest_first([a, b], 128 * 128) == {b}
    assert batch.oldest_first([], cap) == set()


def test_buffer_pool_best_fit_and_limit():
    """The engine's pool of freed rings / staging buffers: smallest free
    buffer within the slack, a new one otherwise, free buffers released
    (largest first) so that they and the new one fit the limit."""
    from pysweep_amd import batch

    class Buf:
        def __init__(self, n):
            self.nbytes, self.closed = n, False

        def close(self):
            self.closed = True

    made = []

    def make(n):
        made.append(Buf(n))
        return made[-1]

    pool = batch._Pool(make, slack=1.25)
    x1, x2, x3 = pool.take(100), pool.take(120), pool.take(500)
    assert len(made) == 3 and pool.bytes == 0
    for x in (x1, x2, x3):
        pool.give(x)
    assert pool.bytes == 720
    assert pool.take(110) is x2            # 100 < 110; 120 <= 137.5
    assert pool.take(90) is x1             # smallest >= 90 within 112.5
    assert pool.bytes == 500
    y = pool.take(200)                     # nothing within [200, 250]: new
    assert y is made[-1] and y.nbytes == 200 and pool.bytes == 500
    # a new 300 under a limit of 600 frees the 500 (free + new must fit)
    z = pool.take(300, limit=600)
    assert x3.closed and pool.bytes == 0 and z.nbytes == 300
    pool.give(y)
    pool.give(z)
    pool.trim(250)                         # largest first: z (300) goes
    assert z.closed and not y.closed and pool.bytes == 200


def test_short_first_selection():
    """The engine's round selection (pysweep_amd.batch.short_first): the
    shortest remaining work first, packed up to the capacity in cells; a
    problem left out for max_wait seconds joins every round from then on.
    The remaining-work arithmetic follows the round structure (swept rounds
    cover (G+1)P stages, then single standard stages)."""
    from pysweep_amd import batch, schedule

    def ticket(times, n, c=0):
        t = batch.Ticket(problem=None, times=times, cells=n * n, c=c)
        t.c_end = times.G + 2 + times.tail
        return t

    big = ticket(schedule.swept_times(1000, 16, 2, 2, exact=True), 2048)
    small = ticket(schedule.standard_times(100, 16, 2, 2), 128, c=1)
    twin = ticket(schedule.standard_times(150, 16, 2, 2), 128, c=1)
    assert small.remaining_work == 200 * 128 * 128
    assert big.remaining_work == (big.times.stages + big.times.tail) * 2048 * 2048
    big.c = big.times.G + 1  # at LastPrism
    assert big.remaining_work == (big.times.P / 2 + big.times.tail) * 2048 * 2048
    big.c = 0
    now = 10.0
    for t in (big, small, twin):
        t.last_active = now
    assert batch.short_first([big, small, twin], now) == {small, twin}
    assert batch.short_first([big, small, twin], now + 1.0) == {small, twin}
    # big has now sat out for max_wait: it joins, and stays in
    assert batch.short_first([big, small, twin], now + 2.5) == {small, twin, big}
    assert big.boosted and batch.short_first([big, small, twin], now + 2.6) == {small, twin, big}
    # alone, the long problem always runs
    lone = ticket(schedule.swept_times(1000, 16, 2, 2, exact=True), 2048)
    assert batch.short_first([lone], now) == {lone}
    assert batch.short_first([], now) == set()


def test_staging_ring_first_in_first_out():
    """The engine's page-locked staging ring: pieces are handed out in order,
    wrap to the start when the end is taken, and never overlap a live piece;
    the oldest pieces are given back first."""
    from pysweep_amd import batch

    class Buf:  # stands in for the page-locked buffer
        def __init__(self, n):
            self.a = np.zeros(n, np.uint8)
            self.ptr = 0

        def view(self, dtype, shape):
            return self.a.view(dtype).reshape(shape)

    A = batch._StagingRing.ALIGN
    r = batch._StagingRing(8 * A, make=Buf)
    assert r.alloc(9 * A) is None            # larger than the ring
    assert r.alloc(3 * A) == 0
    assert r.alloc(A + 1) == 3 * A           # rounded up to two units
    assert r.alloc(3 * A) == 5 * A           # up to the end
    assert r.alloc(1) is None                # full
    r.release(1)                             # [0, 3A) free again
    assert r.alloc(4 * A) is None            # does not fit before the oldest live piece
    assert r.alloc(2 * A) == 0               # wraps
    assert r.alloc(A) == 2 * A               # fills [2A, 3A): full again
    assert r.alloc(1) is None
    r.release(2)                             # [3A, 8A) free; live: [0, 3A)
    assert r.alloc(5 * A) == 3 * A
    r.release(3)
    assert list(r.live) == []
    assert r.alloc(8 * A) == 0               # empty ring: from the start

    # random traffic against a byte map of the live pieces
    rng = np.random.default_rng(3)
    r = batch._StagingRing(64 * A, make=Buf)
    owner = np.full(64 * A, -1)
    live = []
    for i in range(2000):
        if live and (rng.uniform() < 0.45 or len(live) > 20):
            k = int(rng.integers(1, len(live) + 1))
            r.release(k)
            for off, n in live[:k]:
                owner[off:off + n] = -1
            live = live[k:]
            continue
        n = int(rng.integers(1, 20 * A))
        off = r.alloc(n)
        if off is None:
            continue
        m = -(-n // A) * A
        assert off % A == 0 and off + m <= 64 * A
        assert (owner[off:off + m] == -1).all(), i
        owner[off:off + m] = i
        live.append((off, m))
    assert [tuple(x) for x in r.live] == live


def test_context_api_without_a_current_context():
    """The per-thread context stack: nothing is current on a fresh thread, a
    pop with nothing pushed is an error with a message, and (without a GPU)
    creating a context reports that no device is visible."""
    import threading

    from pysweep_amd import hip
    seen = {}

    def probe():
        seen["current"] = hip.current_context()
        try:
            hip.check(hip._ctx_pop(), "ps_context_pop")
        except hip.HipError as e:
            seen["pop"] = str(e)

    t = threading.Thread(target=probe)
    t.start()
    t.join()
    assert seen["current"] is None
    assert "no context is current" in seen["pop"]
    if hip.device_count() == 0:
        with pytest.raises(hip.HipError, match="no HIP device"):
            hip.Context(0)


def test_workgroups_per_cu_bounds():
    """Resident workgroups per CU from LDS bytes, VGPRs and lanes (MI355X:
    160 KB of LDS, 4 SIMDs of 8 waves and 512 VGPRs per lane)."""
    from pysweep_amd.autotune import workgroups_per_cu
    assert workgroups_per_cu(28 * 1024, 64, 256) == 5       # LDS-bound
    assert workgroups_per_cu(8 * 1024, 128, 256) == 4       # 4 waves per SIMD by VGPRs
    assert workgroups_per_cu(8 * 1024, 60, 512) == 4        # 32 waves per CU, 8 per workgroup
    assert workgroups_per_cu(130 * 1024, 40, 512) == 1


def test_conservation_record():
    """The conserved-sum record: per variable |final - initial| over the
    initial sum of magnitudes, and their maximum as the drift."""
    import numpy as np

    from pysweep_amd.device import conservation
    c = conservation((np.array([10.0, -2.0]), np.array([20.0, 4.0])), (np.array([10.5, -2.0]),))
    assert c["relative_change"] == [0.025, 0.0] and c["drift"] == 0.025
    assert c["initial"] == [10.0, -2.0] and c["final"] == [10.5, -2.0]
    z = conservation((np.array([0.0]), np.array([0.0])), (np.array([1e-20]),))
    assert z["drift"] == 1e-20  # an all-zero field: the absolute change


def test_tuner_records_and_shapes(tmp_path):
    """Tuner records: defaults without a file, launch shapes only for the
    block size they were tuned with, the method of method="auto"
    problems, and the candidate shapes as class_specs arguments."""
    import json

    from pysweep_amd import batch
    from pysweep_amd.autotune import CANDIDATES, Tuner, shape_of
    t = Tuner(path=None)
    assert t.block_size("heat", "fp64", 1, 96, 64) == 32 and t.block_size("euler", "fp64", 2, 96, 96) == 16
    assert t.block_size("heat", "fp64", 1, 48, 80) == 16
    assert t.method("euler", "fp64", 2) == batch.DEFAULT_METHOD[("euler", "fp64")]
    path = tmp_path / "tuning.json"
    path.write_text(json.dumps({"euler/fp64/its2": {"b": 16, "method": "standard", "stile": [8, 32],
                                                    "sthreads": 512}}))
    t = Tuner(path=str(path))
    assert t.config("euler", "fp64", 2, 16) == {"stile": (8, 32), "sthreads": 512}
    assert t.config("euler", "fp64", 2, 32) == {}
    assert t.method("euler", "fp64", 2) == "standard"
    for plugin, its in (("euler", 2), ("heat", 1)):
        for cand in CANDIDATES[plugin]:
            sp, ssp = batch.class_specs(plugin, "fp64", its, cand["b"], **shape_of(cand))
            if "stile" in cand:
                assert (ssp.bx, ssp.by, ssp.threads) == (*cand["stile"], cand["sthreads"])
            if "slots" in cand:
                assert sp.slots == cand["slots"]


def test_verbose_summary_names_the_swept_kernel():
    """The solve summary says which swept kernel ran (no GPU needed: a
    stand-in run object carries the module spec)."""
    import types

    from pysweep_amd import Solver, reporting
    s = Solver(np.zeros((4, 64, 64)), sendWarning=False)
    s.globals = [0.0, 0.1, 0.01, 0.1, 0.1, 1.4]
    s.cpuStr, s.gpuStr = "euler.py", "euler.cu"
    assert "swept, blocksize" in reporting.solver_print(s)
    spec = kernels.band_spec(eq_path("euler.cu"), 4, 2, 2)
    s._run = types.SimpleNamespace(spec=spec, rows=10)
    assert "band kernel, 2 levels per launch, 10 rows per wavefront" in reporting.solver_print(s)
    s._run = types.SimpleNamespace(spec=kernels.swept_spec(eq_path("euler.cu"), 4, 2, 2, 16), rows=None)
    assert "(phase kernels)" in reporting.solver_print(s)


def test_band_forms_and_launch_shapes():
    """Which plugins have a band form, their bands' output columns, and the
    row segments the wave-balance model picks for the measured workloads
    (Euler 1024^2 and 2048^2, heat 8192^2: the segments the MI355X
    measurements found best or within 1 % of it)."""
    heat, euler = eq_path("heat.cu"), eq_path("euler.cu")
    assert kernels.band_capable(heat, 1, 1, 1) and kernels.band_capable(heat, 1, 1, 2)
    assert kernels.band_capable(euler, 4, 2, 2)
    assert not kernels.band_capable(eq_path("example.cu"), 1, 1, 1)
    h = kernels.band_spec(heat, 1, 1, 1)
    e = kernels.band_spec(euler, 4, 2, 2)
    assert (h.levels, kernels.band_out_columns(h)) == (12, 104)
    assert (e.levels, kernels.band_out_columns(e)) == (2, 56)
    assert kernels.band_spec(eq_path("heat_fp32.cu"), 1, 1, 1, real="float").levels == 8
    assert kernels.band_rows(e, 1024, 1024) == 10
    assert kernels.band_rows(e, 2048, 2048) == 40
    assert kernels.band_rows(h, 8192, 8192) == 216
    for n in (3, 5, 64, 130):
        assert 1 <= kernels.band_rows(e, n, n) <= max(n, 4)
    with pytest.raises(ValueError):
        kernels.band_spec(euler, 4, 2, 2, levels=16)  # 64 - 2*2*16 < 8 output columns
    src = h.source()
    assert src.index("ps_band_grid") < src.index("ps_point") < src.index("Band(PS_REAL")


def test_abi_structs_match_the_header(tmp_path):
    """Every ctypes mirror of a C-ABI struct has the size and field offsets
    of the declaration in include/pysweep_hip.h (compiled with gcc)."""
    import shutil
    import subprocess
    from pysweep_amd import hip
    if shutil.which("gcc") is None:
        pytest.skip("gcc not found")
    mirrors = {"ps_grid": hip.Grid, "ps_swept_plan": hip.SweptPlan, "ps_standard_plan": hip.StandardPlan,
               "ps_band_plan": hip.BandPlan, "ps_inst": hip.Inst, "ps_multi_plan": hip.MultiPlan,
               "ps_run_stats": hip.RunStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "pysweep_hip.h"', "int main(void) {"]
    for cname, cls in mirrors.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        cname, field, value = line.split()
        got[(cname, field)] = int(value)
    for cname, cls in mirrors.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)


def test_band_multi_rows_and_whole_step_rounds():
    """The instance-table band launch: each problem's row segments cover its
    grid in equal parts; a batch with more wavefronts takes longer segments;
    only a plugin band sweep of one step per launch has the form; host steps
    run whole steps while a band-method problem takes part."""
    from types import SimpleNamespace
    from pysweep_amd import batch
    spec = kernels.band_spec(eq_path("euler.cu"), 4, 2, 2, multi=True)
    assert spec.multi and spec.threads == 64 and spec.levels == 2
    shapes = [(2048, 2048), (64, 96), (100, 56), (1, 8)]
    for (nx, ny), r in zip(shapes, kernels.band_multi_rows(spec, shapes)):
        segs = -(-nx // r)
        assert 1 <= r <= nx and (segs - 1) * r < nx <= segs * r, (nx, r)
    assert kernels.band_multi_rows(spec, []) == []
    few = kernels.band_multi_rows(spec, [(1024, 1024)])[0]
    many = kernels.band_multi_rows(spec, [(1024, 1024)] * 16)[0]
    assert many > few
    with pytest.raises(ValueError, match="band form"):
        kernels.band_spec(eq_path("example.cu"), 1, 1, 1, multi=True)
    with pytest.raises(ValueError, match="one step"):
        kernels.band_spec(eq_path("euler.cu"), 4, 2, 2, levels=4, multi=True)
    assert not kernels.band_multi_capable(eq_path("euler.cu"), 4, 2, 1)  # a band sweep takes RK2 steps
    assert kernels.band_multi_capable(eq_path("euler_fp32.cu"), 4, 2, 2)
    assert kernels.band_multi_capable(eq_path("heat.cu"), 1, 1, 1) and kernels.band_multi_capable(eq_path("heat.cu"), 1, 1, 2)
    assert kernels.band_multi_levels(eq_path("heat.cu"), 1) == 12 and kernels.band_multi_levels(eq_path("euler.cu"), 2) == 2
    assert kernels.band_multi_levels(eq_path("heat_fp32.cu"), 1, "float") == 8
    heat = kernels.band_spec(eq_path("heat.cu"), 1, 1, 1, multi=True)
    assert heat.levels == 12 and heat.threads == 64
    assert all(r >= 1 for r in kernels.band_multi_rows(heat, [(64, 64), (1024, 96), (8, 200)]))
    band = SimpleNamespace(method="band", cls=SimpleNamespace(its=2, band_h=2))
    std = SimpleNamespace(method="standard", cls=SimpleNamespace(its=2, band_h=2))
    hband = SimpleNamespace(method="band", cls=SimpleNamespace(its=1, band_h=12))
    h8 = SimpleNamespace(method="band", cls=SimpleNamespace(its=1, band_h=8))
    assert batch.Engine._band_rounds([std], 3) == 3
    assert batch.Engine._band_rounds([std, band], 3) == 4
    assert batch.Engine._band_rounds([band], 6) == 6
    assert batch.Engine._band_rounds([band, hband], 5) == 12
    assert batch.Engine._band_rounds([hband, h8], 5) == 24  # whole launches of both
    assert any(sp.kind == "band" and sp.multi for sp in batch.shipped_class_specs())


def test_plugin_source_cache_follows_edits(tmp_path):
    """A plugin source (and its includes) is read again once a file it was
    read from has changed, and so are the macro tests on it."""
    inc = tmp_path / "part.cu"
    top = tmp_path / "plug.cu"
    inc.write_text("// nothing yet\n")
    top.write_text('#include "part.cu"\n__device__ void step(double *s, int i, int g) {}\n')
    assert not kernels.has_band_sweep(str(top))
    first = kernels.read_plugin_source(str(top))
    assert kernels.read_plugin_source(str(top)) is first  # unchanged files: the kept text
    st = os.stat(inc)
    inc.write_text("#define PS_HAS_BAND 1\n")
    os.utime(inc, ns=(st.st_atime_ns, st.st_mtime_ns + 10 ** 9))
    assert "PS_HAS_BAND" in kernels.read_plugin_source(str(top))
    assert kernels.has_band_sweep(str(top))


def test_auto_method_falls_back_where_the_class_has_no_band_form():
    """method="auto" resolves to the default table's method, and to the
    standard one where that is "band" but the class cannot run it."""
    from types import SimpleNamespace
    from pysweep_amd import batch
    eng = SimpleNamespace(tuner=None)
    p = SimpleNamespace(method="auto", plugin="euler", precision="fp64", its=2)
    assert batch.DEFAULT_METHOD[("euler", "fp64")] == "band"
    assert batch.Engine._method_for(eng, p, SimpleNamespace(band_multi_ok=True)) == "band"
    assert batch.Engine._method_for(eng, p, SimpleNamespace(band_multi_ok=False)) == "standard"
    heat = SimpleNamespace(method="auto", plugin="heat", precision="fp64", its=1)
    assert batch.Engine._method_for(eng, heat, SimpleNamespace(band_multi_ok=False)) == "standard"
    fixed = SimpleNamespace(method="swept", plugin="euler", precision="fp64", its=2)
    assert batch.Engine._method_for(eng, fixed, SimpleNamespace(band_multi_ok=True)) == "swept"
