"""Kernel module specifications, launch-shape defaults and loading.

A ``ModuleSpec`` fixes everything compiled into one hiprtc module: the
plugin's device source, the stencil geometry (NV, OPS, ITS), and the tile
shape / LDS window / workgroup size.  ``build_module_code`` compiles (or
fetches from the cache) without a GPU; ``load`` needs one.
"""
import dataclasses
import os
import re

from pysweep_amd import compiler
from pysweep_amd.equations import path as eq_path

#: LDS per workgroup the default shapes aim for; 160 KB per CU leaves room
#: for two workgroups (eight wavefronts) per CU.
LDS_BUDGET = 64 * 1024
LDS_MAX = 160 * 1024


_INCLUDE_RE = re.compile(r'^[ \t]*#[ \t]*include[ \t]+"([^"]+)"[ \t]*$', re.M)


_SOURCES = {}  # absolute path -> ((file, mtime) of every file read, expanded text)


def _mtime(path):
    try:
        return os.stat(path).st_mtime_ns
    except OSError:
        return None


_MACROS = {}  # (absolute path, macro) -> (source text it was found in, answer)


def _defines_macro(path, name):
    """Whether the plugin source (with its includes) defines macro ``name``
    on a line of its own; kept with the source text it was read from."""
    text = read_plugin_source(path)
    key = (os.path.abspath(path), name)
    hit = _MACROS.get(key)
    if hit is None or hit[0] is not text:
        hit = _MACROS[key] = (text, re.search(r"^[ \t]*#[ \t]*define[ \t]+" + name + r"\b", text, re.M) is not None)
    return hit[1]


def read_plugin_source(path, _seen=None):
    """Plugin device source with ``#include "file"`` of sibling files
    expanded in place (so the code-object cache key covers them).  Kept
    while none of the files it was read from has changed (the launch-shape
    helpers ask for it on every table change of the batched engine)."""
    path = os.path.abspath(path)
    if _seen is None:
        hit = _SOURCES.get(path)
        if hit is not None and all(_mtime(f) == m for f, m in hit[0]):
            return hit[1]
        seen = set()
        text = read_plugin_source(path, seen)
        _SOURCES[path] = (tuple((f, _mtime(f)) for f in seen), text)
        return text
    seen = _seen
    if path in seen:
        return ""
    seen.add(path)
    with open(path) as f:
        text = f.read()
    here = os.path.dirname(path)

    def sub(m):
        inc = os.path.join(here, m.group(1))
        if not os.path.exists(inc):
            return m.group(0)
        return f'#line 1 "{m.group(1)}"\n' + read_plugin_source(inc, seen) + f'\n#line 1 "{os.path.basename(path)}"\n'

    return _INCLUDE_RE.sub(sub, text)


@dataclasses.dataclass(frozen=True)
class ModuleSpec:
    kind: str            # "swept" or "standard"
    user_path: str       # plugin device source
    nv: int
    ops: int
    its: int
    b: int = 0           # swept tile edge
    slots: int = 0       # swept LDS level slots
    bx: int = 0          # standard tile rows
    by: int = 0          # standard tile columns
    threads: int = 256
    real: str = "double"  # state type in HBM and LDS ("double" or "float")
    multi: bool = False   # instance-table kernels with per-instance constants
    mpss: int = 1         # pyramid height (standard modules of a multi class)
    generic: bool = False  # force the per-cell step even if the plugin has a level update
    pad: int = 0          # swept: LDS row padding (SY = b + pad)
    pfd: int = 1          # swept: levels of face prefetch in flight (level update)
    unroll: int = 14      # swept: phases of at most this many levels run their level loop unrolled
    defines: tuple = ()   # macros defined before the plugin source (e.g. EU_EXACT_DIVISION)
    levels: int = 0       # band: most levels per launch

    @property
    def real_bytes(self):
        return 8 if self.real == "double" else 4

    def source(self):
        user = read_plugin_source(self.user_path)
        name = os.path.basename(self.user_path)
        if self.kind == "swept":
            return compiler.swept_source(user, name, self.nv, self.ops, self.its, self.b, self.slots, self.threads,
                                         real=self.real, multi=self.multi, generic=self.generic, pad=self.pad,
                                         pfd=self.pfd, defines=self.defines, unroll=self.unroll)
        if self.kind == "band":
            return compiler.band_source(user, name, self.nv, self.ops, self.its, self.levels, real=self.real,
                                        defines=self.defines, multi=self.multi)
        if self.kind == "standard":
            return compiler.standard_source(user, name, self.nv, self.ops, self.its, self.bx, self.by, self.threads,
                                            mpss=self.mpss, real=self.real, multi=self.multi, generic=self.generic,
                                            defines=self.defines)
        raise ValueError(self.kind)

    def const_layout(self):
        return compiler.constant_layout(read_plugin_source(self.user_path))

    def lds_bytes(self):
        if self.kind == "band":
            return 0
        if self.kind == "swept":
            tile = self.slots * self.nv * self.b * (self.b + self.pad) * self.real_bytes
            return tile + level_scratch_bytes(self.user_path, self.generic, self.b, self.b + self.pad,
                                              swept_faces(self.b, self.ops))
        sx, sy = self.bx + 2 * self.ops, self.by + 2 * self.ops
        tile = standard_slots(self.user_path, self.its, self.generic) * self.nv * sx * sy * self.real_bytes
        return tile + level_scratch_bytes(self.user_path, self.generic, sx, sy,
                                          max((self.bx + 1) * self.by, self.bx * (self.by + 1)), "standard")

    def describe(self):
        g = (f"b={self.b} slots={self.slots}" + (f" pad={self.pad}" if self.pad else "")
             if self.kind == "swept" else f"levels={self.levels}" if self.kind == "band"
             else f"tile={self.bx}x{self.by}")
        return (f"{self.kind}{'-multi' if self.multi else ''}{'-generic' if self.generic else ''} "
                f"{os.path.basename(self.user_path)} {self.real} "
                f"nv={self.nv} k={self.ops} s={self.its} {g} nt={self.threads}"
                + "".join(f" -D{d}" for d in self.defines))


def has_level_update(user_path):
    """True when the plugin provides the cooperative level update."""
    return _defines_macro(user_path, "PS_HAS_LEVEL")


def has_point_update(user_path):
    """True when the plugin provides the point update of the band kernel
    (PS_HAS_POINT: a single-variable radius-1 stencil whose cell update
    takes its neighbours' values)."""
    return _defines_macro(user_path, "PS_HAS_POINT")


def band_levels(its, levels=None, real="double"):
    """Levels per band launch: a multiple of ITS (so that every launch but
    the last starts on a full step), even, at most 16.  Measured on MI355X
    (heat 8192^2, forward Euler, 1000 steps; GCUPS with 64 / 128 / 192 / 256
    output rows per wavefront): FP64 4 levels 1064 / 1000 / 1091 / 914, 8
    levels 1757 / 1820 / 1720 / 1652, 10 levels 1845 / 1924 / 1825 / 1813,
    12 levels 1833 / 1970 / 1868 / 1856, 16 levels 1302 / 1356 / 1427 /
    1465 (the halo columns recomputed grow with the depth, and past 12 the
    register history leaves fewer wavefronts per SIMD); FP32 state, whose
    stochastic rounding costs two integer multiplies per cell, 8 levels
    906 / 935 / - / 766, 12 levels 793 / 851 / - / 712, 16 levels 511 /
    519 / - / 560."""
    if levels is None:
        levels = 12 if real == "double" else 8
    h = int(levels)
    if h < 2 or h > 16 or h % 2 or h % its:
        raise ValueError(f"band levels {h} must be even, a multiple of intermediate_steps={its}, and 2..16")
    return h


def has_band_sweep(user_path):
    """True when the plugin provides a band sweep of its own (PS_HAS_BAND,
    the Euler plugin), compiled when the band kernel's declarations are
    present."""
    return _defines_macro(user_path, "PS_HAS_BAND")


def band_capable(user_path, nv, ops, its):
    """Whether the band kernel can run this plugin's swept solve."""
    if its > 2:
        return False
    if has_point_update(user_path):
        return nv == 1 and ops == 1
    return has_band_sweep(user_path)


def band_multi_capable(user_path, nv, ops, its):
    """Whether problems of this plugin can run on the instance-table band
    kernel (BandMulti): a point update (launches of band_levels levels), or
    a plugin band sweep of one whole step per launch (s = 2)."""
    if not band_capable(user_path, nv, ops, its):
        return False
    return has_point_update(user_path) or its == 2


def band_multi_levels(user_path, its, real="double"):
    """Levels per instance-table band launch: the point update's default
    depth (band_levels), or one step for a plugin band sweep.  A problem on
    the band method keeps that many levels plus one in its ring."""
    return band_levels(its, None, real) if has_point_update(user_path) else its


#: cells from which a plugin's own band sweep (Euler: one column per lane,
#: a 2-column halo per level) is the faster swept kernel: its 64-column
#: bands leave small grids too few wavefronts.  Measured on MI355X (Euler,
#: GCUPS, band / phase kernels b=16): 128^2 0.67 / 0.83, 256^2 2.64 / 3.20,
#: 384^2 5.81 / 4.82, 512^2 7.55 / 7.52, 1024^2 14.41 / 9.73.  Heat's band
#: kernel led at every size measured (64^2 to 8192^2).
BAND_SWEEP_MIN_CELLS = 1 << 17
#: cells from which the Euler band sweep outruns the standard kernel
#: (GCUPS, band / standard: 512^2 7.60 / 10.22, 768^2 11.42 / 12.42,
#: 1024^2 14.57 / 13.77, 1536^2 17.86 / 14.89); heat's band kernel outruns
#: it at every size (64^2 2.98 / 1.09 .. 2048^2 1179 / 286).
BAND_OVER_STANDARD_MIN_CELLS = 1 << 20


def band_default(user_path, nv, ops, its, nx, ny):
    """Whether a swept solve of this plugin and grid runs on the band kernel
    by default (else on the phase kernels)."""
    if not band_capable(user_path, nv, ops, its):
        return False
    return has_point_update(user_path) or nx * ny >= BAND_SWEEP_MIN_CELLS


def band_beats_standard(user_path, nv, ops, its, nx, ny):
    """Whether the band kernel outruns the standard kernel on this grid (the
    engine then runs a problem alone on it)."""
    if not band_capable(user_path, nv, ops, its):
        return False
    return has_point_update(user_path) or nx * ny >= BAND_OVER_STANDARD_MIN_CELLS


def band_out_columns(spec):
    """Output columns per band of a band module: a point-update band is 128
    columns (two per lane) less one halo column per level on each side; a
    plugin sweep is 64 columns (one per lane) less k per level."""
    if has_point_update(spec.user_path):
        return 128 - 2 * spec.levels
    return 64 - 2 * spec.ops * spec.levels


#: wavefront slots the band kernels' launches are balanced over: MI355X has
#: 256 CUs of 4 SIMDs, and a band wavefront keeps a SIMD busy by itself
BAND_SIMDS = 1024


def band_rows(spec, nx, ny, simds=BAND_SIMDS):
    """Output rows per wavefront for a band launch over nx x ny: the row
    segment minimising the busiest SIMD's time (the larger segment on a
    tie).  That SIMD runs m = ceil(wavefronts / SIMDs) wavefronts.  A
    wavefront's steps are its segment plus the x halo of the launch, rounded
    up to the step loop's unrolling (3); level l computes from the step its
    first valid row appears on; a few steps' worth of set-up is added.  Two
    or more wavefronts share their SIMD's issue slots in full; one alone
    stalls on its own dependences and issues about 0.815 of the time.  The
    constants are fitted to these MI355X measurements, whose ranking the
    model reproduces (GCUPS): Euler 1024^2, rows 12 11.2, 16 10.0, 18 8.8,
    20 12.0, 22 12.3, 24 11.1, 28 10.2; Euler 2048^2, rows 16 15.9, 20 15.6,
    24 14.2, 26 16.0, 28 16.3, 32 14.2; heat 8192^2 at 12 levels, rows 64
    1833, 128 1970, 192 1868, 256 1856."""
    wout = band_out_columns(spec)
    bands = -(-ny // wout)
    work_of = band_wave_work(spec)
    best = None
    for rows in range(min(4, nx), min(nx, 256) + 1):
        waves = bands * -(-nx // rows)
        work = work_of(rows)
        m = -(-waves // simds)
        cost = work / 0.815 if m == 1 else m * work
        if best is None or cost <= best[0]:
            best = (cost, rows)
    return best[1]


def band_wave_work(spec):
    """The work of one band wavefront as a function of its output rows, in
    sweep steps (each computing every level of the launch for one row): the
    segment plus the launch's x halo, rounded up to the step loop's
    unrolling (3); level l computes from the step its first valid row appears
    on; a few steps' worth of set-up is added."""
    H, k = spec.levels, spec.ops
    point = has_point_update(spec.user_path)

    def work(rows):
        steps = 3 * -(-(rows + 2 * k * H) // 3)
        if point:
            return sum(steps - 2 * l for l in range(1, H + 1)) + 2 * H
        return sum(steps - (2 * k * l - 1) for l in range(1, H + 1)) + 4
    return work


#: row caps an instance-table band launch chooses from
BAND_MULTI_CAPS = (4, 6, 8, 10, 12, 16, 20, 24, 28, 32, 40, 48, 56, 64, 80, 96, 112, 128, 160, 192, 224, 256)


def band_multi_rows(spec, shapes, simds=BAND_SIMDS):
    """Output rows per wavefront of each problem (nx, ny) of an
    instance-table band launch.  One cap R serves them all: a problem of nx
    rows takes ceil(nx / R) segments of equal length.  R minimises the
    busiest SIMD's time, as band_rows does for one problem: with W
    wavefronts of at most w_max steps and m = ceil(W / SIMDs) per SIMD:
    with m < 4, m * w_max (every SIMD running its share of the longest
    ones, band_rows' measure); from 4 on, where wavefronts of different
    lengths even out, the total work spread over the SIMDs plus one longest
    wavefront (the last ones to start); one wavefront alone on its SIMD
    issues about 0.815 of the time.  Measured on MI355X (the 256 Euler problems of
    the 512-problem mix, engine seconds by cap): 24 0.940, 32 0.916, 48
    0.894, 64 0.866, 96 0.849 (chosen), 128 0.855, 192 0.861, 256 0.908."""
    import numpy
    if not shapes:
        return []
    wout = band_out_columns(spec)
    work_of = band_wave_work(spec)
    nx = numpy.array([s[0] for s in shapes], dtype=numpy.int64)
    bands = -(-numpy.array([s[1] for s in shapes], dtype=numpy.int64) // wout)
    top = BAND_MULTI_CAPS[-1]
    table = numpy.array([0] + [work_of(r) for r in range(1, top + 1)], dtype=numpy.int64)
    best = None
    for cap in BAND_MULTI_CAPS:
        segs = -(-nx // cap)
        rows = -(-nx // segs)
        work = table[rows]
        waves = bands * segs
        W, w_max, total = int(waves.sum()), int(work.max()), float((waves * work).sum())
        m = -(-W // simds)
        if m == 1:
            cost = w_max / 0.815
        elif m < 4:
            cost = m * w_max
        else:
            cost = total / simds + w_max
        if best is None or cost <= best[0]:
            best = (cost, rows)
        if cap >= int(nx.max()):
            break
    return [int(r) for r in best[1]]


def band_spec(user_path, nv, ops, its, real="double", levels=None, defines=(), multi=False):
    """The band module of a plugin; ``multi``: its instance-table form
    (BandMulti, one wavefront per workgroup)."""
    if not band_capable(user_path, nv, ops, its):
        raise ValueError(f"{os.path.basename(user_path)} (nv={nv}, k={ops}, s={its}) has no band form: a point "
                         "update (PS_HAS_POINT, one variable, k=1) or a band sweep (PS_HAS_BAND), s <= 2")
    if levels is None and has_band_sweep(user_path) and not has_point_update(user_path):
        levels = its  # one step per launch: the registers hold a few rows of every level
    if multi and not band_multi_capable(user_path, nv, ops, its):
        raise ValueError(f"{os.path.basename(user_path)} has no instance-table band form: a point update, or a "
                         "band sweep (PS_HAS_BAND) of one whole step per launch (s = 2)")
    if multi and levels is not None and not has_point_update(user_path) and levels != its:
        raise ValueError(f"an instance-table band launch of a band sweep takes one step ({its} levels), not {levels}")
    spec = ModuleSpec("band", os.path.abspath(user_path), nv, ops, its, threads=64 if multi else 256, real=real,
                      multi=bool(multi), levels=band_levels(its, levels, real), defines=tuple(defines))
    if band_out_columns(spec) < 8:
        raise ValueError(f"band of {spec.levels} levels leaves fewer than 8 output columns")
    return spec


def swept_faces(b, ops):
    """Most faces of one direction over the swept sets of a b x b tile."""
    m = b - 2 * ops
    return (m + 1) * m


_SCRATCH_RE = re.compile(r"^[ \t]*#[ \t]*define[ \t]+PS_LEVEL_SCRATCH\((\w+),\s*(\w+)\)(.*)$", re.M)


def level_scratch(user_path):
    """The plugin's PS_LEVEL_SCRATCH(cells, faces) as a Python function
    (doubles of LDS scratch), or None without a level update."""
    if not has_level_update(user_path):
        return None
    m = _SCRATCH_RE.search(read_plugin_source(user_path))
    if m is None:
        return lambda cells, faces: 0
    a, b, expr = m.group(1), m.group(2), m.group(3).strip()
    if not re.fullmatch(r"[\w\s()+*/-]*", expr):
        raise ValueError(f"PS_LEVEL_SCRATCH must be integer arithmetic of {a} and {b}: {expr!r}")
    code = compile(expr.replace("/", "//"), "PS_LEVEL_SCRATCH", "eval")
    return lambda cells, faces: int(eval(code, {"__builtins__": {}}, {a: cells, b: faces}))


def level_update_in_standard(user_path):
    """False when the plugin asks the standard kernel to keep its per-cell
    step (PS_LEVEL_SWEPT_ONLY)."""
    return re.search(r"^[ \t]*#[ \t]*define[ \t]+PS_LEVEL_SWEPT_ONLY\b", read_plugin_source(user_path),
                     re.M) is None


def level_in_place(user_path, generic=False):
    """True when the plugin's level update may write level r over level r-2
    (PS_LEVEL_INPLACE: it reads ``prev`` only at the cell it writes, before
    writing it), so a swept window of ITS levels suffices."""
    if generic or not has_level_update(user_path):
        return False
    return re.search(r"^[ \t]*#[ \t]*define[ \t]+PS_LEVEL_INPLACE\b", read_plugin_source(user_path),
                     re.M) is not None


def min_slots(user_path, its, generic=False):
    """Fewest LDS level slots a swept window may have: ITS+1, or ITS (at
    least 2) when the level update works in place."""
    return max(2, its) if level_in_place(user_path, generic) else its + 1


def standard_slots(user_path, its, generic=False):
    """LDS level slots of the standard kernel: with a plugin level update,
    level A alone (the update reads level A-1 from the ring at each lane's
    own cells and hands level A+1 to the kernel's HBM stores); with the
    per-cell step, the ITS input levels and the output."""
    if not generic and has_level_update(user_path) and level_update_in_standard(user_path):
        return 1
    return its + 1


def level_scratch_bytes(user_path, generic, sx, sy, faces, kind="swept"):
    """LDS bytes of a plugin's level update for an sx x sy tile."""
    f = None if generic or (kind == "standard" and not level_update_in_standard(user_path)) \
        else level_scratch(user_path)
    return 0 if f is None else 8 * f(sx * sy, faces)


def default_slots(nv, its, b, P, real_bytes=8, scratch=0, least=None):
    """LDS window for a swept tile: the largest within ~28 KB including the
    level scratch, at least ``least`` (default ITS+1) levels.  Smaller
    windows slide more often but let more tiles share a CU, which hides the
    HBM latency of each level's face loads.  Measured on MI355X: Euler
    1024^2 b=16 with the level update (12 KB scratch) 3 levels 4.7 GCUPS,
    4 levels 3.9, and in a later build 2 levels in place 6.5 against 3
    levels 6.0; heat 8192^2 b=32 2 levels 175, 3 levels 180, 5 levels 163;
    b=64 2 levels 245."""
    level = nv * b * b * real_bytes
    full = 2 * P + its
    least = its + 1 if least is None else least
    return max(least, min(full, (28 * 1024 - scratch) // level))


def default_pfd(nv, b, threads, real_bytes):
    """Levels of face prefetch in flight for a swept tile (PS_PFD in
    swept.inc).  Measured on MI355X, 300 steps: heat 8192^2 b=64 292 GCUPS
    with 1 level, 321 with 2, 326 with 4, 323 with 6 (the bridges' levels,
    each waiting on its faces, gain most: YBridge -23 %); heat b=32 and the
    Euler vortex at 1024^2 and 2048^2 with b=16 within 2 % for 1..6 levels
    (their faces per level are few, or their levels long enough to hide the
    fetch)."""
    return 4 if nv == 1 and b >= 64 else 1


def default_unroll(nv, b):
    """Most levels of a swept phase whose level loop is unrolled (every
    level's rectangles and index divisors then compile-time constants).
    Measured on MI355X: Euler b=16 (6-level Octahedron) 105 -> 87 us per
    launch and b=32 (14 levels, 70 KB of code) 1053 -> 1022 us; heat 8192^2
    b=32 (15- and 30-level phases, all unrolled) FP32 state 191 -> 245
    GCUPS, FP64 183 either way; heat b=64 with its 31-level bridges unrolled
    FP64 +1 %, FP32 -5 %, and its 62-level Octahedron unrolled (69 KB of
    code) 10 % slower.  The Euler level update is ten times the heat one's
    code, so its phases stop at 14 levels."""
    return 31 if nv == 1 and b <= 32 else 14


def default_pad(nv, b):
    """Doubles of padding per LDS row of a swept tile (SY = b + pad).  Rows of
    16 doubles put every other row on the same LDS banks, and the level
    passes read several rows per wavefront.  Measured on MI355X (Euler
    vortex, 100-200 steps): b=16 FP64 1024^2 9.52 GCUPS with no padding,
    9.65 with 1, 9.75 with 2, 9.68 with 3; 2048^2 9.85 -> 9.99 with 2; FP32
    state b=16 10.68 -> 10.79 with 2; b=32 within 0.3 % for 0..2.  Heat's
    single-variable level update measured no gain (b=64, 324 against 330)."""
    return 2 if nv > 1 and b == 16 else 0


def swept_spec(user_path, nv, ops, its, b, slots=None, threads=None, lds_budget=LDS_BUDGET, real="double",
               multi=False, generic=False, pad=None, pfd=None, defines=(), unroll=None):
    if b % (2 * ops) != 0:
        raise ValueError(f"blocksize {b} must be a multiple of 2*operating_points={2 * ops}")
    P = b // (2 * ops) - 1
    if P < 1:
        raise ValueError(f"blocksize {b} too small for operating_points={ops} (needs b >= {4 * ops})")
    rb = 8 if real == "double" else 4
    if pad is None:
        pad = default_pad(nv, b)
    level = nv * b * (b + pad) * rb
    full = 2 * P + its
    scratch = level_scratch_bytes(user_path, generic, b, b + pad, swept_faces(b, ops))
    least = min_slots(user_path, its, generic)
    if slots is None:
        sc = None if generic else level_scratch(user_path)
        if sc is not None and sc(0, 1) > sc(0, 0) and level_in_place(user_path, generic):
            # a level update that shares faces through LDS is bound by VALU
            # issue, not by the face loads a longer window would hide, and
            # each level it keeps costs a workgroup's LDS: the fewest levels
            # (Euler 2048^2 b=16 with FP32 state: 2 levels 11.9 GCUPS, 4
            # levels 11.0)
            slots = least
        else:
            slots = default_slots(nv, its, b, P, rb, scratch, least)
    slots = min(slots, full)
    if slots < least:
        raise ValueError(f"an LDS window needs at least {least} levels here "
                         "(intermediate_steps+1, or intermediate_steps for an in-place level update)")
    if slots * level + scratch > LDS_MAX:
        raise ValueError(f"swept tile b={b} with {slots} levels needs {slots * level + scratch} B of LDS "
                         f"(> {LDS_MAX})")
    if threads is None:
        threads = max(64, min(256, (b * b + 63) // 64 * 64))
        # a tile whose LDS leaves room for no second one on a CU gets 512
        # lanes (Euler 2048^2 b=32 with FP32 state, 93 KB: 8.94 GCUPS
        # against 7.06 with 256)
        if level > 16 * 1024 or 2 * (slots * level + scratch) > LDS_MAX:
            threads = 512
    if pfd is None:
        pfd = default_pfd(nv, b, threads, rb)
    if unroll is None:
        unroll = default_unroll(nv, b)
    return ModuleSpec("swept", os.path.abspath(user_path), nv, ops, its, b=b, slots=slots, threads=threads,
                      real=real, multi=multi, generic=generic, pad=pad, pfd=int(pfd), defines=tuple(defines),
                      unroll=int(unroll))


def standard_tile(nv, ops, its, nx, ny, lds_budget=LDS_BUDGET, real_bytes=8, scratch=None, threads=256,
                  slots=None):
    """Tile (bx rows, by columns) for the standard kernel: most cells within
    the LDS budget, then rows of at least 32 doubles (coalesced wavefront
    loads), then the least halo re-reading.  ``scratch`` is the plugin's
    level-update scratch function (``level_scratch``) or None.

    When the level update shares faces through LDS (its scratch grows with
    the faces), every pass of the update (the cells' pressures on the
    cross-shaped footprint, the x faces, the y faces) must fit in one sweep
    of the workgroup's lanes: a second, mostly idle sweep costs as much as
    the first, and among tiles of as many cells the one with the least halo
    wins over rows of 32 values.  Measured on MI355X (Euler 1024^2, level A
    alone in LDS): 16x16 on 512 lanes 13.6 GCUPS, 8x32 on 512 lanes 13.4,
    8x16 on 256 lanes 12.8, 16x8 on 256 lanes 12.7 (with levels A-1 and A
    in LDS, 256 lanes: 8x16 7.1, 8x32 6.6, 16x16 6.5).  Rows of 128 values
    are tried with 512 lanes or more: heat 8192^2 32x128 on 512 lanes 318
    GCUPS, 16x256 308, 32x64 on 256 lanes 300 (its one-value y halos cost
    half a line per 64-value row), 32x128 on 256 lanes 225."""
    faces_shared = scratch is not None and scratch(0, 1) > scratch(0, 0)
    slots = its + 1 if slots is None else slots
    best = None
    for by in ((128,) if threads >= 512 else ()) + (64, 32, 16, 8):
        for bx in (32, 16, 8, 4, 2):
            if bx > nx or by > ny:
                continue
            sx, sy = bx + 2 * ops, by + 2 * ops
            lds = slots * nv * sx * sy * real_bytes
            if scratch is not None:
                lds += 8 * scratch(sx * sy, max((bx + 1) * by, bx * (by + 1)))
            if faces_shared:
                foot = sx * sy - 4 * ops * ops
                if max((bx + 1) * by, bx * (by + 1), foot) > threads:
                    continue
            if lds > lds_budget:
                continue
            halo = sx * sy / (bx * by)
            score = (bx * by, -halo, by >= 32) if faces_shared else (bx * by, by >= 32, -halo)
            if best is None or score > best[0]:
                best = (score, bx, by)
    if best is None:
        return max(1, min(nx, 4)), max(1, min(ny, 16))
    return best[1], best[2]


def standard_spec(user_path, nv, ops, its, nx, ny, bx=None, by=None, threads=None, lds_budget=LDS_BUDGET,
                  real="double", multi=False, mpss=1, generic=False, defines=()):
    sc = None if generic or not level_update_in_standard(user_path) else level_scratch(user_path)
    if threads is None:
        # a level update gets 512 lanes: a 16x16 Euler tile's passes fit one
        # sweep, and heat takes 32x128 tiles (measured in standard_tile)
        threads = 512 if sc is not None else 256
    if bx is None or by is None:
        bx, by = standard_tile(nv, ops, its, nx, ny, lds_budget, 8 if real == "double" else 4, sc, threads,
                               standard_slots(user_path, its, generic))
    if bx > nx or by > ny:
        raise ValueError("standard tile larger than the grid")
    return ModuleSpec("standard", os.path.abspath(user_path), nv, ops, its, bx=bx, by=by, threads=threads,
                      real=real, multi=multi, mpss=mpss, generic=generic, defines=tuple(defines))


def build_module_code(spec):
    return compiler.compile_cached(spec.source(), os.path.basename(spec.user_path) + f".{spec.kind}")


_loaded = {}


def _forget_context(handle):
    """Drop (and unload) the modules loaded for a context being closed."""
    for key in [k for k in _loaded if k[1] == handle]:
        _loaded.pop(key).unload()


def load(spec):
    """Load (once per context) the module for ``spec``; returns hip.Module.
    A thread with a ``hip.Context`` current gets that context's own copy, so
    solves on distinct contexts never share a module's ``__constant__``
    globals."""
    from pysweep_amd import hip
    if _forget_context not in hip.CONTEXT_CLOSE_HOOKS:
        hip.CONTEXT_CLOSE_HOOKS.append(_forget_context)
    key = (spec, hip.current_context())
    m = _loaded.get(key)
    if m is None:
        m = hip.Module(build_module_code(spec))
        _loaded[key] = m
    return m


SWEPT_KERNELS = ("UpPyramid", "YBridge", "XBridge", "Octahedron", "DownPyramid")


def shipped_module_specs():
    """Modules used by the test suite, smoke() and bench.py."""
    specs = []
    heat, euler = eq_path("heat.cu"), eq_path("euler.cu")
    ex, ch, hf = eq_path("example.cu"), eq_path("checker.cu"), eq_path("half.cu")
    for b in (8, 16, 32):
        specs.append(swept_spec(heat, 1, 1, 1, b))
        specs.append(swept_spec(heat, 1, 1, 2, b))
    for b in (8, 16, 32):
        specs.append(swept_spec(euler, 4, 2, 2, b))
    specs.append(swept_spec(ex, 1, 1, 1, 8))
    specs.append(swept_spec(ex, 1, 2, 2, 16))
    specs.append(swept_spec(ch, 1, 1, 1, 8))
    specs.append(swept_spec(ch, 1, 2, 2, 16))
    specs.append(swept_spec(hf, 1, 1, 2, 8))
    for (path, nv, k, s) in ((heat, 1, 1, 1), (heat, 1, 1, 2), (euler, 4, 2, 2), (ex, 1, 1, 1), (ex, 1, 2, 2),
                             (ch, 1, 1, 1), (ch, 1, 2, 2), (hf, 1, 1, 2)):
        for n in (64, 384, 1024):
            specs.append(standard_spec(path, nv, k, s, n, n))
    # the Euler modules with IEEE division (a solve repeated after a state
    # left the fast division's range)
    specs.append(swept_spec(euler, 4, 2, 2, 16, defines=("EU_EXACT_DIVISION",)))
    specs.append(standard_spec(euler, 4, 2, 2, 64, 64, defines=("EU_EXACT_DIVISION",)))
    for s_ in (1, 2):
        specs.append(band_spec(heat, 1, 1, s_))
    specs.append(band_spec(eq_path("heat_fp32.cu"), 1, 1, 1, real="float"))
    specs.append(band_spec(euler, 4, 2, 2))
    specs.append(band_spec(euler, 4, 2, 2, defines=("EU_EXACT_DIVISION",)))
    specs.append(band_spec(eq_path("euler_fp32.cu"), 4, 2, 2, real="float"))
    for name, nv, k, s in (("euler_fp32.cu", 4, 2, 2), ("heat_fp32.cu", 1, 1, 1)):
        specs.append(standard_spec(eq_path(name), nv, k, s, 1024, 1024, real="float"))
        for b in ((16, 32) if nv > 1 else (16, 32, 64)):
            specs.append(swept_spec(eq_path(name), nv, k, s, b, real="float"))
    seen, out = set(), []
    for s in specs:
        if s not in seen:
            seen.add(s)
            out.append(s)
    return out
