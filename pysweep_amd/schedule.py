"""Time-step bookkeeping of the swept and standard solvers.

Restates createTimeStepData (core/solver.py:113-128) and the output-level
selection of io.sweptWrite / io.standardWrite (core/io.py:134-144, 186-208)
for a device-resident ring of levels:

* absolute level A holds the state after A-(s-1) stages; levels 0..s-1 are
  the initial condition (core/block.py:25-26); stage g (globalTimeStep g,
  1-based) reads level s-2+g and writes s-1+g;
* full time steps are the levels A = s-1 + m*s, m = 0, 1, ...
"""
import dataclasses


@dataclasses.dataclass
class SweptTimes:
    N: int            # requested steps, int((tf-t0)/dt)
    P: int            # max pyramid swept steps, b/(2k)-1
    G: int            # UpPrism cycles
    stages: int       # swept stages, P*(G+1)
    tail: int         # standard stages run after the swept cycles (exact mode)
    nt: int           # full-step levels recorded
    subtraction: int  # the reference's output-parity flag
    window: int       # ring levels, 2P+s


def requested_steps(t0, tf, dt):
    """N = int((tf - t0)/dt) (core/solver.py:115)."""
    return int((tf - t0) / dt)


def swept_times(N, b, k, s, exact=False):
    """Cycle count and stage totals.

    Reference mode (exact=False): G = int(s*N/P - 1), P*(G+1) stages and
    nt = int(P*(G+1)/s + 1) recorded levels, which quantizes the run to a
    multiple of P stages (core/solver.py:119-121).

    Exact mode: the largest G with P*(G+1) <= s*N, then s*N - P*(G+1) stages
    of the standard kernel on the same ring, so exactly N steps are taken
    (G = -1 means no swept cycle fits and every stage is standard).
    """
    P = b // (2 * k) - 1
    if P < 1:
        raise ValueError(f"blocksize {b} gives no pyramid for operating_points={k}")
    if exact:
        G = (s * N) // P - 1
        swept = P * (G + 1) if G >= 0 else 0
        tail = s * N - swept
        nt = N + 1
    else:
        G = int(s * N / P - 1)
        if G < 0:
            raise ValueError("fewer stages than one pyramid; use exact mode or a smaller blocksize")
        swept = P * (G + 1)
        tail = 0
        nt = int(P * (G + 1) / s + 1)
    sub = 1 if (P % 2 != 0 and s % 2 == 0) else 0
    return SweptTimes(N=N, P=P, G=G, stages=swept, tail=tail, nt=nt, subtraction=sub, window=2 * P + s)


def standard_times(N, b, k, s):
    """Every one of the s*N stages on the standard kernel (no swept cycle)."""
    P = b // (2 * k) - 1
    return SweptTimes(N=N, P=P, G=-1, stages=0, tail=s * N, nt=N + 1, subtraction=0, window=2 * P + s)


def complete_after_cycle(c, G, P, s):
    """Highest absolute level whose every cell is final after swept cycle c
    (0 = FirstPrism, 1..G = UpPrism, G+1 = LastPrism)."""
    if c == 0:
        return s - 1
    return s - 1 + min(c, G + 1) * P


def full_step_levels(s, lo, hi):
    """Full-step absolute levels A in (lo, hi]."""
    first = lo + 1
    r = (first - (s - 1)) % s
    if r:
        first += s - r
    return list(range(first, hi + 1, s))


PHASES = ("UpPyramid", "YBridge", "XBridge", "Octahedron", "DownPyramid")


def cycle_phases(c, G, P, s, b):
    """Phases of swept cycle c (0 FirstPrism, 1..G UpPrism, G+1 LastPrism;
    core/functions.py:12-88) as (kind, A0, gts0, xoff, yoff): tile-local
    level r reads absolute level A0+r-1 and writes A0+r with global stage
    gts0+r-1; tiles start at rows xoff + I*b and columns yoff + J*b."""
    h, base = b // 2, s - 1
    xoff = (c % 2) * h
    if c == 0:
        return [("UpPyramid", base, 1, 0, h), ("YBridge", base, 1, 0, 0)]
    if c <= G:
        A, g = base + (c - 1) * P, (c - 1) * P + 1
        return [("XBridge", A, g, xoff, 0), ("Octahedron", A, g, xoff, h), ("YBridge", A + P, g + P, xoff, 0)]
    A, g = base + G * P, G * P + 1
    return [("XBridge", A, g, xoff, 0), ("DownPyramid", A, g, xoff, h)]


def phase_set(kind, r, b, k, P):
    """Active rectangle (x0, x1, y0, y1) of tile-local level r of a phase
    (core/block.py:123-164; core/pysweep.cu:64-249)."""
    h = b // 2
    if kind == "Octahedron":
        if r <= P:
            return (h - r * k, h + r * k, h - r * k, h + r * k)
        q = r - P
        return (q * k, b - q * k, q * k, b - q * k)
    if kind == "UpPyramid":
        return (r * k, b - r * k, r * k, b - r * k)
    if kind == "DownPyramid":
        return (h - r * k, h + r * k, h - r * k, h + r * k)
    if kind == "YBridge":
        return (r * k, b - r * k, h - r * k, h + r * k)
    if kind == "XBridge":
        return (h - r * k, h + r * k, r * k, b - r * k)
    raise ValueError(kind)


def private_set(kind, r, b, k, s, P):
    """Cells of tile-local level r that no later phase reads from the ring
    (the kernels' ``ps_private``): inside level r+1's set shrunk by k on each
    side, and inside the sets of levels r+2..r+s.  None if empty."""
    L = 2 * P if kind == "Octahedron" else P
    if r + s > L:
        return None
    x0, x1, y0, y1 = phase_set(kind, r, b, k, P)
    a = phase_set(kind, r + 1, b, k, P)
    box = [max(x0, a[0] + k), min(x1, a[1] - k), max(y0, a[2] + k), min(y1, a[3] - k)]
    for j in range(2, s + 1):
        c = phase_set(kind, r + j, b, k, P)
        box = [max(box[0], c[0]), min(box[1], c[1]), max(box[2], c[2]), min(box[3], c[3])]
    if box[0] >= box[1] or box[2] >= box[3]:
        return None
    return tuple(box)


def phase_cell_stages(kind, b, k, P):
    """Cell-stages one tile of a phase computes (sum of its set sizes)."""
    L = 2 * P if kind == "Octahedron" else P
    tot = 0
    for r in range(1, L + 1):
        x0, x1, y0, y1 = phase_set(kind, r, b, k, P)
        tot += (x1 - x0) * (y1 - y0)
    return tot
