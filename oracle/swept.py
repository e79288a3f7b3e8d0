"""Swept space-time schedule and standard stage loop on a periodic grid
(test infrastructure only; see oracle/__init__.py).

The reference keeps a window ``sharedArray[t, v, x, y]`` of 2P+s levels,
rolls it by +-b/2 rows between cycles (functions.py:90-123) and moves it
down by P levels after every UpPrism (io.py:134-144).  Here the same
computation is expressed in the unrolled frame:

* absolute level A holds the state after A-(s-1) stages (levels 0..s-1 hold
  the initial condition, block.py:25-26), stored in ring slot A mod (2P+s);
* a phase is a set of b x b tiles whose origins are (I*b + xoff, J*b + yoff)
  mod (nx, ny); xoff alternates 0, b/2 with the cycle (the roll) and yoff is
  b/2 for UpPyramid / Octahedron / DownPyramid (block.py:106-120,
  pysweep.cu:69,168,223) and 0 for the bridges;
* at tile-local level r the active cells are the reference's sets
  (block.py:123-164, pysweep.cu:71-248).

Every stage is evaluated on the whole grid with the plugin step and only the
active cells are committed, which is exactly the reference's per-cell
arithmetic as long as every value read was written first -- which the
read-after-write checker asserts.
"""
import numpy as np

UP, DOWN, YBRIDGE, XBRIDGE, OCT = "UpPyramid", "DownPyramid", "YBridge", "XBridge", "Octahedron"


def time_step_data(t0, tf, dt, b, k, s, swept=True):
    """createTimeStepData (solver.py:113-128).

    Returns a dict with N (requested steps), P (MPSS), G (UpPrism cycles),
    nt (output levels), subtraction, stages (stages actually run).
    """
    N = int((tf - t0) / dt)
    out = {"N": N}
    if swept:
        P = b // (2 * k) - 1
        G = int(s * N / P - 1)
        nt = int(P * (G + 1) / s + 1)
        out.update(P=P, G=G, nt=nt, O=2 * P,
                   subtraction=1 if (P % 2 != 0 and s % 2 == 0) else 0,
                   window=2 * P + s, stages=P * (G + 1))
    else:
        out.update(nt=N + 1, window=s + 1, stages=s * (N + 1))
    return out


def phase_rect(kind, r, b, k, P):
    """Active set of tile-local level r (1-based) as half-open (x0,x1,y0,y1).

    Up       [rk, b-rk)^2                         block.py:123-133
    Down     [b/2-rk, b/2+rk)^2                   block.py:135-147
    YBridge  x [rk, b-rk), y [b/2-rk, b/2+rk)     block.py:149-164
    XBridge  x [b/2-rk, b/2+rk), y [rk, b-rk)     pysweep.cu:138-161
    Oct      Down for r<=P, then Up(r-P)          block.py:54, pysweep.cu:164-217
    """
    h = b // 2
    if kind == OCT:
        if r <= P:
            return (h - r * k, h + r * k, h - r * k, h + r * k)
        q = r - P
        return (q * k, b - q * k, q * k, b - q * k)
    if kind == UP:
        return (r * k, b - r * k, r * k, b - r * k)
    if kind == DOWN:
        return (h - r * k, h + r * k, h - r * k, h + r * k)
    if kind == YBRIDGE:
        return (r * k, b - r * k, h - r * k, h + r * k)
    if kind == XBRIDGE:
        return (h - r * k, h + r * k, r * k, b - r * k)
    raise ValueError(kind)


def phase_levels(kind, P):
    return 2 * P if kind == OCT else P


def phase_yoff(kind, b):
    return b // 2 if kind in (UP, DOWN, OCT) else 0


def schedule(b, k, s, G):
    """Phase launches of sweptSolve (solver.py:158-176; functions.py:12-88).

    Yields (kind, A0, gts0, xoff): A0 is the absolute level read at r=1,
    gts0 the global stage counter of r=1; level r writes A0+r with
    gts0+r-1.  Cycle c uses xoff=(c%2)*b/2.
    """
    P = b // (2 * k) - 1
    h = b // 2
    base = s - 1
    # FirstPrism: Up then YBridge, stages 1..P (functions.py:12-35)
    yield (UP, base, 1, 0)
    yield (YBRIDGE, base, 1, 0)
    gts = 1
    for c in range(1, G + 1):
        xoff = (c % 2) * h
        A = base + (c - 1) * P
        # UpPrism (functions.py:37-63): XBridge, Octahedron at gts, YBridge at gts+P
        yield (XBRIDGE, A, gts, xoff)
        yield (OCT, A, gts, xoff)
        yield (YBRIDGE, A + P, gts + P, xoff)
        gts += P
    # LastPrism (functions.py:65-88): XBridge + DownPyramid
    c = G + 1
    xoff = (c % 2) * h
    A = base + G * P
    yield (XBRIDGE, A, gts, xoff)
    yield (DOWN, A, gts, xoff)


def tile_mask(rect, b, nx, ny, xoff, yoff):
    """Global boolean mask of a tile-local rectangle repeated over all tiles."""
    x0, x1, y0, y1 = rect
    m = np.zeros((b, b), dtype=bool)
    m[x0:x1, y0:y1] = True
    big = np.tile(m, (nx // b, ny // b))
    return np.roll(big, (xoff, yoff), axis=(0, 1))


class Ring:
    """Ring of absolute levels with a read-after-write checker."""

    def __init__(self, ic, T, s, check=True):
        nv, nx, ny = ic.shape
        self.T, self.s = T, s
        self.data = np.zeros((T, nv, nx, ny))
        self.level = np.full(T, -10**9)
        self.valid = np.zeros((T, nx, ny), dtype=bool)
        self.check = check
        for A in range(s):
            self.data[A % T] = ic
            self.level[A % T] = A
            self.valid[A % T] = True

    def get(self, A):
        sl = A % self.T
        assert self.level[sl] == A, f"level {A} evicted (slot holds {self.level[sl]})"
        return self.data[sl]

    def require(self, A, mask):
        if not self.check:
            return
        sl = A % self.T
        assert self.level[sl] == A, f"read of level {A} after eviction"
        bad = mask & ~self.valid[sl]
        assert not bad.any(), f"read of unwritten cells at level {A}: {np.argwhere(bad)[:4].tolist()}"

    def commit(self, A, values, mask):
        sl = A % self.T
        if self.level[sl] != A:
            self.level[sl] = A
            self.valid[sl] = False
        if self.check:
            dup = mask & self.valid[sl]
            assert not dup.any(), f"cell written twice at level {A}"
        self.data[sl][:, mask] = values[:, mask]
        self.valid[sl] |= mask


def _dilate(mask, k):
    """Cells read by a radius-k cross stencil centred on ``mask``."""
    out = mask.copy()
    for d in range(1, k + 1):
        for ax in (0, 1):
            out |= np.roll(mask, d, axis=ax) | np.roll(mask, -d, axis=ax)
    return out


def swept_solve(step, ic, b, k, s, G, check=True, on_level=None):
    """Run FirstPrism, G UpPrisms and LastPrism.

    ``step(q, qb, gts)`` evaluates one stage on the whole grid.  Returns the
    ring and the final absolute level.  ``on_level(A, q)`` is called for
    every absolute level once all of its cells are final.
    """
    nv, nx, ny = ic.shape
    assert b % (2 * k) == 0 and nx % b == 0 and ny % b == 0
    P = b // (2 * k) - 1
    assert P >= 1
    T = 2 * P + s
    ring = Ring(ic, T, s, check)
    full = np.ones((nx, ny), dtype=bool)
    for kind, A0, gts0, xoff in schedule(b, k, s, G):
        yoff = phase_yoff(kind, b)
        for r in range(1, phase_levels(kind, P) + 1):
            A = A0 + r - 1
            mask = tile_mask(phase_rect(kind, r, b, k, P), b, nx, ny, xoff, yoff)
            ring.require(A, _dilate(mask, k))
            for j in range(1, s):
                ring.require(A - j, mask)
            qb = ring.get(A - 1) if s > 1 else None
            new = step(ring.get(A), qb, gts0 + r - 1)
            ring.commit(A + 1, new, mask)
            if on_level is not None and ring.valid[(A + 1) % T].all():
                on_level(A + 1, ring.get(A + 1))
    final = s - 1 + P * (G + 1)
    if check:
        assert ring.valid[final % T].all()
        ring.require(final, full)
    return ring, final


def standard_solve(step, ic, s, stages, on_level=None):
    """Standard decomposition (solver.py:178-189): one stage per pass over the
    whole periodic grid; stage j (gts=j) reads level s-2+j and writes s-1+j."""
    T = s + 1
    ring = Ring(ic, T, s, check=False)
    full = np.ones(ic.shape[1:], dtype=bool)
    for j in range(1, stages + 1):
        A = s - 2 + j
        qb = ring.get(A - 1) if s > 1 else None
        new = step(ring.get(A), qb, j)
        ring.commit(A + 1, new, full)
        if on_level is not None:
            on_level(A + 1, ring.get(A + 1))
    return ring, s - 1 + stages


def full_step_levels(s, last):
    """Absolute levels that hold full steps: A = s-1 + m*s (m = 0, 1, ...)."""
    return list(range(s - 1, last + 1, s))
