"""Pins the swept schedule of the oracle (oracle/swept.py): step-count
arithmetic, swept == standard bitwise for every plugin, the reference's
known-answer tests, and the read-after-write checker."""
import numpy as np
import pytest

from oracle import euler, heat, plugins
from oracle import swept as sw


@pytest.mark.parametrize("args,P,G,stages,nt", [
    # heat 384^2, b=8, k=1, s=1, 100 steps (dt=0.1*dx^2)
    ((0.0, 100 * 0.1 / 384 ** 2, 0.1 / 384 ** 2, 8, 1, 1), 3, 32, 99, 100),
    # euler 1024^2, b=16, k=2, s=2, 500 steps
    ((0.0, 500 * 0.1 * 10 / 1024, 0.1 * 10 / 1024, 16, 2, 2), 3, 332, 999, 500),
    # heat b=32, 1000 steps
    ((0.0, 1000 * 1e-3, 1e-3, 32, 1, 1), 15, 65, 990, 991),
    # euler b=32, 1000 steps
    ((0.0, 1000 * 1e-3, 1e-3, 32, 2, 2), 7, 284, 1995, 998),
])
def test_time_step_data(args, P, G, stages, nt):
    """createTimeStepData (solver.py:113-128) quantization."""
    d = sw.time_step_data(*args)
    assert (d["P"], d["G"], d["stages"], d["nt"]) == (P, G, stages, nt)


def _ic_heat(n):
    return heat.initial_condition(n, seed=11)


def _plugin_cases():
    n = 48
    dxh = 1.0 / n
    dth = 0.1 * dxh * dxh
    c = heat.courant_py(1.0, dth, dxh)
    dxe = 10.0 / n
    yield ("heat_fe", 1, 1, _ic_heat(n), lambda q, qb, g: heat.step_py(q[0], None, g, c, True)[None])
    yield ("heat_rk2", 1, 2, _ic_heat(n), lambda q, qb, g: heat.step_py(q[0], qb[0], g, c, False)[None])
    yield ("euler", 2, 2, euler.initial_condition(n, eps=3.0),
           lambda q, qb, g: euler.step_py(q, qb, g, 1.4, 0.1 * dxe, dxe, dxe))
    yield ("example_fe", 1, 1, np.zeros((2, n, n)), lambda q, qb, g: plugins.example_step(q, qb, g, True))
    yield ("example_rk2", 2, 2, np.zeros((1, n, n)), lambda q, qb, g: plugins.example_step(q, qb, g, False))
    yield ("checker", 1, 1, plugins.checker_ic(1, n, n), plugins.checker_step)
    yield ("half", 1, 2, plugins.half_ic(n, n), plugins.half_step)


@pytest.mark.parametrize("case", list(_plugin_cases()), ids=lambda c: c[0])
@pytest.mark.parametrize("b", [8, 16, 24])
def test_swept_equals_standard_bitwise(case, b):
    name, k, s, ic, step = case
    if b % (2 * k) or (b // (2 * k) - 1) < 1 or ic.shape[1] % b:
        pytest.skip("invalid block size for this stencil")
    P = b // (2 * k) - 1
    G = 3
    ring, fin = sw.swept_solve(step, ic, b, k, s, G, check=True)
    ring2, fin2 = sw.standard_solve(step, ic, s, P * (G + 1))
    assert fin == fin2
    assert np.array_equal(ring.get(fin), ring2.get(fin2))


def _full_step_history(step, ic, b, k, s, G):
    hist = {}
    sw.swept_solve(step, ic, b, k, s, G, on_level=lambda A, q: hist.__setitem__(A, q.copy()))
    return [hist[A] for A in sorted(hist) if (A - (s - 1)) % s == 0] if s > 1 else [ic] + [hist[A] for A in sorted(hist)]


def test_known_answer_simple_one():
    """testSimpleOne (tests.py:255-286): every cell == i at step i."""
    n, b = 32, 8
    levels = _full_step_history(lambda q, qb, g: plugins.example_step(q, qb, g, True),
                                np.zeros((1, n, n)), b, 1, 1, 5)
    for i, lv in enumerate(levels):
        assert np.all(lv == i)


def test_known_answer_simple_two():
    """testSimpleTwo (tests.py:288-320): k=2, s=2, full step i == 2i."""
    n, b = 32, 16
    ic = np.zeros((1, n, n))
    hist = {}
    sw.swept_solve(lambda q, qb, g: plugins.example_step(q, qb, g, False), ic, b, 2, 2, 5,
                   on_level=lambda A, q: hist.__setitem__(A, q.copy()))
    for A, q in hist.items():
        if (A - 1) % 2 == 0:
            assert np.all(q == 2 * ((A - 1) // 2))


def test_known_answer_checker_one():
    """testCheckerOne (tests.py:322-366): the checkerboard shifts by one cell
    in x every step."""
    n, b = 32, 8
    ic = plugins.checker_ic(1, n, n)
    levels = _full_step_history(plugins.checker_step, ic, b, 1, 1, 4)
    p1 = ic[0]
    p2 = np.roll(p1, 1, axis=0)
    for i, lv in enumerate(levels):
        assert np.array_equal(lv[0], p1 if i % 2 == 0 else p2)


def test_known_answer_checker_two():
    """testCheckerTwo (tests.py:369-408): two averaging stages per step leave
    the pattern unchanged at every full step."""
    n, b = 32, 16
    ic = plugins.checker_ic(1, n, n)
    hist = {}
    sw.swept_solve(plugins.checker_step, ic, b, 2, 2, 4, on_level=lambda A, q: hist.__setitem__(A, q.copy()))
    for A, q in hist.items():
        if (A - 1) % 2 == 0:
            assert np.array_equal(q[0], ic[0])


def test_half_against_reference_brute_force():
    """testHalf (tests.py:222-253): rtol=atol=9e-15 vs the in-test formula."""
    n, b = 24, 8
    ic = plugins.half_ic(n, n)
    G = 5
    ring, fin = sw.swept_solve(plugins.half_step, ic, b, 1, 2, G)
    steps = (fin - 1) // 2
    want = plugins.half_brute_force(ic[0], steps)
    assert (fin - 1) % 2 == 0
    assert np.allclose(ring.get(fin)[0], want, rtol=9e-15, atol=9e-15)


def test_raw_checker_catches_wrong_lattice_offset():
    """A schedule whose UpPyramid is not shifted by b/2 in y (planted in a
    copy of the oracle on the CPU) reads cells before they are written."""
    saved = sw.phase_yoff
    try:
        sw.phase_yoff = lambda kind, b: 0
        with pytest.raises(AssertionError):
            sw.swept_solve(lambda q, qb, g: plugins.example_step(q, qb, g, True),
                           np.zeros((1, 32, 32)), 8, 1, 1, 3)
    finally:
        sw.phase_yoff = saved


def test_every_cell_level_written_once():
    """The commit checker asserts single writes; this also checks the final
    level is complete and equals the number of stages."""
    n, b = 32, 8
    ring, fin = sw.swept_solve(lambda q, qb, g: plugins.example_step(q, qb, g, True),
                               np.zeros((1, n, n)), b, 1, 1, 4)
    assert np.all(ring.get(fin) == 3 * 5)


def _hull(a, c):
    if a is None:
        return c
    if c is None:
        return a
    return (min(a[0], c[0]), max(a[1], c[1]), min(a[2], c[2]), max(a[3], c[3]))


@pytest.mark.parametrize("b,k,s", [(8, 1, 1), (8, 1, 2), (16, 2, 2), (16, 1, 2), (12, 1, 3), (24, 2, 3),
                                   (32, 2, 2), (32, 1, 1), (24, 3, 2), (16, 2, 1), (32, 4, 2)])
@pytest.mark.parametrize("G", [0, 1, 3])
def test_private_cells_are_never_loaded(b, k, s, G):
    """The swept kernels do not write the cells of a level that only the same
    tile's next levels read (pysweep_amd.schedule.private_set, the kernels'
    ps_private).  Replay every phase's HBM loads -- for each tile-local level
    l the hull of the radius-k box around level l+1's set and of levels
    l+2..l+s, minus level l's own set -- and check that no private cell below
    the final s levels is ever loaded."""
    from pysweep_amd import schedule as ps
    P = b // (2 * k) - 1
    n = 4 * b
    names = {sw.UP: "UpPyramid", sw.YBRIDGE: "YBridge", sw.XBRIDGE: "XBridge", sw.OCT: "Octahedron",
             sw.DOWN: "DownPyramid"}

    def S(kind, r):
        L = sw.phase_levels(kind, P)
        return sw.phase_rect(kind, r, b, k, P) if 1 <= r <= L else None

    def box(rect):
        m = np.zeros((b, b), dtype=bool)
        if rect is not None:
            m[rect[0]:rect[1], rect[2]:rect[3]] = True
        return m

    last = s - 1 + (G + 1) * P
    loaded = np.zeros((last + 1, n, n), dtype=bool)
    phases = list(sw.schedule(b, k, s, G))
    for kind, A0, _, xoff in phases:
        yoff = sw.phase_yoff(kind, b)
        for l in range(1 - s, sw.phase_levels(kind, P)):
            u = S(kind, l + 1)
            if u is not None:
                u = (max(u[0] - k, 0), min(u[1] + k, b), max(u[2] - k, 0), min(u[3] + k, b))
            for r in range(l + 2, l + s + 1):
                u = _hull(u, S(kind, r))
            m = box(u) & ~box(S(kind, l))
            if A0 + l >= 0:
                loaded[A0 + l] |= np.roll(np.tile(m, (n // b, n // b)), (xoff, yoff), axis=(0, 1))
    skipped = 0
    for kind, A0, _, xoff in phases:
        yoff = sw.phase_yoff(kind, b)
        for r in range(1, sw.phase_levels(kind, P) + 1):
            q = ps.private_set(names[kind], r, b, k, s, P)
            if q is None or A0 + r >= (G + 1) * P:
                continue
            m = sw.tile_mask(q, b, n, n, xoff, yoff)
            assert not (m & loaded[A0 + r]).any(), (names[kind], r)
            skipped += int(m.sum())
    if G >= 1 and P >= 2:
        assert skipped > 0
