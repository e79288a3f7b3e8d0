"""GPU parity: the HIP swept and standard kernels (through the C ABI and the
product Solver) against the CPU oracle on the same seeded inputs.

Tolerance: max|gpu - oracle| / max|oracle| <= 1e-12 per variable (FP64).
The device plugins evaluate the numpy plugin's operation order with no FMA
contraction, so results are also expected to be bitwise equal; for Euler
that is what makes multi-stage agreement possible at all (the scheme turns
a 1-ulp difference into >1e-10 within a few stages, see
tests/test_oracle_euler.py::test_round_off_sensitivity_is_a_property_of_the_scheme).
"""
import numpy as np
import pytest

from gpu_helpers import (euler_case, gpu_solve, heat_case, normwise, o_euler, o_heat, o_plugins, o_swept,
                         oracle_stages, swept_kernel)

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.mark.parametrize("fe", [True, False], ids=["fe", "rk2"])
@pytest.mark.parametrize("swept,b,phases", [(False, 8, False), (True, 8, False), (True, 32, False), (True, 8, True),
                                            (True, 16, True), (True, 32, True)])
def test_heat_matches_oracle(device, fe, swept, b, phases):
    """Heat swept solves on the band kernel (their default) and on the five
    phase kernels; both, and the standard kernel, give the oracle's bits."""
    n, steps = 64, 40
    s = 1 if fe else 2
    ic, g, step = heat_case(n, fe, steps, seed=1)
    sol = gpu_solve("heat", ic, g, 1, s, b, swept=swept, exact=True, tuning=swept_kernel(phases))
    from pysweep_amd import device as dev
    assert isinstance(sol._run, dev.BandRun) == (swept and not phases)
    want = oracle_stages(step, ic, s, s * steps)
    assert normwise(sol.final_state, want) <= TOL
    assert np.array_equal(sol.final_state, want)


@pytest.mark.parametrize("swept,b,phases", [(False, 8, False), (True, 16, False), (True, 8, True), (True, 16, True),
                                            (True, 32, True)])
def test_euler_matches_oracle(device, swept, b, phases):
    """Euler swept solves on the band kernel (the plugin's band sweep; the
    default from 2^17 cells, asked for here) and on the five phase kernels;
    both, and the standard kernel, give the oracle's bits."""
    n, steps = 64, 12
    ic, g, step = euler_case(n, steps, eps=5.0)
    sol = gpu_solve("euler", ic, g, 2, 2, b, swept=swept, exact=True, tuning=swept_kernel(phases))
    from pysweep_amd import device as dev
    assert isinstance(sol._run, dev.BandRun) == (swept and not phases)
    want = oracle_stages(step, ic, 2, 2 * steps)
    assert normwise(sol.final_state, want) <= TOL
    assert np.array_equal(sol.final_state, want)


@pytest.mark.parametrize("swept,b", [(False, 8), (True, 16), (True, 32)])
def test_euler_per_cell_step_contract_matches_oracle(device, swept, b):
    """The per-cell step(double*, int, int) path (what any plugin uses) gives
    the same bits as the plugin's cooperative level update and the oracle."""
    n, steps = 64, 12
    ic, g, step = euler_case(n, steps, eps=5.0)
    sol = gpu_solve("euler", ic, g, 2, 2, b, swept=swept, exact=True, tuning={"generic": True})
    want = oracle_stages(step, ic, 2, 2 * steps)
    assert np.array_equal(sol.final_state, want)


def test_euler_one_stage_parity_with_reference_pow_along_trajectory(device):
    """Each GPU stage against one stage of the faithful numpy path (the C
    library's pow for squares, euler.py:194, 239) applied to the GPU's own
    previous levels."""
    from pysweep_amd import device as dev
    from pysweep_amd.equations import path as eq_path
    from pysweep_amd.plugins import load_cpu_module
    n, stages = 48, 12
    ic, g, _ = euler_case(n, stages // 2, eps=4.0)
    run = dev.StandardRun(eq_path("euler.cu"), 4, 2, 2, n, n, T=stages + 2)
    cpu = load_cpu_module(eq_path("euler.py"))
    cpu.set_globals(*g)
    run.set_globals(cpu, g)
    run.upload_ic(ic)
    levels = [run.level(0), run.level(1)]
    for gts in range(1, stages + 1):
        run.run(gts, gts + 1)
        levels.append(run.level(gts + 1))
    dx, dt = g[3], g[2]
    for gts in range(1, stages + 1):
        A = gts  # level read
        want = o_euler.step_py(levels[A], levels[A - 1], gts, 1.4, dt, dx, dx, sq=o_euler.sq_pow)
        assert normwise(levels[A + 1], want) <= TOL, f"stage {gts}"
    run.close()


def _swept_vs_standard(name, ic, g, k, s, b, steps, tuning=None):
    a = gpu_solve(name, ic, g, k, s, b, swept=True, exact=True, tuning=tuning)
    c = gpu_solve(name, ic, g, k, s, b, swept=False, exact=True)
    return a.final_state, c.final_state


@pytest.mark.parametrize("phases", [True, False], ids=["phases", "band"])
@pytest.mark.parametrize("b", [8, 16, 32])
def test_heat_swept_equals_standard_bitwise(device, b, phases):
    ic, g, _ = heat_case(128, True, 100, seed=2)
    a, c = _swept_vs_standard("heat", ic, g, 1, 1, b, 100, tuning=swept_kernel(phases))
    assert np.array_equal(a, c)


@pytest.mark.parametrize("phases", [True, False], ids=["phases", "band"])
@pytest.mark.parametrize("b", [8, 16, 32])
def test_euler_swept_equals_standard_bitwise(device, b, phases):
    ic, g, _ = euler_case(128, 30, eps=3.0, xc=1.0, yc=-2.0)
    a, c = _swept_vs_standard("euler", ic, g, 2, 2, b, 30, tuning=swept_kernel(phases))
    assert np.array_equal(a, c)


def test_non_square_grid(device):
    """nx != ny (the batched workloads use non-square grids)."""
    nx, ny, steps = 64, 96, 6
    dx = 10.0 / nx
    dt = 0.1 * dx
    X, Y = np.meshgrid(o_euler.grid(nx), np.linspace(-5, 5, ny, endpoint=False), indexing="ij")
    ic = o_euler.vortex(X, Y, 0.0, 1.4, eps=3.0)
    g = [0.0, steps * dt, dt, dx, dx, 1.4]
    sol = gpu_solve("euler", ic, g, 2, 2, 16, swept=True, exact=True, tuning={"band": True})
    ph = gpu_solve("euler", ic, g, 2, 2, 16, swept=True, exact=True, tuning={"phases": True})
    std = gpu_solve("euler", ic, g, 2, 2, 16, swept=False, exact=True)
    want = oracle_stages(lambda q, qb, gg: o_euler.step_py(q, qb, gg, 1.4, dt, dx, dx), ic, 2, 2 * steps)
    assert np.array_equal(sol.final_state, want)
    assert np.array_equal(ph.final_state, want)
    assert np.array_equal(std.final_state, want)


@pytest.mark.parametrize("nx,ny", [(48, 80), (80, 48), (16, 16)])
def test_tile_counts_not_multiple_of_xcds(device, nx, ny):
    """Tile counts that are not multiples of the 8 XCDs (15 and 1 swept b=16
    tiles, 30 and 2 standard 8x16 tiles): the XCD-aware tile order still
    visits every tile once."""
    steps = 5
    dx = 10.0 / nx
    dt = 0.1 * dx
    X, Y = np.meshgrid(o_euler.grid(nx), np.linspace(-5, 5, ny, endpoint=False), indexing="ij")
    ic = o_euler.vortex(X, Y, 0.0, 1.4, eps=3.0)
    g = [0.0, steps * dt, dt, dx, dx, 1.4]
    want = oracle_stages(lambda q, qb, gg: o_euler.step_py(q, qb, gg, 1.4, dt, dx, dx), ic, 2, 2 * steps)
    assert np.array_equal(gpu_solve("euler", ic, g, 2, 2, 16, swept=True, exact=True,
                                    tuning={"phases": True}).final_state, want)
    assert np.array_equal(gpu_solve("euler", ic, g, 2, 2, 16, swept=False, exact=True).final_state, want)


# ---------------------------------------------------------------- known answers

def _kat_globals(steps, scheme=None):
    g = [0.0, steps * 0.1, 0.1, 0.1, 0.1]
    return g + ([scheme] if scheme is not None else [])


def test_known_answer_simple_one(device):
    """testSimpleOne (tests.py:255-286): cell == i at step i, every step."""
    n, steps = 64, 60
    sol = gpu_solve("example", np.zeros((1, n, n)), _kat_globals(steps, True), 1, 1, 8, swept=True,
                    exact=False, history=True)
    for i in range(sol.arrayShape[0]):
        assert np.all(sol.data[i] == i), i


def test_known_answer_simple_two(device):
    """testSimpleTwo (tests.py:288-320): k=2, s=2: cell == 2i at step i."""
    n, steps = 64, 60
    sol = gpu_solve("example", np.zeros((1, n, n)), _kat_globals(steps, False), 2, 2, 16, swept=True,
                    exact=False, history=True)
    for i in range(sol.arrayShape[0]):
        assert np.all(sol.data[i] == 2 * i), i


def test_known_answer_checker_one(device):
    """testCheckerOne (tests.py:322-366)."""
    n, steps = 64, 40
    ic = o_plugins.checker_ic(1, n, n)
    sol = gpu_solve("checker", ic, _kat_globals(steps, True), 1, 1, 8, swept=True, exact=False, history=True)
    p1, p2 = ic[0], np.roll(ic[0], 1, axis=0)
    for i in range(sol.arrayShape[0]):
        assert np.array_equal(sol.data[i, 0], p1 if i % 2 == 0 else p2), i


def test_known_answer_checker_two(device):
    """testCheckerTwo (tests.py:369-408)."""
    n, steps = 64, 40
    ic = o_plugins.checker_ic(1, n, n)
    sol = gpu_solve("checker", ic, _kat_globals(steps, False), 2, 2, 16, swept=True, exact=False, history=True)
    for i in range(sol.arrayShape[0]):
        assert np.array_equal(sol.data[i, 0], ic[0]), i


@pytest.mark.parametrize("swept", [True, False])
def test_known_answer_half(device, swept):
    """testHalf (tests.py:222-253): rtol=atol=9e-15 against the in-test
    restatement, at every recorded step."""
    n, steps = 48, 30
    ic = o_plugins.half_ic(n, n)
    sol = gpu_solve("half", ic, _kat_globals(steps), 1, 2, 8, swept=swept, exact=False, history=True)
    a = ic[0].copy()
    for i in range(1, sol.arrayShape[0]):
        a = o_plugins.half_brute_force(a, 1)
        assert np.allclose(sol.data[i, 0], a, rtol=9e-15, atol=9e-15), i


# ---------------------------------------------------------------- configurations

@pytest.mark.parametrize("phases", [True, False], ids=["phases", "band"])
def test_config_heat_384_swept_cpu_case(device, phases):
    """Heat 384^2, b=8, forward Euler, 100 steps with the reference's
    quantization (99 stages): oracle bitwise, exact discrete Fourier solution
    and the analytic decay of the sin(2 pi x) sin(2 pi y) mode."""
    n, steps = 384, 100
    ic, g, step = heat_case(n, True, steps, seed=0)
    sol = gpu_solve("heat", ic, g, 1, 1, 8, swept=True, exact=False, tuning=swept_kernel(phases))
    assert sol.times.stages == 99 and sol.times.G == 32
    want = oracle_stages(step, ic, 1, 99)
    assert np.array_equal(sol.final_state, want)
    c = o_heat.courant_py(1.0, g[2], g[3])
    four = o_heat.fourier_solution(ic[0], 99, c, True)
    assert np.abs(sol.final_state[0] - four).max() <= 1e-12 * np.abs(four).max()
    x = o_heat.grid(n)
    mode = np.outer(np.sin(2 * np.pi * x), np.sin(2 * np.pi * x))
    amp = (sol.final_state[0] * mode).sum() / (mode * mode).sum()
    amp0 = (ic[0] * mode).sum() / (mode * mode).sum()
    t = 99 * g[2]
    assert abs(amp / amp0 - np.exp(-8 * np.pi ** 2 * t)) < 1e-4


def test_config_euler_1024_full_size(device):
    """Euler 1024^2, b=16, 500 steps (exact): swept == standard bitwise,
    conservation of the four integrals, and the vortex error of the same
    size as the oracle's on a coarser grid."""
    n, steps = 1024, 500
    ic, g, _ = euler_case(n, steps)
    sw = gpu_solve("euler", ic, g, 2, 2, 16, swept=True, exact=True)
    ph = gpu_solve("euler", ic, g, 2, 2, 16, swept=True, exact=True, tuning={"phases": True})
    st = gpu_solve("euler", ic, g, 2, 2, 16, swept=False, exact=True)
    assert np.array_equal(sw.final_state, st.final_state)
    assert np.array_equal(ph.final_state, st.final_state)
    for v in range(4):
        assert abs(sw.final_state[v].sum() - ic[v].sum()) <= 1e-11 * np.abs(ic[v]).sum()
    X, Y = np.meshgrid(o_euler.grid(n), o_euler.grid(n), indexing="ij")
    ex = o_euler.vortex(X, Y, steps * g[2], 1.4)
    err = np.abs(sw.final_state[0] - ex[0]).max()
    assert err < 0.05


def test_standard_ragged_tiles(device):
    """Grid sides that are not multiples of the standard tile (partial tiles
    at the high edges)."""
    nx, ny, steps = 40, 24, 5
    dx = 10.0 / nx
    dt = 0.1 * dx
    X, Y = np.meshgrid(o_euler.grid(nx), np.linspace(-5, 5, ny, endpoint=False), indexing="ij")
    ic = o_euler.vortex(X, Y, 0.0, 1.4, eps=2.5)
    g = [0.0, steps * dt, dt, dx, dx, 1.4]
    sol = gpu_solve("euler", ic, g, 2, 2, 8, swept=False, exact=True)
    want = oracle_stages(lambda q, qb, gg: o_euler.step_py(q, qb, gg, 1.4, dt, dx, dx), ic, 2, 2 * steps)
    assert np.array_equal(sol.final_state, want)


@pytest.mark.parametrize("fe", [True, False])
def test_standard_ragged_tiles_heat(device, fe):
    """Heat on the standard kernel's 32x64 tiles with both sides ragged
    (40 x 72): the level update reads level A-1 from the ring, and the
    cells of the partial tiles beyond the grid must neither be read nor
    stored; forward Euler and RK2 (which reads level A-1)."""
    nx, ny, steps = 40, 72, 7
    dx = 1.0 / nx
    dt = 0.1 * dx * dx
    x = np.linspace(0, 1, nx, endpoint=False)
    y = np.linspace(0, 1, ny, endpoint=False)
    ic = (1.3 * np.outer(np.sin(2 * np.pi * x), np.sin(4 * np.pi * y)))[None]
    its = 1 if fe else 2
    g = [0.0, steps * dt, dt, dx, dx, 1.0, fe]
    c = o_heat.courant_py(1.0, dt, dx)
    sol = gpu_solve("heat", ic, g, 1, its, 8, swept=False, exact=True)
    want = oracle_stages(lambda q, qb, gg: o_heat.step_py(q[0], None if qb is None else qb[0], gg, c, fe)[None],
                         ic, its, its * steps)
    assert np.array_equal(sol.final_state, want)


@pytest.mark.parametrize("b,phases", [(32, True), (64, True), (64, False)])
def test_large_heat_swept_equals_standard_and_decays(device, b, phases):
    """Heat 8192^2 (the large-instance size), 200 forward-Euler steps:
    swept == standard bitwise; the sin(2 pi x) sin(2 pi y) mode decays as
    exp(-8 pi^2 t) to O(dx^2)."""
    from pysweep_amd import workloads
    n, steps = 8192, 200
    dx = 1.0 / n
    dt = 0.1 * dx * dx
    ic = workloads.heat_mode(n, n)
    g = [0.0, steps * dt, dt, dx, dx, 1.0, True]
    sw = gpu_solve("heat", ic, g, 1, 1, b, swept=True, exact=True, tuning=swept_kernel(phases))
    st = gpu_solve("heat", ic, g, 1, 1, b, swept=False, exact=True)
    assert np.array_equal(sw.final_state, st.final_state)
    ratio = sw.final_state[0, n // 4, n // 4] / ic[0, n // 4, n // 4]
    assert abs(ratio - np.exp(-8 * np.pi ** 2 * steps * dt)) < 1e-9


@pytest.mark.parametrize("phases", [True, False], ids=["phases", "band"])
def test_large_euler_swept_equals_standard_conserves(device, phases):
    """Euler 2048^2, b=32, 100 steps: swept == standard bitwise and the four
    integrals conserved to round-off."""
    from pysweep_amd import workloads
    n, steps = 2048, 100
    dx = 10.0 / n
    dt = 0.1 * dx
    ic = workloads.vortex_state(n, n)
    g = [0.0, steps * dt, dt, dx, dx, 1.4]
    sw = gpu_solve("euler", ic, g, 2, 2, 32, swept=True, exact=True, tuning=swept_kernel(phases))
    st = gpu_solve("euler", ic, g, 2, 2, 32, swept=False, exact=True)
    assert np.array_equal(sw.final_state, st.final_state)
    for v in range(4):
        assert abs(sw.final_state[v].sum() - ic[v].sum()) <= 1e-11 * np.abs(ic[v]).sum()


# ------------------------------------------- the reference's own test sweeps
# tests.py `testing` decorator (tests.py:54-77): block sizes 8, 12, 16, 24,
# swept and standard; the heat and vortex tests compare with the analytic
# solutions through numpy.allclose (tests.py:410-519, 106-163).

@pytest.mark.parametrize("swept", [True, False])
@pytest.mark.parametrize("b", [8, 12, 16, 24])
def test_reference_sweep_heat_rk2_384(device, b, swept):
    """testHeatRungeKuttaTwo (tests.py:466-519): heat RK2 on 384^2 over the
    reference's block sizes, from the analytic initial condition: the
    oracle's bits, and numpy.allclose against the analytic decay at its
    default tolerances, as the reference asserts."""
    n, steps = 384, 60
    dx = 1.0 / n
    dt = 0.1 * dx * dx
    x = o_heat.grid(n)
    ic = np.outer(np.sin(2 * np.pi * x), np.sin(2 * np.pi * x))[None]
    g = [0.0, steps * dt, dt, dx, dx, 1.0, False]
    c = o_heat.courant_py(1.0, dt, dx)
    step = lambda q, qb, gg: o_heat.step_py(q[0], qb[0], gg, c, False)[None]
    want = oracle_stages(step, ic, 2, 2 * steps)
    for phases in ((True, False) if swept else (False,)):
        sol = gpu_solve("heat", ic, g, 1, 2, b, swept=swept, exact=True, tuning=swept_kernel(phases))
        assert np.array_equal(sol.final_state, want)
        assert np.allclose(sol.final_state[0], ic[0] * np.exp(-8 * np.pi ** 2 * steps * dt))


@pytest.mark.parametrize("b", [12, 24])
def test_reference_block_sizes_euler(device, b):
    """Euler with the reference's non-power-of-two block sizes (k=2: b=12
    gives P=2, b=24 gives P=5): swept == standard == oracle, bitwise."""
    n, steps = 96 if b == 12 else 120, 15
    ic, g, step = euler_case(n, steps, eps=4.0)
    want = oracle_stages(step, ic, 2, 2 * steps)
    for phases in (True, False):
        sw = gpu_solve("euler", ic, g, 2, 2, b, swept=True, exact=True, tuning=swept_kernel(phases))
        assert np.array_equal(sw.final_state, want), phases


def test_reference_heat_forward_euler_120_b24(device):
    """testHeatForwardEuler (tests.py:410-464): 120^2, b=24, against the
    analytic solution with numpy.allclose's default tolerances."""
    n = 120
    dx = 1.0 / n
    dt = 0.1 * dx * dx
    steps = 200
    x = o_heat.grid(n)
    ic = np.outer(np.sin(2 * np.pi * x), np.sin(2 * np.pi * x))[None]
    g = [0.0, steps * dt, dt, dx, dx, 1.0, True]
    exact = ic[0] * np.exp(-8 * np.pi ** 2 * steps * dt)
    for phases in (True, False):
        sol = gpu_solve("heat", ic, g, 1, 1, 24, swept=True, exact=True, tuning=swept_kernel(phases))
        assert np.allclose(sol.final_state[0], exact)


def test_reference_euler_vortex_120_b24(device):
    """testEulerVortex (tests.py:106-163): 120^2, 200 steps, b=24: the
    oracle's bits, and the advected vortex within the scheme's truncation
    error (the reference prints its maximum error)."""
    n, steps = 120, 200
    ic, g, step = euler_case(n, steps)
    sol = gpu_solve("euler", ic, g, 2, 2, 24, swept=True, exact=True, tuning={"band": True})
    assert np.array_equal(sol.final_state, oracle_stages(step, ic, 2, 2 * steps))
    ph = gpu_solve("euler", ic, g, 2, 2, 24, swept=True, exact=True, tuning={"phases": True})
    assert np.array_equal(ph.final_state, sol.final_state)
    X, Y = np.meshgrid(o_euler.grid(n), o_euler.grid(n), indexing="ij")
    ex = o_euler.vortex(X, Y, steps * g[2], 1.4)
    assert np.abs(sol.final_state[0] - ex[0]).max() < 0.2


@pytest.mark.parametrize("pad", [0, 1, 2, 3])
def test_padded_lds_rows(device, pad):
    """Swept tiles whose LDS rows are padded (SY = b + pad) give the same bits."""
    from pysweep_amd import device as dev, hip, plugins, schedule
    from pysweep_amd.equations import path as eq_path
    n, steps = 64, 7
    ic, g, step = euler_case(n, steps, eps=4.0)
    cpu = plugins.load_cpu_module(eq_path("euler.py"))
    cpu.set_globals(*g)
    run = dev.SweptRun(eq_path("euler.cu"), 4, 2, 2, n, n, 16, pad=pad)
    run.set_globals(cpu, g)
    run.upload_ic(ic)
    times = schedule.swept_times(steps, 16, 2, 2, exact=True)
    last = run.run(times)
    hip.synchronize()
    assert np.array_equal(run.level(last), oracle_stages(step, ic, 2, 2 * steps))
    run.close()


@pytest.mark.parametrize("name,b,slots", [("euler", 16, 2), ("euler", 16, 3), ("euler", 32, 2), ("euler", 32, 3),
                                          ("heat", 16, 2), ("heat", 32, 2), ("heat", 32, 4)])
def test_in_place_window(device, name, b, slots):
    """A window of ITS levels (level r written over level r-2 in LDS, for
    level updates that work in place) gives the same bits as larger windows
    and the oracle, with RK2 (two intermediate steps)."""
    from pysweep_amd import device as dev, hip, plugins, schedule
    from pysweep_amd.equations import path as eq_path
    n, steps = 64, 9
    if name == "euler":
        ic, g, step = euler_case(n, steps, eps=4.0)
        nv, k = 4, 2
    else:
        ic, g, step = heat_case(n, False, steps, seed=3)
        nv, k = 1, 1
    cpu = plugins.load_cpu_module(eq_path(name + ".py"))
    cpu.set_globals(*g)
    run = dev.SweptRun(eq_path(name + ".cu"), nv, k, 2, n, n, b, slots=slots)
    assert run.spec.slots == slots
    run.set_globals(cpu, g)
    run.upload_ic(ic)
    times = schedule.swept_times(steps, b, k, 2, exact=True)
    last = run.run(times)
    hip.synchronize()
    assert np.array_equal(run.level(last), oracle_stages(step, ic, 2, 2 * steps))
    run.close()


def test_fp32_state_error_is_the_schemes_one_ulp_sensitivity(device):
    """The FP32-state solve's error against FP64 is of the size the FP64
    solve itself moves when every initial value moves by one ulp: the
    limiter amplifies any perturbation (Euler vortex 128^2, 60 steps, swept
    b=16).  So no storage precision below FP64 can meet a 1e-5 bound here."""
    from pysweep_amd import device as dev, hip, plugins, schedule, workloads
    from pysweep_amd.equations import path as eq_path
    n, steps, b = 128, 60, 16
    ic = workloads.vortex_state(n, n)
    ulp = np.nextafter(ic, np.where(ic >= 0, np.inf, -np.inf))
    dx = 10.0 / n
    dt = 0.1 * dx
    g = [0.0, steps * dt, dt, dx, dx, 1.4]
    cpu = plugins.load_cpu_module(eq_path("euler.py"))
    cpu.set_globals(*g)
    times = schedule.swept_times(steps, b, 2, 2, exact=True)
    res = {}
    for name, cu, real, init in (("fp64", "euler.cu", "double", ic), ("fp32", "euler_fp32.cu", "float", ic),
                                 ("ulp", "euler.cu", "double", ulp)):
        r = dev.SweptRun(eq_path(cu), 4, 2, 2, n, n, b, real=real)
        r.set_globals(cpu, g)
        r.upload_ic(init)
        last = r.run(times)
        hip.synchronize()
        res[name] = r.level(last)
        r.close()
    ref = res["fp64"]

    def rel(x):
        e = x - ref
        return float(np.sqrt((e * e).sum() / (ref * ref).sum()))

    e32, eulp = rel(res["fp32"]), rel(res["ulp"])
    assert eulp > 1e-4          # one ulp of the input grows past 1e-4 in 60 steps
    assert 0.2 < e32 / eulp < 10.0


def test_fp32_heat_with_sub_ulp_increments(device):
    """FP32-state heat where each step changes u by less than one FP32 ulp
    (512^2, dt = dx^2/640: the mode decays by 8 pi^2 alpha dt = 4.7e-8 of u
    per step): the stochastic storage rounding keeps the error a random walk
    (rounding to nearest would round every increment the same way, about
    1e-5 after these 800 steps), and swept and standard FP32 solves give
    the same bits."""
    from pysweep_amd import device as dev, hip, plugins, schedule, workloads
    from pysweep_amd.equations import path as eq_path
    n, steps, b = 512, 800, 32
    dx = 1.0 / n
    dt = dx * dx / 640
    g = [0.0, steps * dt, dt, dx, dx, 1.0, True]
    ic = workloads.heat_mode(n, n)
    cpu = plugins.load_cpu_module(eq_path("heat.py"))
    cpu.set_globals(*g)
    out = {}
    for name, cu, real in (("fp64", "heat.cu", "double"), ("fp32", "heat_fp32.cu", "float")):
        r = dev.StandardRun(eq_path(cu), 1, 1, 1, n, n, real=real)
        r.set_globals(cpu, g)
        r.upload_ic(ic)
        r.run(1, steps + 1)
        out[name] = r.level(steps)
        r.close()
    sw = dev.SweptRun(eq_path("heat_fp32.cu"), 1, 1, 1, n, n, b, real="float")
    sw.set_globals(cpu, g)
    sw.upload_ic(ic)
    last = sw.run(schedule.swept_times(steps, b, 1, 1, exact=True))
    hip.synchronize()
    swept32 = sw.level(last)
    sw.close()
    ref = out["fp64"]
    e = out["fp32"] - ref
    l2 = float(np.sqrt((e * e).sum() / (ref * ref).sum()))
    linf = float(np.abs(e).max() / np.abs(ref).max())
    print(f"fp32 heat 512^2, 800 sub-ulp steps: rel L2 {l2:.2e}, Linf {linf:.2e}")
    assert l2 < 4e-6 and linf < 1e-5
    assert np.array_equal(swept32, out["fp32"])


@pytest.mark.parametrize("pfd", [1, 2, 4, 6])
def test_face_prefetch_depth_heat(device, pfd):
    """Faces fetched 1..6 levels ahead (PS_PFD) give the oracle's bits:
    heat 256^2, b=64 (31-level phases, the rolled level loop unrolled by
    the depth, including depths that do not divide the level count)."""
    ic, g, step = heat_case(256, 1, 100, seed=2)
    s = gpu_solve("heat", ic, g, 1, 1, 64, tuning={"pfd": pfd, "phases": True})
    assert np.array_equal(s.final_state, oracle_stages(step, ic, 1, 100))


@pytest.mark.parametrize("pfd", [2, 3])
def test_face_prefetch_depth_euler(device, pfd):
    """The same for Euler 128^2, b=16 (6-level phases, fully unrolled)."""
    ic, g, step = euler_case(128, 20)
    s = gpu_solve("euler", ic, g, 2, 2, 16, tuning={"pfd": pfd, "phases": True})
    assert np.array_equal(s.final_state, oracle_stages(step, ic, 2, 40))


@pytest.mark.parametrize("swept,phases", [(True, False), (True, True), (False, False)])
def test_state_outside_the_division_range_is_solved_with_ieee_division(device, swept, phases):
    """An Euler state scaled by 2^-130 (rho, E and p below the range in which
    the plugin's fast quotients are the IEEE ones) trips EU_RANGE_FAULT; the
    Solver solves again with IEEE division and square root and gives the
    oracle's bits.  The same problem unscaled does not trip it."""
    import warnings
    ic, g, step = euler_case(64, 6)
    small = ic * 2.0 ** -130
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        s = gpu_solve("euler", small, g, 2, 2, 16, swept=swept, tuning=swept_kernel(phases))
    assert s.division_rerun and any("division" in str(x.message) for x in w)
    assert np.array_equal(s.final_state, oracle_stages(step, small, 2, 12))
    s2 = gpu_solve("euler", ic, g, 2, 2, 16, swept=swept, tuning=swept_kernel(phases))
    assert not s2.division_rerun


def test_engine_marks_tickets_of_a_class_that_left_the_division_range(device):
    """The batched engine cannot re-run one problem of a class, so it marks
    the class's tickets and warns; a class that stayed in range is not
    marked."""
    import warnings

    from pysweep_amd.batch import Engine, Problem
    ic, g, _ = euler_case(64, 4)
    hic, hg, _ = heat_case(64, 1, 10)
    eng = Engine()
    with warnings.catch_warnings(record=True):
        warnings.simplefilter("always")
        t1 = eng.submit(Problem("euler", ic * 2.0 ** -130, g, 2), arrival=0.0)
        t2 = eng.submit(Problem("heat", hic, hg, 1), arrival=0.0)
        eng.run()
    assert t1.division_exact is False and t2.division_exact is True
    assert ("euler", "fp64", 2, t1.cls.b) in eng.division_faults


@pytest.mark.parametrize("name,b,k,s,fe", [("heat", 8, 1, 1, 1), ("heat", 32, 1, 2, 0), ("euler", 16, 2, 2, None),
                                           ("euler", 32, 2, 2, None)])
def test_nan_poisoned_ring_swept_equals_oracle(device, name, b, k, s, fe):
    """Debug mode: the ring is filled with NaN before the initial condition
    goes in, so any read of a cell no phase wrote (a skipped private cell,
    a face outside the sets) would turn the result into NaN.  The swept and
    standard results still equal the oracle bitwise."""
    if name == "heat":
        ic, g, step = heat_case(128, fe, 40)
        stages = 40 * s
    else:
        ic, g, step = euler_case(128, 12)
        stages = 24
    want = oracle_stages(step, ic, s, stages)
    for swept, phases in ((True, True), (True, False), (False, False)):
        r = gpu_solve(name, ic, g, k, s, b, swept=swept, tuning={"poison": True, **swept_kernel(phases)})
        assert np.isfinite(r.final_state).all()
        assert np.array_equal(r.final_state, want), (swept, phases)


def test_nan_poisoned_engine_rings(device):
    """The engine's debug mode: every ring NaN before its initial condition;
    the batched mix (swept classes with private cells skipped) still gives
    the oracle's bits."""
    from pysweep_amd.batch import Engine, Problem
    cases = [heat_case(64, 1, 30), heat_case(128, 0, 20, seed=1), euler_case(64, 9), euler_case(128, 5)]
    probs = [Problem("heat", cases[0][0], cases[0][1], 1), Problem("heat", cases[1][0], cases[1][1], 2),
             Problem("euler", cases[2][0], cases[2][1], 2), Problem("euler", cases[3][0], cases[3][1], 2)]
    eng = Engine(poison=True)
    tickets = [eng.submit(p, arrival=0.0) for p in probs]
    eng.run()
    for t, p, (ic, g, step) in zip(tickets, probs, cases):
        steps = int(round((g[1] - g[0]) / g[2]))
        assert np.array_equal(t.result, oracle_stages(step, ic, p.its, p.its * steps)), p.plugin
