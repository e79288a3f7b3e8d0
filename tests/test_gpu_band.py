"""GPU parity of the band kernel (the swept solve of plugins with a point
update, heat, or a band sweep of their own, Euler) through the C ABI:
against the oracle on ragged, odd and tiny grids (periodic wrap of rows and
columns several times over), for every launch depth and row segment, FP64
and FP32 state, forward Euler and RK2, with and without the full history.
Bitwise throughout: the point update and the band sweep evaluate the
plugins' arithmetic in its order."""
import numpy as np
import pytest

from gpu_helpers import gpu_solve, heat_case, o_heat, oracle_stages, swept_kernel

pytestmark = pytest.mark.gpu


def _heat_grid(nx, ny, fe, steps, seed=0):
    rng = np.random.default_rng(seed)
    dx = 1.0 / nx
    dt = 0.1 * dx * dx
    x = np.linspace(0, 1, nx, endpoint=False)
    y = np.linspace(0, 1, ny, endpoint=False)
    ic = (np.outer(np.sin(2 * np.pi * x), np.cos(4 * np.pi * y)) + 1e-3 * rng.random((nx, ny)))[None]
    g = [0.0, steps * dt, dt, dx, dx, 1.0, fe]
    c = o_heat.courant_py(1.0, dt, dx)
    step = lambda q, qb, gg: o_heat.step_py(q[0], None if qb is None else qb[0], gg, c, fe)[None]
    return ic, g, step


def _band(ic, g, its, steps, levels=None, rows=None, real="double", cu="heat.cu", on_level=None):
    from pysweep_amd import device as dev, hip, plugins, schedule
    from pysweep_amd.equations import path as eq_path
    nv, nx, ny = ic.shape
    cpu = plugins.load_cpu_module(eq_path("heat.py"))
    cpu.set_globals(*g)
    run = dev.BandRun(eq_path(cu), 1, 1, its, nx, ny, levels=levels, rows=rows, real=real)
    run.set_globals(cpu, g)
    run.upload_ic(ic)
    times = schedule.standard_times(steps, 8, 1, its)
    last = run.run(times, on_level=on_level)
    hip.synchronize()
    out = run.level(last)
    assert last == its - 1 + its * steps
    return run, out


@pytest.mark.parametrize("rows", [48, None], ids=["rows48", "rows_auto"])
@pytest.mark.parametrize("fe", [True, False], ids=["fe", "rk2"])
@pytest.mark.parametrize("nx,ny", [(64, 96), (130, 66), (37, 45), (3, 200), (200, 3), (5, 7), (96, 257), (1, 2)])
def test_band_matches_oracle_on_ragged_and_tiny_grids(device, nx, ny, fe, rows):
    """Odd ny (no paired loads), ny below one band (columns wrap several
    times), nx below the halo (rows wrap several times, down to one row),
    bands and row segments that do not divide the grid, with a fixed segment
    and the model's; 13 steps, so the last launch is shorter than the
    others."""
    its = 1 if fe else 2
    steps = 13
    ic, g, step = _heat_grid(nx, ny, fe, steps, seed=nx + ny)
    run, got = _band(ic, g, its, steps, rows=rows)
    assert np.array_equal(got, oracle_stages(step, ic, its, its * steps))
    run.close()


@pytest.mark.parametrize("levels", [2, 4, 6, 8, 12, 16])
def test_band_launch_depths_give_the_same_bits(device, levels):
    """Every depth (levels per launch) and several row segments equal the
    standard kernel bitwise (heat 384 x 320, forward Euler, 50 steps)."""
    from pysweep_amd import device as dev, plugins
    from pysweep_amd.equations import path as eq_path
    nx, ny, steps = 384, 320, 50
    ic, g, _ = _heat_grid(nx, ny, True, steps, seed=3)
    cpu = plugins.load_cpu_module(eq_path("heat.py"))
    cpu.set_globals(*g)
    st = dev.StandardRun(eq_path("heat.cu"), 1, 1, 1, nx, ny)
    st.set_globals(cpu, g)
    st.upload_ic(ic)
    st.run(1, steps + 1)
    want = st.level(steps)
    st.close()
    for rows in (1, 17, 128, 1000):
        run, got = _band(ic, g, 1, steps, levels=levels, rows=rows)
        assert np.array_equal(got, want), rows
        run.close()


def test_band_rk2_levels_are_multiples_of_its(device):
    """RK2 launches start on full steps: depths that are not a multiple of
    the two stages are refused before anything runs."""
    from pysweep_amd import kernels
    from pysweep_amd.equations import path as eq_path
    with pytest.raises(ValueError):
        kernels.band_spec(eq_path("heat.cu"), 1, 1, 2, levels=6 + 1)
    kernels.band_spec(eq_path("heat.cu"), 1, 1, 2, levels=6)


def test_band_run_rejects_a_start_off_a_full_step(device):
    """ps_band_run checks its arguments on the host and launches nothing
    when they are invalid."""
    from pysweep_amd import hip
    ic, g, _ = _heat_grid(32, 32, False, 4)
    run, _ = _band(ic, g, 2, 4)
    with pytest.raises(hip.HipError, match="full-step"):
        hip.band_run(run.plan, 2, 4)
    bad = hip.BandPlan(state=run.plan.state, nx=32, ny=32, its=2, T=run.T, levels=7, wout=114, rows=8, store_all=0,
                       band=run.plan.band)
    with pytest.raises(hip.HipError, match="levels"):
        hip.band_run(bad, 1, 4)
    run.close()


@pytest.mark.parametrize("fe", [True, False], ids=["fe", "rk2"])
def test_band_history_equals_oracle_every_step(device, fe):
    """With on_level (the Solver's history and HDF5 output) every level is
    written and each full step handed over is the oracle's."""
    its = 1 if fe else 2
    steps = 21
    ic, g, step = _heat_grid(96, 64, fe, steps, seed=7)
    seen = {}
    holder = {}

    def on_level(A):
        seen[A] = holder["run"].level(A)

    from pysweep_amd import device as dev, hip, plugins, schedule
    from pysweep_amd.equations import path as eq_path
    cpu = plugins.load_cpu_module(eq_path("heat.py"))
    cpu.set_globals(*g)
    run = dev.BandRun(eq_path("heat.cu"), 1, 1, its, 96, 64, rows=40)
    holder["run"] = run
    run.set_globals(cpu, g)
    run.upload_ic(ic)
    run.run(schedule.standard_times(steps, 8, 1, its), on_level=on_level)
    hip.synchronize()
    assert sorted(seen) == [its - 1 + its * m for m in range(1, steps + 1)]
    for m in (1, 2, 8, 9, 17, steps):
        assert np.array_equal(seen[its - 1 + its * m], oracle_stages(step, ic, its, its * m)), m
    run.close()


def test_solver_history_through_the_band_kernel(device):
    """The Solver's history (the reference's output dataset) with the
    reference's step quantization, through the band kernel."""
    n, steps = 64, 30
    ic, g, step = heat_case(n, False, steps, seed=4)
    sol = gpu_solve("heat", ic, g, 1, 2, 8, swept=True, exact=False, history=True)
    from pysweep_amd import device as dev
    assert isinstance(sol._run, dev.BandRun)
    for m in range(sol.data.shape[0]):
        assert np.array_equal(sol.data[m], oracle_stages(step, ic, 2, 2 * m)), m


def test_band_fp32_state_equals_the_other_kernels(device):
    """FP32 state (stochastic storage rounding keyed on the value and the
    stage): the band kernel gives the bits of the FP32 standard kernel and
    of the FP32 phase kernels, 1024^2, 100 steps."""
    from pysweep_amd import device as dev, hip, plugins, schedule, workloads
    from pysweep_amd.equations import path as eq_path
    n, steps = 1024, 100
    dx = 1.0 / n
    dt = 0.1 * dx * dx
    g = [0.0, steps * dt, dt, dx, dx, 1.0, True]
    ic = workloads.heat_mode(n, n)
    cpu = plugins.load_cpu_module(eq_path("heat.py"))
    cpu.set_globals(*g)
    st = dev.StandardRun(eq_path("heat_fp32.cu"), 1, 1, 1, n, n, real="float")
    st.set_globals(cpu, g)
    st.upload_ic(ic)
    st.run(1, steps + 1)
    want = st.level(steps)
    st.close()
    run, got = _band(ic, g, 1, steps, real="float", cu="heat_fp32.cu")
    assert np.array_equal(got, want)
    run.close()
    sw = dev.SweptRun(eq_path("heat_fp32.cu"), 1, 1, 1, n, n, 32, real="float")
    sw.set_globals(cpu, g)
    sw.upload_ic(ic)
    last = sw.run(schedule.swept_times(steps, 32, 1, 1, exact=True))
    hip.synchronize()
    assert np.array_equal(sw.level(last), want)
    sw.close()


# ------------------------------------------------------------------ Euler

def _euler_band(ic, g, steps, rows=None, real="double", cu="euler.cu"):
    from pysweep_amd import device as dev, hip, plugins, schedule
    from pysweep_amd.equations import path as eq_path
    nv, nx, ny = ic.shape
    cpu = plugins.load_cpu_module(eq_path("euler.py"))
    cpu.set_globals(*g)
    run = dev.BandRun(eq_path(cu), 4, 2, 2, nx, ny, rows=rows, real=real)
    run.set_globals(cpu, g)
    run.upload_ic(ic)
    last = run.run(schedule.standard_times(steps, 16, 2, 2))
    hip.synchronize()
    assert last == 1 + 2 * steps
    return run, run.level(last)


@pytest.mark.parametrize("nx,ny,rows", [(64, 64, 16), (48, 80, 7), (80, 48, 32), (20, 36, 128), (130, 70, 40),
                                         (3, 40, None), (36, 20, None), (130, 70, None)])
def test_euler_band_matches_oracle(device, nx, ny, rows):
    """The Euler plugin's band sweep (one column per lane, the radius-2
    faces and pressures carried in registers) against the oracle bitwise on
    grids whose sides are not multiples of the band or the row segment, and
    on grids narrower than one band (columns wrap)."""
    from oracle import euler as o_euler
    steps = 6
    dx = 10.0 / nx
    dt = 0.1 * dx
    X, Y = np.meshgrid(o_euler.grid(nx), np.linspace(-5, 5, ny, endpoint=False), indexing="ij")
    ic = o_euler.vortex(X, Y, 0.0, 1.4, eps=3.0)
    g = [0.0, steps * dt, dt, dx, dx, 1.4]
    want = oracle_stages(lambda q, qb, gg: o_euler.step_py(q, qb, gg, 1.4, dt, dx, dx), ic, 2, 2 * steps)
    run, got = _euler_band(ic, g, steps, rows=rows)
    assert np.array_equal(got, want)
    assert not run.range_fault()
    run.close()


def test_euler_band_equals_standard_and_phases(device):
    """256^2 vortex, 40 steps: band == standard == five phases (b=16), and
    through the Solver with the reference's step quantization."""
    from gpu_helpers import euler_case
    ic, g, _ = euler_case(256, 40, eps=4.0, xc=1.0, yc=-1.5)
    std = gpu_solve("euler", ic, g, 2, 2, 16, swept=False, exact=True)
    sw = gpu_solve("euler", ic, g, 2, 2, 16, swept=True, exact=True, tuning={"phases": True})
    run, got = _euler_band(ic, g, 40, rows=64)
    assert np.array_equal(got, std.final_state)
    assert np.array_equal(sw.final_state, std.final_state)
    run.close()
    bs = gpu_solve("euler", ic, g, 2, 2, 16, swept=True, exact=False, tuning=swept_kernel(False))
    from pysweep_amd import device as dev
    assert isinstance(bs._run, dev.BandRun)
    ps = gpu_solve("euler", ic, g, 2, 2, 16, swept=True, exact=False, tuning={"phases": True})
    assert np.array_equal(bs.final_state, ps.final_state)


def test_euler_band_fp32_equals_standard_fp32(device):
    """FP32 state: the band sweep keeps each stage rounded to float in its
    register rows, as the other kernels keep it in LDS and HBM."""
    from gpu_helpers import euler_case
    from pysweep_amd import device as dev, plugins
    from pysweep_amd.equations import path as eq_path
    n, steps = 192, 25
    ic, g, _ = euler_case(n, steps, eps=4.0)
    cpu = plugins.load_cpu_module(eq_path("euler.py"))
    cpu.set_globals(*g)
    st = dev.StandardRun(eq_path("euler_fp32.cu"), 4, 2, 2, n, n, real="float")
    st.set_globals(cpu, g)
    st.upload_ic(ic)
    st.run(1, 2 * steps + 1)
    want = st.level(2 * steps + 1)
    st.close()
    run, got = _euler_band(ic, g, steps, rows=48, real="float", cu="euler_fp32.cu")
    assert np.array_equal(got, want)
    run.close()


def test_euler_band_flags_states_outside_the_division_range(device):
    """The band sweep range-checks the cells it reads, as the other kernels
    do: a state scaled by 2^-130 trips EU_RANGE_FAULT and the Solver solves
    again with IEEE division."""
    import warnings

    from gpu_helpers import euler_case
    ic, g, step = euler_case(64, 6)
    small = ic * 2.0 ** -130
    with warnings.catch_warnings(record=True):
        warnings.simplefilter("always")
        s = gpu_solve("euler", small, g, 2, 2, 16, swept=True, tuning=swept_kernel(False))
    assert s.division_rerun
    assert np.array_equal(s.final_state, oracle_stages(step, small, 2, 12))
    s2 = gpu_solve("euler", ic, g, 2, 2, 16, swept=True, tuning=swept_kernel(False))
    assert not s2.division_rerun


def test_solver_picks_the_band_kernel_by_plugin_and_size(device):
    """Heat's swept solves take the band kernel at every size; Euler's from
    2^17 cells (below, the phase kernels: the band's 64-column bands leave
    small grids too few wavefronts)."""
    from gpu_helpers import euler_case
    from pysweep_amd import device as dev
    ic, g, _ = heat_case(64, True, 5)
    assert isinstance(gpu_solve("heat", ic, g, 1, 1, 8)._run, dev.BandRun)
    ic, g, _ = euler_case(256, 2)
    assert isinstance(gpu_solve("euler", ic, g, 2, 2, 16)._run, dev.SweptRun)
    ic, g, _ = euler_case(384, 2)
    assert isinstance(gpu_solve("euler", ic, g, 2, 2, 16)._run, dev.BandRun)


def test_engine_runs_lone_problems_on_the_band_kernel(device):
    """A problem running alone in the engine takes its standard stages on
    the band kernel where that outruns the standard kernel (Euler from
    1024^2 cells) and its ring holds a launch's levels: an Euler 1024^2
    solve in one run of rounds, and again in runs of three rounds, so that
    every other run starts on a half step (one standard stage realigns it,
    and a run may end on one).  Results equal the plain solves bitwise;
    heat problems, whose band launches are 12 levels deep, keep the
    standard kernel in rings of s+1 levels."""
    from gpu_helpers import euler_case
    from pysweep_amd import baseline
    from pysweep_amd.batch import Engine, Problem
    ic, g, _ = euler_case(1024, 10)
    plain = None
    for rounds, lo, hi in ((1 << 20, 18, 20), (3, 1, 19)):
        eng = Engine(tuner=None, max_rounds_ahead=rounds)
        t = eng.submit(Problem("euler", ic, g, 2, method="standard"), arrival=0.0)
        eng.run()
        assert lo <= eng.band_stages <= hi, eng.band_stages
        plain = baseline.solve_plain(t.problem) if plain is None else plain
        assert np.array_equal(t.result, plain)
    eng = Engine(tuner=None, band_solo=False)
    t = eng.submit(Problem("euler", ic, g, 2, method="standard"), arrival=0.0)
    eng.run()
    assert eng.band_stages == 0 and np.array_equal(t.result, plain)
    h_ic, h_g, h_step = heat_case(64, False, 10, seed=2)
    eng = Engine(tuner=None)
    th = eng.submit(Problem("heat", h_ic, h_g, 2, method="standard"), arrival=0.0)
    eng.run()
    assert eng.band_stages == 0 and np.array_equal(th.result, oracle_stages(h_step, h_ic, 2, 20))


def test_euler_history_through_the_band_kernel(device):
    """The Solver's history of an Euler solve on the band kernel (384^2, the
    default there): every full step, written by launches that record every
    level, equals the standard solve's, with the reference's quantization
    (an odd stage count: the last launch takes one stage)."""
    from gpu_helpers import euler_case
    from pysweep_amd import device as dev
    n, steps = 384, 8
    ic, g, _ = euler_case(n, steps, eps=4.0)
    sw = gpu_solve("euler", ic, g, 2, 2, 16, swept=True, exact=False, history=True)
    assert isinstance(sw._run, dev.BandRun) and sw.times.stages == 15
    st = gpu_solve("euler", ic, g, 2, 2, 16, swept=False, exact=True, history=True)
    assert sw.data.shape[0] == 8
    for m in range(sw.data.shape[0]):
        assert np.array_equal(sw.data[m], st.data[m]), m
