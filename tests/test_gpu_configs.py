"""GPU tests at the workloads' own scale: the 512-problem batched mix, a
Poisson stream of Euler solves of up to 2048^2 with FP32-state requests, and
the FP32-state precision study at 8192^2 (heat) and 2048^2 (Euler).  The CPU
oracle cannot run these sizes in a test; size-independent properties stand
in for it: every engine result equals the same problem's one-at-a-time plain
solve bitwise (the plain solve being bitwise the oracle at the sizes the
parity tests reach), the periodic schemes conserve their sums to rounding,
and swept and standard solves give the same bits."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", ["swept", "auto"])
def test_batched_mix_of_512_problems_equals_plain_solves(device, method):
    """Half heat, half Euler, sides 64..1024 (multiples of 32, non-square
    allowed), 200 steps, FP64, through the default (tuner-driven) engine, on
    the swept method or "auto" (Euler on the band method, heat on the
    standard one): all 512 results equal the one-at-a-time plain solves
    bitwise; conserved sums move by rounding only; HBM in use stays within
    the budget."""
    from pysweep_amd import baseline, batch, workloads
    probs = workloads.batched_mix(n=512, steps=200, seed=0)
    for p in probs:
        p.method = method
    eng = batch.Engine()
    tickets = [eng.submit(p, arrival=0.0) for p in probs]
    eng.run()
    assert all(t.done for t in tickets)
    assert eng.peak_in_use <= eng.hbm_budget
    for t in tickets:
        assert t.conservation["drift"] < 1e-13, (t.problem.tag, t.conservation)
        assert t.division_exact
    if method == "auto":
        assert {t.method for t in tickets if t.problem.plugin == "euler"} == {"band"}
    for t, p in zip(tickets, probs):
        assert np.array_equal(t.result, baseline.solve_plain(p)), p.tag


def test_served_stream_with_fp32_requests_equals_plain_solves(device):
    """Poisson arrivals (compressed to 30 per second) of 24 Euler vortex
    solves, sides 128..2048, 100..2000 steps, ~30 % with FP32 state, on the
    serving policy: each result equals the same request served alone by the
    plain solver, bitwise (FP32 ones too), p50 <= p99, and the FP32 results'
    conserved sums move by less than 1e-5."""
    from pysweep_amd import baseline, batch, workloads
    stream = workloads.served_stream(n=24, rate=30.0, seed=1)
    assert any(p.precision == "fp32" for _, p in stream)
    for _, p in stream:
        p.method = "auto"
    eng = batch.Engine(policy="oldest_first", capacity=1 << 19)
    tickets = [eng.submit(p, arrival=t) for t, p in stream]
    eng.run()
    lat = [t.latency for t in tickets]
    assert batch.percentile(lat, 50) <= batch.percentile(lat, 99)
    for t, (_, p) in zip(tickets, stream):
        assert np.array_equal(t.result, baseline.solve_plain(p)), (p.tag, p.precision, p.ic.shape)
        assert t.conservation["drift"] < (1e-5 if p.precision == "fp32" else 1e-13), t.conservation


def test_fp32_heat_at_8192_meets_the_precision_target(device):
    """The precision study's heat case: 8192^2, 1000 forward-Euler steps
    (dt = 0.1 dx^2, each step's increment about one FP32 ulp): the FP32-state
    solve is within 1e-5 relative L2 of the FP64 one, and its swept (b=64)
    and standard solves give the same bits."""
    from pysweep_amd import device as dev, hip, plugins, schedule, workloads
    from pysweep_amd.equations import path as eq_path
    n, steps = 8192, 1000
    dx = 1.0 / n
    dt = 0.1 * dx * dx
    g = [0.0, steps * dt, dt, dx, dx, 1.0, True]
    cpu = plugins.load_cpu_module(eq_path("heat.py"))
    cpu.set_globals(*g)
    ic = workloads.heat_mode(n, n)
    out = {}
    for name, cu, real in (("fp64", "heat.cu", "double"), ("fp32", "heat_fp32.cu", "float")):
        r = dev.StandardRun(eq_path(cu), 1, 1, 1, n, n, real=real)
        r.set_globals(cpu, g)
        r.upload_ic(ic)
        r.run(1, steps + 1)
        out[name] = r.level(steps)
        r.close()
    sw = dev.SweptRun(eq_path("heat_fp32.cu"), 1, 1, 1, n, n, 64, real="float")
    sw.set_globals(cpu, g)
    sw.upload_ic(ic)
    last = sw.run(schedule.swept_times(steps, 64, 1, 1, exact=True))
    hip.synchronize()
    assert np.array_equal(sw.level(last), out["fp32"])
    sw.close()
    ref = out["fp64"]
    e = out["fp32"] - ref
    l2 = float(np.sqrt((e * e).sum() / (ref * ref).sum()))
    print(f"fp32 heat 8192^2, 1000 steps: rel L2 {l2:.2e}")
    assert l2 < 1e-5


def test_fp32_euler_at_2048_error_is_the_one_ulp_sensitivity(device):
    """The precision study's Euler case at its size (2048^2, b=16, 100
    steps): the FP32-state solve's error against FP64 is of the size the
    FP64 solve itself moves when every initial value moves by one ulp, and
    the FP32 swept and standard solves give the same bits."""
    from pysweep_amd import device as dev, hip, plugins, schedule, workloads
    from pysweep_amd.equations import path as eq_path
    n, steps, b = 2048, 100, 16
    dx = 10.0 / n
    dt = 0.1 * dx
    g = [0.0, steps * dt, dt, dx, dx, 1.4]
    cpu = plugins.load_cpu_module(eq_path("euler.py"))
    cpu.set_globals(*g)
    ic = workloads.vortex_state(n, n)
    ulp = np.nextafter(ic, np.where(ic >= 0, np.inf, -np.inf))
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
    st = dev.StandardRun(eq_path("euler_fp32.cu"), 4, 2, 2, n, n, real="float")
    st.set_globals(cpu, g)
    st.upload_ic(ic)
    st.run(1, 2 * steps + 1)
    assert np.array_equal(st.level(2 * steps + 1), res["fp32"])
    st.close()
    ref = res["fp64"]

    def rel(x):
        e = x - ref
        return float(np.sqrt((e * e).sum() / (ref * ref).sum()))

    e32, eulp = rel(res["fp32"]), rel(res["ulp"])
    print(f"euler 2048^2, 100 steps: fp32 rel L2 {e32:.2e}, fp64 one-ulp input {eulp:.2e}")
    assert 0.2 < e32 / eulp < 10.0
