"""Per-kernel launch timing through the C ABI (ps_*_run_timed): every
launch, or a sample of the launches of chosen kernels; timing never
changes results."""
import numpy as np
import pytest

from gpu_helpers import euler_case

pytestmark = pytest.mark.gpu


def _swept(steps=10, b=16):
    from pysweep_amd import device as dev, plugins, schedule
    from pysweep_amd.equations import path as eq_path
    n = 64
    ic, g, _ = euler_case(n, steps, eps=4.0)
    cpu = plugins.load_cpu_module(eq_path("euler.py"))
    cpu.set_globals(*g)
    run = dev.SweptRun(eq_path("euler.cu"), 4, 2, 2, n, n, b)
    run.set_globals(cpu, g)
    run.keep_ic_on_device(ic)
    return run, schedule.swept_times(steps, b, 2, 2, exact=True)


def test_every_launch_and_sampled_launches(device):
    from pysweep_amd import hip
    run, times = _swept()
    G = times.G
    assert G == 5 and times.tail > 0
    run.reset()
    last = run.run(times)
    hip.synchronize()
    plain = run.level(last)

    full = hip.RunStats()
    run.reset()
    run.run(times, stats=full)
    hip.synchronize()
    d = full.as_dict()
    assert np.array_equal(run.level(last), plain)
    want = {"UpPyramid": 1, "YBridge": G + 1, "XBridge": G + 1, "Octahedron": G, "DownPyramid": 1,
            "Standard": times.tail, "Band": 0}
    assert {k: v["launches"] for k, v in d.items()} == want
    assert all(v["ms"] > 0 for v in d.values() if v["launches"])

    samp = hip.RunStats.sampled(["Octahedron", "Standard"], 4)
    run.reset()
    run.run(times, stats=samp)
    hip.synchronize()
    d = samp.as_dict()
    assert np.array_equal(run.level(last), plain)
    # launches 0, 4, 8, ... of each sampled kind within a run call
    assert {k: v["launches"] for k, v in d.items()} == {
        "UpPyramid": 0, "YBridge": 0, "XBridge": 0, "Octahedron": -(-G // 4), "DownPyramid": 0,
        "Standard": -(-times.tail // 4), "Band": 0}
    assert d["Octahedron"]["ms"] > 0 and d["YBridge"]["ms"] == 0
    run.close()


def test_band_launches_are_timed(device):
    """The band solve's launches (one per RK2 step, the last one shorter
    when the stage count is odd) are counted and timed under PS_K_BAND."""
    from pysweep_amd import device as dev, hip, plugins, schedule
    from pysweep_amd.equations import path as eq_path
    n, steps = 64, 10
    ic, g, _ = euler_case(n, steps, eps=4.0)
    cpu = plugins.load_cpu_module(eq_path("euler.py"))
    cpu.set_globals(*g)
    run = dev.BandRun(eq_path("euler.cu"), 4, 2, 2, n, n)
    run.set_globals(cpu, g)
    run.keep_ic_on_device(ic)
    run.reset()
    times = schedule.swept_times(11, 16, 2, 2, exact=False)  # 21 stages: ten launches of 2, one of 1
    assert times.stages == 21
    st = hip.RunStats()
    last = run.run(times, stats=st)
    hip.synchronize()
    assert last == 1 + times.stages
    d = st.as_dict()
    assert d["Band"]["launches"] == -(-times.stages // 2) and d["Band"]["ms"] > 0
    assert sum(v["launches"] for v in d.values()) == d["Band"]["launches"]
    run.close()
