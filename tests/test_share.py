"""The CPU/GPU share split (core/process.py:81-176): with share = 0 the
whole swept solve runs on the host through the plugin's Python step (runs
here, no GPU); with 0 < share < 1 the GPU and the host split the block rows
and must give the GPU-only solve's bits."""
import numpy as np
import pytest

from gpu_helpers import euler_case, heat_case, oracle_stages


def _solve(name, ic, g, k, s, b, share, exact=True, history=False):
    from pysweep_amd import Solver
    from pysweep_amd.equations import path as eq_path
    sol = Solver(ic, sendWarning=False)
    sol.simulation, sol.operating, sol.intermediate, sol.blocksize = True, k, s, (b, b, 1)
    sol.share = share
    sol.globals = list(g)
    sol.setCPU(eq_path(name + ".py"))
    sol.setGPU(eq_path(name + ".cu"))
    sol.exact = exact
    sol.history = history
    sol.loadCPUModule()
    sol()
    return sol


def test_share_zero_heat_on_the_host_matches_oracle():
    ic, g, step = heat_case(32, True, 20, seed=3)
    sol = _solve("heat", ic, g, 1, 1, 8, 0.0)
    assert sol.share == 0.0
    assert np.array_equal(sol.final_state, oracle_stages(step, ic, 1, 20))


def test_share_zero_euler_rk2_on_the_host_matches_oracle():
    ic, g, step = euler_case(32, 5, eps=3.0)
    sol = _solve("euler", ic, g, 2, 2, 8, 0.0)
    assert np.array_equal(sol.final_state, oracle_stages(step, ic, 2, 10))


def test_share_zero_reference_quantization_and_history():
    """Reference step quantization (no tail) and every full step recorded."""
    ic, g, step = heat_case(32, False, 12, seed=4)
    sol = _solve("heat", ic, g, 1, 2, 16, 0.0, exact=False, history=True)
    assert sol.data.shape[0] == sol.arrayShape[0]
    for i in (1, sol.arrayShape[0] - 1):
        assert np.array_equal(sol.data[i], oracle_stages(step, ic, 2, 2 * i)), i


def test_share_standard_is_rejected():
    from pysweep_amd import Solver
    from pysweep_amd.equations import path as eq_path
    ic, g, _ = heat_case(32, True, 4)
    sol = Solver(ic, sendWarning=False)
    sol.simulation, sol.blocksize, sol.share, sol.globals = False, (8, 8, 1), 0.5, list(g)
    sol.setCPU(eq_path("heat.py"))
    sol.setGPU(eq_path("heat.cu"))
    with pytest.raises(NotImplementedError):
        sol()


@pytest.mark.gpu
@pytest.mark.parametrize("share", [0.25, 0.5, 0.75])
@pytest.mark.parametrize("exact", [True, False])
def test_share_split_equals_gpu_only(device, share, exact):
    ic, g, step = euler_case(64, 13, eps=4.0)
    split = _solve("euler", ic, g, 2, 2, 16, share, exact=exact)
    whole = _solve("euler", ic, g, 2, 2, 16, 1.0, exact=exact)
    assert 0 < split.share < 1
    assert np.array_equal(split.final_state, whole.final_state)


@pytest.mark.gpu
def test_share_split_heat_history(device):
    ic, g, step = heat_case(64, True, 30, seed=9)
    split = _solve("heat", ic, g, 1, 1, 16, 0.5, history=True)
    for i in (1, 15, 30):
        assert np.array_equal(split.data[i], oracle_stages(step, ic, 1, i)), i


@pytest.mark.gpu
@pytest.mark.parametrize("b", [8, 12, 16, 24])
def test_reference_sweep_heat_rk2_384_at_share_half(device, b):
    """The reference's parameter sweep (tests.py:54-77: 384^2, b in
    {8, 12, 16, 24}, shares 0, 0.5, 1) at share 0.5 for heat RK2: the split
    solve gives the GPU-only solve's bits, and that one the oracle's."""
    ic, g, step = heat_case(384, False, 30, seed=5)
    split = _solve("heat", ic, g, 1, 2, b, 0.5)
    whole = _solve("heat", ic, g, 1, 2, b, 1.0)
    assert split.share == 0.5
    assert np.array_equal(split.final_state, whole.final_state)
    assert np.array_equal(whole.final_state, oracle_stages(step, ic, 2, 60))
