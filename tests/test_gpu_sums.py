"""FP64 sums of a level on the device (ps_level_sums) and the conserved-sum
records of solves: sums agree with numpy's FP64 sums, repeat bitwise, and
the periodic schemes conserve every sum up to rounding."""
import numpy as np
import pytest

from gpu_helpers import euler_case, gpu_solve, heat_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype,nv,n", [(np.float64, 4, 1000), (np.float32, 4, 777), (np.float64, 1, 1 << 20),
                                        (np.float32, 8, 3), (np.float64, 2, 0)])
def test_level_sums_match_numpy(device, dtype, nv, n):
    from pysweep_amd import hip
    rng = np.random.default_rng(n)
    a = (rng.standard_normal((nv, n)) * 10 ** rng.uniform(-3, 3, (nv, 1))).astype(dtype)
    d = hip.DeviceArray(a.nbytes)
    if n:
        d.upload(a)
    s1, m1 = hip.level_sums(d.ptr, dtype == np.float32, nv, n)
    s2, m2 = hip.level_sums(d.ptr, dtype == np.float32, nv, n)
    assert np.array_equal(s1, s2) and np.array_equal(m1, m2)  # fixed order: same bits every call
    a64 = a.astype(np.float64)
    want, wabs = a64.sum(axis=1), np.abs(a64).sum(axis=1)
    assert np.allclose(m1, wabs, rtol=1e-13, atol=0)
    assert np.all(np.abs(s1 - want) <= 1e-13 * wabs)


def test_level_sums_rejects_bad_arguments(device):
    from pysweep_amd import hip
    d = hip.DeviceArray(1024)
    with pytest.raises(hip.HipError, match="nv must be 1..8"):
        hip.level_sums(d.ptr, False, 9, 4)


@pytest.mark.parametrize("swept", [True, False])
def test_solver_reports_conserved_sums(device, swept):
    ic, g, _ = euler_case(128, 30)
    s = gpu_solve("euler", ic, g, 2, 2, 16, swept=swept)
    c = s.conservation
    want0 = ic.astype(np.float64).sum(axis=(1, 2))
    assert np.allclose(c["initial"], want0, rtol=1e-13)
    assert np.allclose(c["final"], s.final_state.sum(axis=(1, 2)), rtol=1e-12)
    assert c["drift"] < 1e-13, c
    h, gh, _ = heat_case(128, 1, 50)
    sh = gpu_solve("heat", h, gh, 1, 1, 8, swept=swept)
    assert sh.conservation["drift"] < 1e-13, sh.conservation
