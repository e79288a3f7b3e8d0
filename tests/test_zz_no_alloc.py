"""The entry point that allocates nothing (ps_level_sums_noalloc): its sums
equal ps_level_sums' bit for bit, on consecutive calls that share the
library's scratch array (the second one smaller), for FP64 and FP32 levels;
sizes beyond the array are rejected."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fp32", [False, True])
def test_noalloc_sums_equal_the_allocating_entry_point(device, fp32):
    from pysweep_amd import hip
    dtype = np.float32 if fp32 else np.float64
    rng = np.random.default_rng(7)
    for nv, cells in ((8, 300_001), (3, 1_000)):  # the second call smaller
        a = (rng.standard_normal((nv, cells)) * 10 ** rng.uniform(-2, 2, (nv, 1))).astype(dtype)
        lvl = hip.DeviceArray(a.nbytes)
        lvl.upload(a)
        out_a, out_b = hip.DeviceArray(16 * 8), hip.DeviceArray(16 * 8)
        hip.level_sums_async(lvl.ptr, fp32, nv, cells, out_a.ptr)
        hip.level_sums_noalloc_async(lvl.ptr, fp32, nv, cells, out_b.ptr)
        ha, hb = np.empty(2 * nv), np.empty(2 * nv)
        out_a.download(ha)
        out_b.download(hb)
        assert np.array_equal(ha, hb), (nv, cells)
        a64 = a.astype(np.float64)
        assert np.allclose(hb[nv:], np.abs(a64).sum(axis=1), rtol=1e-13, atol=0)
        assert np.all(np.abs(hb[:nv] - a64.sum(axis=1)) <= 1e-13 * hb[nv:])


def test_noalloc_rejects_more_variables_than_its_scratch(device):
    from pysweep_amd import hip
    d = hip.DeviceArray(9 * 64 * 8)
    o = hip.DeviceArray(18 * 8)
    with pytest.raises(hip.HipError, match="nv must be 1..8"):
        hip.level_sums_noalloc_async(d.ptr, False, 9, 64, o.ptr)
