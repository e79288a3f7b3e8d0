"""The Euler plugin's division and square-root helpers against the GPU's
own IEEE division and square root.

euler.cu forms n/d from a shared, Newton-refined reciprocal of d
(eu_recip / eu_quo) instead of the compiler's full division sequence, and
the reciprocal of 2d by halving that of d (eu_twice).  The claim is that
the quotients are bitwise those of `n / d` and `n / (2*d)` whenever the
operands are normal, not within 2^53 of the subnormal range and less than
2^768 apart, and that zeros, infinities and NaNs come out as IEEE division
gives them.  Likewise eu_sqrt is bitwise `sqrt` for x >= 2^-767 and for
zeros, infinities, NaNs and negative x.  The kernels below evaluate both
forms side by side, compiled with the plugin's own source and options.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CHECK_KERNEL = r"""
extern "C" __global__ void ps_div_check(const double *num, const double *den, double *out, int cnt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt)
        return;
    const double a = num[i], b = den[i];
    const eu_rcp r = eu_recip(b);
    out[6 * i + 0] = a / b;
    out[6 * i + 1] = eu_quo(a, r);
    out[6 * i + 2] = a / (2 * b);
    out[6 * i + 3] = eu_quo(a, eu_twice(r));
    out[6 * i + 4] = eu_quo_pos(a, r);
    out[6 * i + 5] = eu_quo_pos(a, eu_twice(r));
}

extern "C" __global__ void ps_sqrt_check(const double *x, double *out, int cnt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt)
        return;
    out[2 * i + 0] = sqrt(x[i]);
    out[2 * i + 1] = eu_sqrt(x[i]);
}
"""


def _module():
    from pysweep_amd import compiler, hip, kernels
    from pysweep_amd.equations import path as eq_path
    spec = kernels.standard_spec(eq_path("euler.cu"), 4, 2, 2, 64, 64)
    code = compiler.compile_cached(spec.source() + CHECK_KERNEL, "euler_div_check")
    return hip.Module(code)


def _run(num, den):
    import ctypes
    from pysweep_amd import hip
    num = np.ascontiguousarray(num, dtype=np.float64)
    den = np.ascontiguousarray(den, dtype=np.float64)
    n = len(num)
    assert den.shape == (n,)
    mod = _module()
    fn = mod.get_function("ps_div_check")
    dn, dd, do = hip.DeviceArray(8 * n), hip.DeviceArray(8 * n), hip.DeviceArray(48 * n)
    dn.upload(num)
    dd.upload(den)
    block = 256
    grid = (n + block - 1) // block  # the kernel returns for i >= cnt
    hip.launch(fn, (grid,), (block,), [ctypes.c_void_p(dn.ptr), ctypes.c_void_p(dd.ptr),
                                       ctypes.c_void_p(do.ptr), ctypes.c_int(n)])
    hip.synchronize()
    out = do.download(np.empty((n, 6)))
    for a in (dn, dd, do):
        a.close()
    mod.unload()
    return out


def _same_bits(a, b):
    """Bitwise equal, any NaN matching any NaN."""
    nan = np.isnan(a) & np.isnan(b)
    return nan | (a.view(np.uint64) == b.view(np.uint64))


def _random_normals(rng, n, emin, emax):
    mant = rng.integers(0, 1 << 52, n, dtype=np.uint64)
    exp = (rng.integers(emin, emax + 1, n) + 1023).astype(np.uint64)
    sign = rng.integers(0, 2, n, dtype=np.uint64)
    return ((sign << np.uint64(63)) | (exp << np.uint64(52)) | mant).view(np.float64)


def test_quotients_match_ieee_division_over_the_normal_range(device):
    rng = np.random.default_rng(20261202)
    n = 1 << 20
    num = _random_normals(rng, n, -380, 380)
    den = _random_normals(rng, n, -380, 380)
    out = _run(num, den)
    assert np.array_equal(out[:, 0], num / den)  # the GPU's division is IEEE
    assert _same_bits(out[:, 1], out[:, 0]).all()
    assert _same_bits(out[:, 3], out[:, 2]).all()


def test_quotients_match_on_euler_like_values(device):
    """Magnitudes of densities, momenta, energies and of round-off-sized
    differences of neighbouring pressures (the limiter's ratios)."""
    rng = np.random.default_rng(7)
    n = 1 << 19
    big = rng.uniform(1e-3, 10.0, n) * rng.choice([-1.0, 1.0], n)
    p = rng.uniform(0.1, 2.0, n)
    tiny = (p + rng.integers(-4, 5, n) * np.spacing(p)) - p   # 0 or a few ulps
    num = np.concatenate([big, tiny, big, tiny])
    den = np.concatenate([rng.uniform(1e-3, 10.0, n), rng.uniform(0.1, 2, n), tiny, tiny])
    with np.errstate(divide="ignore", invalid="ignore"):
        want = num / den
    out = _run(num, den)
    assert _same_bits(out[:, 0], want).all()
    assert _same_bits(out[:, 1], out[:, 0]).all()
    assert _same_bits(out[:, 3], out[:, 2]).all()


def test_special_operands_follow_ieee_division(device):
    vals = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, -1.0, 3.5, -2.25e-300, 7e300])
    num, den = [a.ravel() for a in np.meshgrid(vals, vals, indexing="ij")]
    with np.errstate(divide="ignore", invalid="ignore", over="ignore", under="ignore"):
        want = num / den
    out = _run(num, den)

    def extreme(x):   # finite and outside the range the claim covers
        return np.isfinite(x) & ((np.abs(x) > 1e299) | ((np.abs(x) < 1e-299) & (x != 0)))

    # normal and special operands: bitwise IEEE (the sign of zero included)
    ok = ~extreme(num) & ~extreme(den)
    assert ok.sum() == 64
    assert _same_bits(out[ok, 0], want[ok]).all()
    assert _same_bits(out[ok, 1], want[ok]).all()


def _run_sqrt(x):
    import ctypes
    from pysweep_amd import hip
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = len(x)
    mod = _module()
    fn = mod.get_function("ps_sqrt_check")
    dx, do = hip.DeviceArray(8 * n), hip.DeviceArray(16 * n)
    dx.upload(x)
    block = 256
    grid = (n + block - 1) // block  # the kernel returns for i >= cnt
    hip.launch(fn, (grid,), (block,), [ctypes.c_void_p(dx.ptr), ctypes.c_void_p(do.ptr), ctypes.c_int(n)])
    hip.synchronize()
    out = do.download(np.empty((n, 2)))
    dx.close()
    do.close()
    mod.unload()
    return out


def test_square_root_matches_ieee(device):
    rng = np.random.default_rng(11)
    n = 1 << 20
    x = np.abs(_random_normals(rng, n, -760, 1000))
    x[: n // 4] = rng.uniform(1e-3, 20.0, n // 4)       # densities, sound speeds
    x[n // 4: n // 4 + 16] = [0.0, -0.0, np.inf, -np.inf, np.nan, -1.0, -4.0, 1.0, 4.0, 2.0,
                              0.5, 1e300, 1e-300, 2.0 ** -767, -(2.0 ** -767), 1.5]
    out = _run_sqrt(x)
    with np.errstate(invalid="ignore"):
        want = np.sqrt(x)
    covered = ~(np.isfinite(x) & (x > 0) & (x < 2.0 ** -767))
    assert covered.sum() == n - 1  # only 1e-300 lies below the range
    assert _same_bits(out[covered, 1], out[covered, 0]).all()
    assert _same_bits(out[:, 0], want).all()  # and the GPU's sqrt is IEEE


def test_quotients_without_fixup_inside_the_guarded_range(device):
    """eu_quo_pos (no special-case fix-up), used for the positive normal
    denominators of the step, is bitwise the IEEE quotient for operands of
    the magnitudes the division-range guard admits: numerators 0 or of
    magnitude in [2^-520, 2^250], denominators in [2^-130, 2^130]."""
    rng = np.random.default_rng(99)
    n = 1 << 20
    num = _random_normals(rng, n, -520, 250)
    num[:1024] = 0.0
    den = np.abs(_random_normals(rng, n, -130, 130))
    out = _run(num, den)
    assert _same_bits(out[:, 4], out[:, 0]).all()
    assert _same_bits(out[:, 5], out[:, 2]).all()
