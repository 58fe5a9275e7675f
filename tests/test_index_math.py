"""The device index arithmetic's FP32-reciprocal quotient (fast_div / quot in
kronmult_amd/csrc/kron_device.hpp): int((float(x) + 0.5f) * rcp(d)) equals
x // d for 0 <= x < 2^21, provided rcp(d) is within one ulp of 1/d (the
accuracy of v_rcp_f32).  Checked in float32 arithmetic (round-to-nearest
product, truncating conversion, as the kernels do it) with the reciprocal
correctly rounded and one ulp below and above, at the dividends where an
error would show first: k*d - 1 and k*d, the values closest to a quotient
boundary from either side, for every k with k*d < 2^21."""
import numpy as np
import pytest

LIMIT = 1 << 21


def _quot(x, r):
    xf = x.astype(np.float32) + np.float32(0.5)
    return (xf * r).astype(np.float32).astype(np.int64)


def _reciprocals(d):
    r = np.float32(1.0) / np.float32(d)
    return [np.nextafter(r, np.float32(0)), r, np.nextafter(r, np.float32(np.inf))]


def _boundary_dividends(d):
    k = np.arange(1, (LIMIT - 1) // d + 1, dtype=np.int64)
    x = np.concatenate([k * d - 1, k * d, np.array([0, LIMIT - 1], dtype=np.int64)])
    return x[(x >= 0) & (x < LIMIT)]


@pytest.mark.parametrize("dlo,dhi", [(1, 64), (64, 512), (512, 2048), (2048, 4097)])
def test_fp32_reciprocal_quotient_exact(dlo, dhi):
    for d in range(dlo, dhi):
        x = _boundary_dividends(d)
        want = x // d
        for r in _reciprocals(d):
            got = _quot(x, r)
            bad = np.nonzero(got != want)[0]
            assert bad.size == 0, (d, int(x[bad[0]]), float(r))


def test_fp32_reciprocal_quotient_random():
    rng = np.random.default_rng(2024)
    d = rng.integers(1, 1 << 16, size=200000)
    x = rng.integers(0, LIMIT, size=200000)
    want = x // d
    r = (np.float32(1.0) / d.astype(np.float32)).astype(np.float32)
    for rr in (np.nextafter(r, np.float32(0)), r, np.nextafter(r, np.float32(np.inf))):
        assert np.array_equal(_quot(x, rr), want)


def test_fp32_reciprocal_quotient_fails_past_the_bound():
    """The bound is real: far past 2^21 the same formula goes wrong, so the
    checks above are sensitive to it."""
    d = 3
    k = np.arange((1 << 25) // d, (1 << 25) // d + 4096, dtype=np.int64)
    x = np.concatenate([k * d - 1, k * d])
    r = np.nextafter(np.float32(1.0) / np.float32(d), np.float32(np.inf))
    assert np.any(_quot(x, r) != x // d)
