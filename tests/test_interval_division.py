"""Dividing by a power-of-two spacing as a multiply (pow2_reciprocal and div_by in interval.h, used by the edge hull
of the certificate): for h = 2^k the IEEE quotient x / h and the product x * 2^-k round the same real number, so they
are the same double for every x, subnormals, huge values, signed zeros and infinities included. Checked on NumPy
binary64 arithmetic over random bit patterns; spacings that are not powers of two keep the division."""
import numpy as np


def pow2_reciprocal(h):
    b = np.float64(h).view(np.uint64)
    e = int((b >> np.uint64(52)) & np.uint64(0x7ff))
    if int(b >> np.uint64(63)) or int(b & np.uint64((1 << 52) - 1)) or e < 2 or e > 2044:
        return 0.
    return float(np.uint64((2046 - e) << 52).view(np.float64))


def test_power_of_two_division_is_the_product():
    rng = np.random.default_rng(3)
    x = rng.integers(0, 1 << 64, size=400_000, dtype=np.uint64, endpoint=False).view(np.float64)
    x = np.concatenate((x[np.isfinite(x)], [0., -0., np.inf, -np.inf, 5e-324, -5e-324, 1.7976931348623157e308]))
    for h in (1., 2., .5, .25, 8., 2. ** -30, 2. ** 40, 2. ** -1000, 2. ** 1000):
        inv = pow2_reciprocal(h)
        assert inv == 1. / h
        with np.errstate(over="ignore", under="ignore"):
            q, p = x / h, x * inv
        assert np.array_equal(q.view(np.uint64), p.view(np.uint64)), h
    for h in (3., 1.3, .8, 0., -1., np.inf, np.nan, 2. ** -1060, 2. ** 1023):
        assert pow2_reciprocal(h) == 0.
