"""The site-index division of the kernels (lceq_device.hpp div_small): the
single-precision estimate s * (1/d) is within one of floor(s / d) for
s < 2^31 and quotients below 2^22, and one correction step either way makes
it exact.  Checked here in IEEE single precision (numpy float32: round to
nearest on conversion and multiply, truncation on the conversion back, as the
gfx950 instructions do) over every divisor up to 4096, 3000 random divisors
up to 2^25, and quotients at and around every multiple."""
import numpy as np


def div_small(s, d):
    rd = np.float32(1.0) / np.float32(d)
    q = (s.astype(np.float32) * rd).astype(np.int64)
    r = s - q * d
    return q + (r >= d) - (r < 0)


def test_div_small_is_exact_over_the_lattice_ranges():
    rng = np.random.default_rng(0)
    divisors = list(range(1, 4097)) + [int(x) for x in rng.integers(4097, 2 ** 25, 3000)]
    for d in divisors:
        qmax = min(2 ** 22, (2 ** 31 - 1) // d)
        q = np.unique(np.concatenate([np.arange(0, min(qmax, 1024)),
                                      rng.integers(0, qmax, 500), [qmax - 1]]))
        s = np.concatenate([q * d, q * d + d - 1, np.maximum(q * d - 1, 0),
                            q * d + rng.integers(0, d, q.size)])
        s = s[s < 2 ** 31]
        assert np.array_equal(div_small(s, d), s // d), d
