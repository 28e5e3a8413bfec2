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


def order_site(P, lx, ly, lz, sy):
    """lceq_device.hpp order_site in the kernels' integer arithmetic: the site
    at position P of the strip order with sy rows per strip (0: P itself)."""
    if sy == 0:
        return P
    h1, h2 = lx // 2, lx * ly // 2
    lsy = sy.bit_length() - 1
    rows = sy * lz
    r = div_small(P, h1)
    k = P - r * h1
    strip = div_small(r, rows)
    q = r - strip * rows
    z = q >> lsy
    y = (strip << lsy) + (q - (z << lsy))
    return k + h1 * y + h2 * z


def test_strip_site_order_is_a_permutation_with_local_z_neighbours():
    """Every strip order the library accepts maps the positions 0..N-1 of a
    launch one to one onto the sites of the parity (so every load and store
    stays in bounds and every site is updated once), keeps a row's x run
    contiguous, and puts a site's z neighbour sy rows later in the order."""
    for shape in ((2, 2, 2), (2, 8, 4), (4, 4, 4), (8, 4, 6), (64, 64, 64), (160, 160, 160),
                  (32, 32, 128), (24, 48, 8), (6, 16, 2)):
        lx, ly, lz = shape
        n = lx * ly * lz // 2
        P = np.arange(n, dtype=np.int64)
        for sy in (0, 2, 4, 8, 16):
            if sy and ly % sy:
                continue
            s = order_site(P, lx, ly, lz, sy)
            assert np.array_equal(np.sort(s), P), (shape, sy)
            if sy:
                h1, h2 = lx // 2, lx * ly // 2
                # one row of the order is a contiguous run of h1 sites
                assert np.all(np.diff(s.reshape(-1, h1), axis=1) == 1)
                # position P + sy * h1 is the z neighbour of P within a strip
                inner = P[: n - sy * h1]
                zn = order_site(inner + sy * h1, lx, ly, lz, sy)
                same_strip = (inner // h1) // (sy * lz) == ((inner + sy * h1) // h1) // (sy * lz)
                assert np.all(zn[same_strip] - s[inner][same_strip] == h2), (shape, sy)
