"""RANLUX (f/luxury.f) in liblceq.so against the reference's own known
answers (f/luxtst.f:88-146), bit patterns of the reference's luxury.f built
single precision (tests/golden/ranlux_ref.npz, oracle/_ref/refranlux when it
is built), the seeding of f/measinit.f:214-216, and the subtract-with-borrow
recurrence itself.  Host code only: runs without a GPU."""
import json
import os
import subprocess

import numpy as np
import pytest

from helpers import golden

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REFRANLUX = os.path.join(ROOT, "oracle", "_ref", "refranlux")


@pytest.fixture(scope="module")
def lux():
    from lceqcd_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from lceqcd_amd import build
        build.build()
    import lceqcd_amd
    return lceqcd_amd


def bits(x):
    return np.asarray(x, np.float32).view(np.int32)


def test_luxtst_known_answers(lux):
    """f/luxtst.f's program, replayed; its listed output is F12.8 (F10.6 for
    the skip test), so agreement is to half a unit of the last printed digit."""
    kat = json.load(open(os.path.join(HERE, "golden", "ranlux_luxtst.json")))
    seq = iter(kat["sequences"])
    R = lux.Ranlux()

    def check(v, tol=5e-9):
        want = np.array(next(seq)["values"])
        assert np.abs(v[:5] - want).max() <= tol

    check(R.ranlux(100))
    check(R.ranlux(100))
    R.rluxgo(0, 0, 0, 0)
    check(R.ranlux(100))
    check(R.ranlux(100))
    R.rluxgo(389, 1, 0, 0)
    check(R.ranlux(100))
    check(R.ranlux(100))
    R.rluxgo(75, 0, 0, 0)
    check(R.ranlux(100))
    check(R.ranlux(100))
    restart = R.rluxut()
    assert restart.tolist() == kat["restart"]
    check(R.ranlux(100))
    check(R.ranlux(100))
    R.rluxin(restart)
    check(R.ranlux(100))
    check(R.ranlux(100))
    # restarting by skipping (the block luxtst keeps commented out, with its
    # listed output)
    R.rluxgo(4, 7674985, 0, 0)
    assert list(R.rluxat()) == kat["rluxat"][0]
    for _ in range(10):
        R.ranlux(1000)
    at = R.rluxat()
    assert list(at) == kat["rluxat"][1]
    v = R.ranlux(200)
    assert abs(v[0] - kat["skip_pairs"][0][0]) <= 5e-7
    assert abs(v[199] - kat["skip_pairs"][0][1]) <= 5e-7
    R.rluxgo(*at)
    w = R.ranlux(200)
    assert (bits(w) == bits(v)).all()


def test_reference_build_bit_patterns(lux):
    z = golden("ranlux_ref")
    for c, (luxury, seed, k1) in enumerate(z["cases"]):
        R = lux.Ranlux()
        R.rluxgo(int(luxury), int(seed), int(k1), 0)
        got = np.concatenate([R.ranlux(int(n)) for n in str(z["case_lengths"][c]).split(",")])
        assert (bits(got) == z["case%d_bits" % c]).all(), c
        assert (R.rluxut() == z["case%d_rluxut" % c]).all(), c
        assert list(R.rluxat()) == z["case%d_rluxat" % c].tolist(), c


def test_measinit_seeds(lux):
    z = golden("ranlux_ref")
    for rank, want in zip(z["measinit_ranks"], z["measinit_seeds"]):
        assert (lux.measinit_seeds(int(rank)) == want).all(), rank


@pytest.mark.skipif(not os.path.exists(REFRANLUX), reason="oracle/_ref/refranlux not built")
def test_against_reference_build_random_calls(lux):
    """Random luxuries, seeds, call lengths and save/restore points, the
    reference's luxury.f and the library driven by the same script."""
    rng = np.random.default_rng(7)
    for trial in range(6):
        luxury = int(rng.choice([0, 1, 2, 3, 4, 24, 31, 100, 389, 1999]))
        seed = int(rng.integers(1, 2**31 - 1))
        script, expect = "1\n%d %d 0 0\n" % (luxury, seed), []
        R = lux.Ranlux()
        R.rluxgo(luxury, seed)
        for _ in range(5):
            op = rng.integers(0, 3)
            if op == 0:
                n = int(rng.integers(0, 700))
                script += "2\n%d\n" % n
                expect.append(bits(R.ranlux(n)))
            elif op == 1:
                v = R.rluxut()
                script += "3\n"
                expect.append(v)
                script += "4\n%s\n" % " ".join(map(str, v))
                R.rluxin(v)
            else:
                script += "5\n"
                expect.append(np.array(R.rluxat()))
        out = subprocess.run([REFRANLUX], input=script, capture_output=True, text=True,
                             check=True).stdout
        ref = np.array([int(l[1:]) for l in out.splitlines() if l.startswith("D")])
        got = np.concatenate([np.asarray(e, np.int64) for e in expect]) if expect else ref
        assert got.shape == ref.shape and (got == ref).all(), (trial, luxury, seed)


def test_subtract_with_borrow_recurrence(lux):
    """At luxury 0 every generated number is delivered: the raw sequence
    obeys x_n = x_{n-10} - x_{n-24} - c_{n-1} (mod 1) with c_n = 2^-24 when
    that difference was negative (Marsaglia-Zaman, lags 24 and 10).  Numbers
    below 2^-12 are padded on output, so the recurrence is checked where all
    three terms are unpadded."""
    R = lux.Ranlux()
    R.rluxgo(0, 271828)
    x = R.ranlux(20000).astype(np.float64)
    t = 2.0 ** -24
    ok = 0
    carry = None
    for n in range(24, len(x)):
        if min(x[n], x[n - 10], x[n - 24]) < 2.0 ** -12:
            carry = None
            continue
        d = x[n - 10] - x[n - 24]
        if carry is None:  # infer the carry from this step
            carry = 0.0 if abs((d % 1.0) - x[n]) < 0.5 * t else t
        y = d - carry
        c = t if y < 0 else 0.0
        assert abs((y % 1.0) - x[n]) < 0.5 * t, n
        carry = c
        ok += 1
    assert ok > 19000


def test_luxury_skips_whole_blocks(lux):
    """At level p, of every p numbers generated 24 are delivered: 24 numbers
    at luxury 4 then advance the level-0 stream of the same seed by 389."""
    R4, R0 = lux.Ranlux(), lux.Ranlux()
    R4.rluxgo(4, 1234)
    R0.rluxgo(0, 1234)
    a = R4.ranlux(48)
    b = R0.ranlux(389 + 24)
    assert (bits(a[:24]) == bits(b[:24])).all()
    assert (bits(a[24:]) == bits(b[389:389 + 24])).all()


def test_error_fallbacks(lux):
    R, ref = lux.Ranlux(), lux.Ranlux()
    with pytest.raises(lux.LceqError, match="luxury"):
        R.rluxgo(10, 5)  # illegal: level 4 is used
    assert R.rluxat()[0] == 4
    ref.rluxgo(4, 5)
    assert (bits(R.ranlux(100)) == bits(ref.ranlux(100))).all()
    with pytest.raises(lux.LceqError, match="seed"):
        R.rluxgo(2, -3)  # negative seed: the default seed is used
    ref.rluxgo(2, 0)
    assert (bits(R.ranlux(100)) == bits(ref.ranlux(100))).all()
    # f/luxtst.f:141-146: a restart point that cannot occur at level 4
    with pytest.raises(lux.LceqError, match="cannot occur"):
        R.rluxgo(4, 11111, 31, 0)
    with pytest.raises(ValueError):
        R.rluxin(np.zeros(24, np.int32))


def test_default_initialisation_matches_rluxgo_default(lux):
    a, b = lux.Ranlux(), lux.Ranlux()
    b.rluxgo(3, 314159265)
    assert (bits(a.ranlux(500)) == bits(b.ranlux(500))).all()
    assert a.rluxat() == (3, 314159265, 500 + 199 * (500 // 24), 0)


def test_measinit_key_of_the_reference_seeds():
    """The device streams' key from a rank's measinit seeds: iseed(1) +
    2^32 iseed(2) (both below 10^6, so distinct seed pairs give distinct keys)."""
    from helpers import golden
    from lceqcd_amd import measinit_key
    z = golden("ranlux_ref")
    keys = set()
    for rank, s in zip(z["measinit_ranks"], z["measinit_seeds"]):
        k = measinit_key(s)
        assert k == int(s[0]) + (int(s[1]) << 32)
        assert 0 <= s[0] < 10 ** 6 and 0 <= s[1] < 10 ** 6
        keys.add(k)
    assert len(keys) == len(z["measinit_ranks"])
