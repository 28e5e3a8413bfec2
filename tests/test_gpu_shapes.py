"""Launch shapes of the single-lattice link and Higgs updates (threads per
block, XCD-aware block order): the fields are the same bit for bit for every
shape, and the autotuner leaves the chain where it was."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

COUP = (9.0, 0.175, -0.45, 0.025)


def _run(shape_, lat=(10, 6, 12), tune=False):
    import lceqcd_amd
    L = lceqcd_amd.Lattice(*lat)
    L.set_couplings(*COUP)
    L.hotstart(4, 0.5)
    L.rng_philox(4, 0)
    d = L.sweep(0, 2, n_or=2, delta=0.4, measflag=0)
    if shape_:
        L.set_shape(*shape_)
    ms = L.autotune(n_or=2, trials=2) if tune else None
    d = L.sweep(0, 3, n_or=2, delta=d, measflag=1)
    out = L.get_links(), L.get_higgs(), L.history(3), d, L.get_shape(), ms
    L.close()
    return out


@pytest.mark.parametrize("shape", [(64, False), (64, True), (128, True), (512, False)])
def test_shapes_give_the_same_fields(shape):
    print("single-lattice shape", shape, flush=True)
    u0, a0, h0, d0, _, _ = _run(None)
    u1, a1, h1, d1, got, _ = _run(shape)
    assert got == shape
    assert np.array_equal(u0, u1) and np.array_equal(a0, a1)
    assert d0 == d1
    np.testing.assert_allclose(h1, h0, rtol=1e-13, atol=1e-15)


def test_autotune_restores_the_chain():
    u0, a0, h0, d0, _, _ = _run(None)
    u1, a1, h1, d1, got, ms = _run(None, tune=True)
    print("autotune", {k: round(v, 4) for k, v in ms.items()}, "->", got, flush=True)
    assert ms[got] == min(ms.values())
    assert np.array_equal(u0, u1) and np.array_equal(a0, a1)
    np.testing.assert_allclose(h1, h0, rtol=1e-13, atol=1e-15)


def test_copy_bandwidth_probe():
    """The plain-copy probe that the benchmark reports beside the 8 TB/s peak
    returns a finite time and a rate below that peak for a 64 MiB copy."""
    import ctypes as C
    from lceqcd_amd import _lib
    us = C.c_double()
    nbytes = 64 << 20
    print("copy probe", nbytes, flush=True)
    _lib.check(_lib.load().lceq_copy_bandwidth(nbytes, 10, C.byref(us)))
    rate = 2 * nbytes / (us.value * 1e-6) / 1e9
    print("copy probe GB/s", rate, flush=True)
    assert np.isfinite(us.value) and 0.0 < rate < 20000.0


def _run_order(order, lat, tune=False):
    import lceqcd_amd
    L = lceqcd_amd.Lattice(*lat)
    L.set_couplings(*COUP)
    L.hotstart(6, 0.5)
    L.rng_philox(6, 0)
    d = L.sweep(0, 2, n_or=2, delta=0.4, measflag=0)
    L.set_site_order(order)
    ms = L.autotune_site_order(n_or=2, trials=2) if tune else None
    d = L.sweep(0, 3, n_or=2, delta=d, measflag=1)
    h = L.history(3)
    L.sweep(1, 2, measflag=1)  # the reference schedule: ahiggs2, ahiggs3, agauge, renorm
    h = np.concatenate([h, L.history(2)])
    out = L.get_links(), L.get_higgs(), h, d, L.get_site_order(), ms
    L.close()
    return out


@pytest.mark.parametrize("lat,order", [((16, 16, 16), 4), ((16, 16, 16), 16), ((10, 8, 12), 2),
                                       ((8, 32, 6), 8)])
def test_site_orders_give_the_same_fields(lat, order):
    """Strips of y rows advancing through z (lceq_set_site_order) visit every
    site of a parity once: the same fields bit for bit as the index order,
    through production sweeps (KP, overrelaxation, ahiggs1) and reference
    sweeps (ahiggs2, ahiggs3, KP, renorm)."""
    print("site order", lat, order, flush=True)
    u0, a0, h0, d0, _, _ = _run_order(0, lat)
    u1, a1, h1, d1, got, _ = _run_order(order, lat)
    assert got == order
    assert np.array_equal(u0, u1) and np.array_equal(a0, a1)
    assert d0 == d1
    np.testing.assert_allclose(h1, h0, rtol=1e-13, atol=1e-15)


def test_site_order_tuner_restores_the_chain():
    import lceqcd_amd
    lat = (16, 16, 16)
    u0, a0, h0, d0, _, _ = _run_order(0, lat)
    u1, a1, h1, d1, got, ms = _run_order(0, lat, tune=True)
    print("site order tuner", ms, "->", got, flush=True)
    assert set(ms) == {0, 4, 8, 16} and ms[got] == min(ms.values())
    assert np.array_equal(u0, u1) and np.array_equal(a0, a1)
    L = lceqcd_amd.Lattice(16, 16, 16)
    with pytest.raises(lceqcd_amd.LceqError, match="strip rows"):
        L.set_site_order(3)
    with pytest.raises(lceqcd_amd.LceqError, match="strip rows"):
        L.set_site_order(32)
    L.close()
    # ly * lz = 2^22 rows per parity: the row division would no longer be exact
    L = lceqcd_amd.Lattice(2, 2048, 2048)
    with pytest.raises(lceqcd_amd.LceqError, match="ly \\* lz < 2\\^22"):
        L.set_site_order(2)
    assert L.get_site_order() == 0
    L.close()


def test_stream_bandwidth_probe():
    """The 30-stream read probe that the benchmark reports beside the copy
    ceiling returns a finite time and a rate below the 8 TB/s peak for 30
    streams of 2^20 16-byte values."""
    import ctypes as C
    from lceqcd_amd import _lib
    us = C.c_double()
    n = 1 << 20
    print("stream probe", n, flush=True)
    _lib.check(_lib.load().lceq_stream_bandwidth(30, n, 5, C.byref(us)))
    rate = 31 * n * 16 / (us.value * 1e-6) / 1e9
    print("stream probe GB/s", rate, flush=True)
    assert np.isfinite(us.value) and 0.0 < rate < 20000.0
