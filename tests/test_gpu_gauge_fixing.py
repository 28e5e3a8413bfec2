"""landau_gauge's device-side loop (f/landau_gauge.f:54-59) against the
reference's own loop written on the host: or_gauge_fixing + u_f_renorm
while da1 > limit_or.  The device loop runs its iterations in batches and
tests the loop condition inside the next iteration's first launch; the two
must stop after the same number of iterations with the same da1 and the same
fixed fields, bit for bit (the same kernels run on the same values).  The
iteration cap stops the loop after exactly max_iter iterations."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _thermalised(shape, seed):
    import lceqcd_amd
    L = lceqcd_amd.Lattice(*shape)
    L.set_couplings(9.0, 0.175, -0.45, 0.025)
    L.hotstart(seed, 0.5)
    L.rng_philox(seed, 0)
    L.sweep(lceqcd_amd.SWEEP_KP_OR_METRO, 20, n_or=4, delta=0.4, measflag=0)
    u, a = L.get_links(), L.get_higgs()
    L.close()
    return u, a


def _host_loop(shape, u, a, lam, alpha, limit, nmax):
    import lceqcd_amd
    H = lceqcd_amd.Lattice(*shape)
    H.set_gauge_fixing(lam, alpha, limit)
    H.put_links_fixed(u)
    H.put_higgs_fixed(a)
    da, n = 1e4, 0
    while da > limit and n < nmax:
        da = H.or_gauge_fixing()
        H.u_f_renorm()
        n += 1
    return H, da, n


@pytest.mark.parametrize("shape,alpha,limit", [((8, 8, 8), 1.0, 1e-8), ((6, 4, 8), 0.7, 1e-9),
                                               ((16, 16, 16), 1.0, 1e-7), ((4, 4, 4), 1.0, 1e-12),
                                               ((128, 128, 80), 1.0, 1e-4)])
def test_device_loop_equals_host_loop(shape, alpha, limit):
    """128 x 128 x 80 has 2560 blocks per launch: the loop condition is then
    tested by a one-block launch after every iteration instead of inside the
    next iteration's first launch."""
    import lceqcd_amd
    print("landau_gauge device vs host loop", shape, alpha, limit, flush=True)
    u, a = _thermalised(shape, 41)
    G = lceqcd_amd.Lattice(*shape)
    G.put_links(u)
    G.put_higgs(a)
    G.set_gauge_fixing(1.7, alpha, limit)
    da, it = G.landau_gauge()
    H, dah, nh = _host_loop(shape, u, a, 1.7, alpha, limit, 100000)
    print("iterations device %d host %d, da %.3e" % (it, nh, da), flush=True)
    assert it == nh and it > 8
    assert da == dah and da <= limit
    assert np.array_equal(G.get_links_fixed(), H.get_links_fixed())
    assert np.array_equal(G.get_higgs_fixed(), H.get_higgs_fixed())
    # the unfixed fields are untouched; a second call repeats the first
    assert np.array_equal(G.get_links(), u)
    assert G.landau_gauge() == (da, it)


@pytest.mark.parametrize("cap,shape", [(1, (8, 8, 8)), (8, (8, 8, 8)), (13, (8, 8, 8)),
                                       (40, (8, 8, 8)), (13, (128, 128, 80))])
def test_iteration_cap(cap, shape):
    import lceqcd_amd
    print("landau_gauge cap", cap, flush=True)
    u, a = _thermalised(shape, 43)
    G = lceqcd_amd.Lattice(*shape)
    G.put_links(u)
    G.put_higgs(a)
    G.set_gauge_fixing(1.7, 1.0, 1e-14, cap)
    with pytest.raises(lceqcd_amd.LceqError, match="max_iter"):
        G.landau_gauge()
    H, _, n = _host_loop(shape, u, a, 1.7, 1.0, 1e-14, cap)
    assert n == cap
    assert np.array_equal(G.get_links_fixed(), H.get_links_fixed())
