"""Field guard (lceq_check_fields): the largest departure of a link from
SU(2) and the number of non-finite link and Higgs values, over the whole
lattice; it reports, and returns LCEQ_ERR_FIELD on, planted faults."""
import numpy as np
import pytest

from helpers import rand_conf

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(8, 8, 8), (6, 4, 10), (64, 64, 64)])
def test_guard_reports_planted_faults(shape):
    import lceqcd_amd
    print("field guard", shape, flush=True)
    L = lceqcd_amd.Lattice(*shape)
    L.set_couplings(9.0, 0.175, -0.45, 0.025)
    L.hotstart(3, 0.5)
    L.rng_philox(3, 0)
    L.sweep(0, 2, n_or=2, delta=0.4, measflag=0)
    dev, bad = L.check_fields(1e-12)
    assert bad == 0 and dev < 1e-13
    u, a = L.get_links(), L.get_higgs()
    # one link scaled off SU(2): |U|^2 = 1.001^2
    u2 = u.copy()
    u2[2, 1, 5] *= 1.001
    L.put_links(u2)
    dev, bad = L.check_fields()
    assert bad == 0 and abs(dev - (1.001 ** 2 - 1.0)) < 1e-12
    with pytest.raises(lceqcd_amd.LceqError, match="off SU"):
        L.check_fields(1e-6)
    L.renorm()
    assert L.check_fields(1e-12)[1] == 0
    # a NaN Higgs component and an infinite link component
    a2 = a.copy()
    a2[0, 7, 1] = np.nan
    u3 = u.copy()
    u3[0, 0, 3, 2] = np.inf
    L.put_links(u3)
    L.put_higgs(a2)
    with pytest.raises(lceqcd_amd.LceqError, match="2 non-finite"):
        L.check_fields()
    L.close()
