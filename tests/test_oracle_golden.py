"""The CPU restatement (oracle/) against golden vectors written by the
reference's own FP64 Fortran (tests/golden/make_golden.py).  These pin the
oracle before it is trusted as the checker of the HIP kernels."""
import numpy as np
import pytest

from helpers import golden
from parity_cases import all_cases, check_all


@pytest.fixture(scope="module")
def make(oracle_lib):
    return oracle_lib.OracleLattice


@pytest.mark.parametrize("cid,case", all_cases(), ids=[c[0] for c in all_cases()])
def test_oracle_matches_reference(make, cid, case):
    check_all(case(make))


def test_geometry_tables_match_adresses(make):
    """f/adresses.f neighbour tables: restatement and closed form == reference."""
    fx = golden("geom_6x4x8")
    L = make(*[int(x) for x in fx["shape"]])
    up, dn = L.site_tables()
    upc, dnc = L.site_tables(closed_form=True)
    assert np.array_equal(up, fx["siteup"]) and np.array_equal(dn, fx["sitedn"])
    assert np.array_equal(upc, fx["siteup"]) and np.array_equal(dnc, fx["sitedn"])


@pytest.mark.skipif(not __import__("oracle.refrun", fromlist=["x"]).available(64),
                    reason="oracle/_ref not built")
@pytest.mark.parametrize("prec", [64, 32])
def test_reference_timing_command_runs_the_sweep(prec):
    """The harness's timing command (the reference CPU baseline of the
    reference-schedule benchmark) runs the same f/main.f:37-41 sweeps as the
    captured-stream sweep command: the same seed gives the same plaquette."""
    from oracle import refrun
    from helpers import rand_conf
    shape = (4, 4, 6)
    u, a = rand_conf(shape, 14, 0.5)
    kw = dict(seed=777, beta=9.0, beta_a=0.2, beta_2=0.6, beta_4=0.05, imode=0, prec=prec)
    sec, plaq = refrun.time_sweeps(shape, u, a, 3, **kw)
    recs = refrun.run("sweep", shape, u, a, nrep=3, **kw)
    assert sec >= 0.0
    # command 11 writes 7 records per sweep: acceptances, then plaq0, ...
    assert plaq == recs[7 * 2 + 3][0]
