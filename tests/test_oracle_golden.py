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
