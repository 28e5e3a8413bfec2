"""25-sweep trajectories of the reference's schedule (f/main.f:37-41:
ahiggs2, ahiggs3, agauge, renorm) with the reference's own uniform stream,
replayed by the CPU restatement and by the GPU in fed mode.  The
trajectories were run through the reference's Fortran built here
(oracle/_ref/refharness64, command 11 captures the stream and checks its
consumption) and committed as tests/golden/trajectory25_*.npz by
tests/golden/make_trajectory_golden.py; where the harness is built, a CPU
test re-runs it and checks the fixture.  The one- and two-sweep golden
fixtures cover single routines; this checks that the restatement and the GPU
follow the reference's Markov chain sweep after sweep: every acceptance and
observable, and the final fields."""
import os
import sys

import numpy as np
import pytest

from helpers import golden

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
from make_trajectory_golden import CASES, name, reference_trajectory  # noqa: E402

IDS = ["%s-imode%d" % ("x".join(map(str, s)), i) for s, i in CASES]


def _replay(L, d):
    L.set_couplings(*d["couplings"])
    L.set_imode(int(d["imode"]))
    L.put_links(d["u"])
    L.put_higgs(d["a"])
    L.rng_feed(d["uniforms"])
    rows = []
    for _ in range(int(d["nrep"])):
        acc2, acc3, acc1 = L.ahiggs2(1), L.ahiggs3(1), L.agauge(1)
        L.renorm()
        ob = L.observables()
        rows.append([acc2, acc3, acc1, ob["plaq0"], ob["a0kin"], ob["a02"], ob["a04"]])
    return np.array(rows), L.get_links().ravel(), L.get_higgs().ravel(), L.rng_state()[0]


@pytest.mark.parametrize("shape,imode", CASES, ids=IDS)
def test_fixture_is_the_reference_chain(shape, imode):
    """The committed trajectory equals a live run of the reference build."""
    from oracle import refrun
    if not refrun.available(64):
        pytest.skip("oracle/_ref not built")
    d, live = golden(name(shape, imode)), reference_trajectory(shape, imode)
    for k in ("u", "a", "obs", "u_out", "a_out", "uniforms"):
        assert np.array_equal(np.asarray(d[k]), np.asarray(live[k])), k


@pytest.mark.parametrize("shape,imode", CASES, ids=IDS)
def test_restatement_follows_the_reference_chain(oracle_lib, shape, imode):
    d = golden(name(shape, imode))
    rows, ug, ag, used = _replay(oracle_lib.OracleLattice(*shape), d)
    assert used == len(d["uniforms"])
    np.testing.assert_allclose(rows, d["obs"], rtol=0, atol=1e-13)
    np.testing.assert_allclose(ug, d["u_out"].ravel(), rtol=0, atol=1e-13)
    np.testing.assert_allclose(ag, d["a_out"].ravel(), rtol=0, atol=1e-13)


@pytest.mark.gpu
@pytest.mark.parametrize("shape,imode", CASES, ids=IDS)
def test_gpu_follows_the_reference_chain(shape, imode):
    import lceqcd_amd
    print("25-sweep fed trajectory", shape, imode, flush=True)
    d = golden(name(shape, imode))
    rows, ug, ag, used = _replay(lceqcd_amd.Lattice(*shape), d)
    print("max deviation: observables %.2e links %.2e higgs %.2e"
          % (np.abs(rows - d["obs"]).max(), np.abs(ug - d["u_out"].ravel()).max(),
             np.abs(ag - d["a_out"].ravel()).max()), flush=True)
    assert used == len(d["uniforms"])
    np.testing.assert_allclose(rows, d["obs"], rtol=0, atol=1e-11)
    np.testing.assert_allclose(ug, d["u_out"].ravel(), rtol=0, atol=1e-11)
    np.testing.assert_allclose(ag, d["a_out"].ravel(), rtol=0, atol=1e-11)
