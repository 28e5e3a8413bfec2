"""25-sweep trajectories of the reference's schedule (f/main.f:37-41:
ahiggs2, ahiggs3, agauge, renorm) with the reference's own uniform stream,
run live through the reference's Fortran (oracle/_ref/refharness64, command
11 captures the stream and checks its consumption), replayed by the CPU
restatement and by the GPU in fed mode.  The golden fixtures cover one or
two sweeps; this checks that the restatement and the GPU follow the
reference's Markov chain sweep after sweep: every acceptance and observable,
and the final fields.  Skipped where the reference harness was not built."""
import numpy as np
import pytest

from helpers import rand_conf

NREP = 25
COUP = dict(beta=9.0, beta_a=0.2, beta_2=0.6, beta_4=0.05)


def _reference(shape, imode):
    from oracle import refrun
    if not refrun.available(64):
        pytest.skip("oracle/_ref not built")
    u, a = rand_conf(shape, 31, 0.6)
    recs = refrun.run("sweep", shape, u, a, nrep=NREP, seed=4321, imode=imode, **COUP)
    obs = np.array([r[0] for r in recs[:7 * NREP]]).reshape(NREP, 7)
    u_out, a_out, chk, uni = recs[7 * NREP:7 * NREP + 4]
    assert chk[0] == 1.0  # the routines consumed exactly the captured stream
    return u, a, obs, u_out, a_out, uni


def _replay(L, u, a, uni, imode):
    L.set_couplings(COUP["beta"], COUP["beta_a"], COUP["beta_2"], COUP["beta_4"])
    L.set_imode(imode)
    L.put_links(u)
    L.put_higgs(a)
    L.rng_feed(uni)
    rows = []
    for _ in range(NREP):
        acc2, acc3, acc1 = L.ahiggs2(1), L.ahiggs3(1), L.agauge(1)
        L.renorm()
        ob = L.observables()
        rows.append([acc2, acc3, acc1, ob["plaq0"], ob["a0kin"], ob["a02"], ob["a04"]])
    return np.array(rows), L.get_links().ravel(), L.get_higgs().ravel(), L.rng_state()[0]


@pytest.mark.parametrize("shape,imode", [((4, 4, 6), 0), ((6, 4, 4), 1)])
def test_restatement_follows_the_reference_chain(oracle_lib, shape, imode):
    u, a, obs, u_out, a_out, uni = _reference(shape, imode)
    rows, ug, ag, used = _replay(oracle_lib.OracleLattice(*shape), u, a, uni, imode)
    assert used == len(uni)
    np.testing.assert_allclose(rows, obs, rtol=0, atol=1e-13)
    np.testing.assert_allclose(ug, u_out, rtol=0, atol=1e-13)
    np.testing.assert_allclose(ag, a_out, rtol=0, atol=1e-13)


@pytest.mark.gpu
@pytest.mark.parametrize("shape,imode", [((4, 4, 6), 0), ((6, 4, 4), 1)])
def test_gpu_follows_the_reference_chain(shape, imode):
    import lceqcd_amd
    print("25-sweep fed trajectory", shape, imode, flush=True)
    u, a, obs, u_out, a_out, uni = _reference(shape, imode)
    rows, ug, ag, used = _replay(lceqcd_amd.Lattice(*shape), u, a, uni, imode)
    print("max deviation: observables %.2e links %.2e higgs %.2e"
          % (np.abs(rows - obs).max(), np.abs(ug - u_out).max(), np.abs(ag - a_out).max()),
          flush=True)
    assert used == len(uni)
    np.testing.assert_allclose(rows, obs, rtol=0, atol=1e-11)
    np.testing.assert_allclose(ug, u_out, rtol=0, atol=1e-11)
    np.testing.assert_allclose(ag, a_out, rtol=0, atol=1e-11)
