"""The benchmark configurations as runnable flows (lceqcd_amd.runs), at
small sizes, against the CPU restatement where a chain can be compared:

- the plain-method flow (cold start, RANLUX-seeded streams, per-sweep
  plaquette, Tr A^2 and Polyakov loop) follows the restatement's run of the
  same schedule from unitconf with the same key;
- the gauge-fixed flow saves configurations in the reference's file format,
  reads them back, fixes them to da < 1e-12 in the restatement's iteration
  count, and measures the correlators;
- the production flow (gauge fixing plus correlators every meas_every
  sweeps) runs, converges at every measurement and gives finite results."""
import os

import numpy as np
import pytest

from helpers import assert_close

pytestmark = pytest.mark.gpu


def test_plain_flow_follows_the_oracle(oracle_lib):
    from lceqcd_amd import runs
    from lceqcd_amd import workloads as W
    print("plain flow 16^3, 3 sweeps", flush=True)
    rec = runs.run_plain(sweeps=3)
    O = oracle_lib.OracleLattice(16, 16, 16)
    O.set_couplings(*W.coupling(W.C1["beta_a"], W.C1["beta_2"]))
    O.set_renorm_on_write(1)
    O.unitconf()
    O.philox(rec["philox_key"], 0)
    assert rec["ranlux_seeds"][0] + (rec["ranlux_seeds"][1] << 32) == rec["philox_key"]
    d = W.C1["delta"]
    for it in range(3):
        O.agauge(0)
        for k in range(4):
            O.link_or(1 if k == 3 else 0)
        d, _ = O.ahiggs1(d, 1)
        ob = O.observables()
        psq, _ = O.polyakov()
        tol = (1e-12, 1e-11, 1e-11)[it]
        assert_close(rec["series"]["plaquette"][it], ob["plaq0"], tol, "plaquette %d" % it)
        assert_close(rec["series"]["tr_a2"][it], ob["a02"] / 2.0, tol, "Tr A^2 %d" % it)
        assert_close(rec["series"]["polyakov"][it], np.sqrt(psq), tol, "polyakov %d" % it)
        assert rec["series"]["delta"][it] == d


def test_gauge_fixed_flow(tmp_path, oracle_lib):
    from lceqcd_amd import runs
    from lceqcd_amd.conf import conf_read
    print("gauge-fixed flow 8^3, 2 configurations", flush=True)
    rec = runs.run_gauge_fixed(configs=2, spacing=5, therm=20, shape=(8, 8, 8),
                               workdir=str(tmp_path))
    assert len(rec["measurements"]) == 2
    for m in rec["measurements"]:
        assert m["da"] <= 1e-12 and m["iterations"] > 0
        assert np.isfinite(m["corr_a"]).all() and np.isfinite(m["scalar_corr"]).all()
    # the first saved configuration, fixed by the restatement
    u, a = conf_read((8, 8, 8), tmp_path / "u_conf_000", tmp_path / "a_conf_000")
    O = oracle_lib.OracleLattice(8, 8, 8)
    O.put_links(u)
    O.put_higgs(a)
    O.set_gauge_fixing(1.7, 1.0, 1e-12, 1000000)
    da, it = O.landau_gauge()
    m = rec["measurements"][0]
    print("iterations GPU %d oracle %d" % (m["iterations"], it), flush=True)
    assert it == m["iterations"]
    ca, cu, c3 = O.correlator()
    assert_close(m["corr_a"], ca, 1e-10, "C_a")
    assert_close(m["corr_u"], cu, 1e-10, "C_u")
    ms, cs = O.scalarcorr()
    assert_close(m["scalar_corr"], cs, 1e-12, "scalar correlator")


def test_production_flow_runs():
    from lceqcd_amd import runs
    print("production flow 8x8x16, 100 sweeps", flush=True)
    rec = runs.run_production((8, 8, 16), 0.33, sweeps=100, meas_every=50)
    assert rec["measurements"] == 2 and rec["gauge_fixing_iterations_mean"] > 0
    assert np.isfinite(rec["plaquette"]).all() and np.isfinite(rec["tr_a2"]).all()
    assert np.isfinite(rec["corr_a_mean"]).all()
    assert 0.5 < rec["plaquette"][0] < 1.0
