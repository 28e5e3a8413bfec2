"""Drop-in test of the Fortran boundary.  oracle/_ref/dropin_harness64 is the
reference harness (oracle/ref/refharness.f, which calls agauge, staples,
heat_kpi, a_force, ahiggs1-3, gauss, renorm, u_f_renorm, or_gauge_fixing and
landau_gauge by their Fortran names through the reference COMMON blocks)
linked against lceqcd_amd/fortran/lceq_dropin.f + liblceq.so instead of the
reference's routines.  Fed the same random_number stream, it must reproduce
the golden vectors the reference itself wrote, through the GPU kernels.
The binary is built in-tree by `make -C oracle/ref` where the reference
sources exist; the test is skipped where it was never built."""
import os

import numpy as np
import pytest

from helpers import golden, rand_conf, relerr

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "dropin_harness64")


def run(cmd, fx, **kw):
    from oracle import refrun
    shape = tuple(int(x) for x in fx["shape"]) if "shape" in fx else (8, 8, 8)
    n = shape[0] * shape[1] * shape[2] // 2
    u = fx["u"] if "u" in fx else np.tile([1.0, 0, 0, 0], (3, 2, n, 1))
    a = fx["a"] if "a" in fx else np.zeros((2, n, 3))
    if "couplings" in fx:
        c = fx["couplings"]
        kw.update(beta=c[0], beta_a=c[1], beta_2=c[2], beta_4=c[3])
    return refrun.run(cmd, shape, u, a, exe=EXE, **kw)


def close(got, exp, tol, what):
    e = relerr(got, exp)
    assert e <= tol, "%s: %.3e" % (what, e)


@pytest.fixture(autouse=True)
def need_exe():
    if not os.access(EXE, os.X_OK):
        pytest.skip("drop-in harness not built (make -C oracle/ref dropin)")


def test_dropin_agauge():
    for name in ("agauge_6x4x8_mixed", "agauge_4x4x4_inputfile", "agauge_8x8x8_production"):
        fx = golden(name)
        nrep = len(fx["accept"])
        r = run("agauge", fx, arg1=1, nrep=nrep)
        assert r[-2][0] == 1.0, name + ": uniform consumption differs"
        assert np.array_equal(r[-1], fx["uniforms"])
        for k in range(nrep):
            close(r[2 * k][0], fx["accept"][k], 1e-12, name + " accept")
            close(r[2 * k + 1][0], fx["plaq0"][k], 1e-12, name + " plaq0")
        close(r[2 * nrep], fx["u_out"], 1e-12, name + " u")


@pytest.mark.parametrize("name", ["ahiggs1_6x4x8_production", "ahiggs2_6x4x8_imode0",
                                  "ahiggs2_6x4x8_imode1", "ahiggs3_6x4x8_imode1",
                                  "ahiggs2_6x4x8_inputfile"])
def test_dropin_higgs(name):
    fx = golden(name)
    r = run(name.split("_")[0], fx, arg1=1, nrep=2, imode=int(fx["imode"]),
            delta=float(fx["delta"]))
    assert r[-2][0] == 1.0
    obs = np.array([[r[5 * k + j][0] for j in range(5)] for k in range(2)])
    close(obs, fx["obs"], 1e-12, name + " obs")
    close(r[10], fx["a_out"], 1e-12, name + " a")


def test_dropin_staples_force_renorm():
    fx = golden("staples_6x4x8")
    n = fx["v"].shape[2]
    for eo in (0, 1):
        for mu in (1, 2, 3):
            r = run("staples", fx, arg1=eo, arg2=mu)
            close(r[0].reshape(3, 2, n, 4)[mu - 1, eo], fx["v"][mu - 1, eo], 1e-13, "staple")
    fx = golden("a_force_6x4x8")
    for eo in (0, 1):
        r = run("a_force", fx, arg1=eo)
        close(r[0].reshape(n, 3), fx["force"][eo], 1e-13, "force")
    fx = golden("renorm_6x4x8")
    close(run("renorm", fx)[0], fx["u_out"], 1e-14, "renorm")


def test_dropin_gauge_fixing():
    fx = golden("or_gauge_fixing_6x4x8")
    r = run("or_gauge_fixing", fx, nrep=3, lam=float(fx["lam"]), alpha=float(fx["alpha"]))
    close([x[0] for x in r[:3]], fx["da"], 1e-11, "da")
    close(r[3], fx["u_f"], 1e-12, "u_f")
    close(r[4], fx["a_f"], 1e-12, "a_f")
    fx = golden("landau_gauge_4x4x4_iso")
    r = run("landau_gauge", fx, lam=float(fx["lam"]), alpha=float(fx["alpha"]),
            limit_or=float(fx["limit_or"]))
    assert abs(r[1][0] - fx["iter_or1"]) <= 1
    if r[1][0] == fx["iter_or1"]:
        close(r[2], fx["u_f"], 1e-10, "landau u_f")


def test_dropin_reference_sweeps():
    for name in ("sweep_4x4x4", "sweep_8x8x8"):
        fx = golden(name)
        nrep = fx["obs"].shape[0]
        r = run("sweep", fx, nrep=nrep)
        assert r[-2][0] == 1.0
        obs = np.array([[r[7 * k + j][0] for j in range(7)] for k in range(nrep)])
        close(obs, fx["obs"], 1e-11, name + " obs")
        close(r[7 * nrep], fx["u_out"], 1e-11, name + " u")
        close(r[7 * nrep + 1], fx["a_out"], 1e-11, name + " a")


def test_dropin_heat_kpi_gauss():
    fx = golden("heat_kpi_256")
    r = run("heat_kpi", {}, ex1=fx["b"], ex2=fx["a0_old"])
    close(r[0].reshape(-1, 4), fx["v"], 1e-13, "heat_kpi v")
    assert r[1][0] == fx["nacch"]
    fx = golden("gauss_256")
    r = run("gauss", {}, ex1=fx["r"])
    close(r[0], fx["x"], 1e-13, "gauss")


def test_dropin_schedule_timed_against_the_device_resident_path():
    """Command 16 of the harness times nrep sweeps of the reference schedule
    (f/main.f:37-41: ahiggs2, ahiggs3, agauge, renorm) inside the Fortran
    program, here through the drop-in layer on 24^3.  Every drop-in routine
    copies the field it changed back into COMMON, and in the default mode it
    also draws the reference's uniforms on the host and uploads them.  Two
    runs of 20 and 420 sweeps give the cost per sweep without the one-off
    device set-up.  In Philox mode (LCEQ_DROPIN_RNG=philox, streams seeded as
    f/measinit.f:214-217 seeds rank 0) the same sweeps of a device-resident
    lattice with the same key end on the same plaquette, with the copies back
    made after every routine and with them deferred (LCEQ_DROPIN_SYNC=
    deferred: the fields stay on the device until lceq_dropin_download)."""
    import time
    import lceqcd_amd
    from oracle import refrun
    shape = (24, 24, 24)
    u, a = rand_conf(shape, 11, 0.5)
    coup = dict(beta=9.0, beta_a=0.175, beta_2=-0.45, beta_4=0.025)
    per_sweep, plaq = {}, {}
    envs = {"philox": {"LCEQ_DROPIN_RNG": "philox"},
            "deferred": {"LCEQ_DROPIN_RNG": "philox", "LCEQ_DROPIN_SYNC": "deferred"},
            "reference": {}}
    for mode in ("philox", "deferred", "reference"):
        env = envs[mode]
        t = {}
        for nrep in (20, 420):
            print("drop-in schedule", mode, nrep, flush=True)
            r = refrun.run("time", shape, u, a, nrep=nrep, exe=EXE, env=env, **coup)
            t[nrep] = float(r[0][0])
            plaq[mode, nrep] = float(r[1][0])
        per_sweep[mode] = (t[420] - t[20]) / 400.0
    L = lceqcd_amd.Lattice(*shape)
    L.set_couplings(coup["beta"], coup["beta_a"], coup["beta_2"], coup["beta_4"])
    L.set_imode(0)
    L.put_links(u)
    L.put_higgs(a)
    L.rng_measinit(0)
    L.sweep(lceqcd_amd.SWEEP_REFERENCE, 20, measflag=1)
    L.history(20)
    L.sync()
    t0 = time.perf_counter()
    L.sweep(lceqcd_amd.SWEEP_REFERENCE, 400, measflag=1)
    L.sync()
    dev = (time.perf_counter() - t0) / 400.0
    p_dev = float(L.history(400)[-1, 0])
    L.close()
    print("24^3 reference schedule, ms per sweep: drop-in %.3f (Philox), %.3f (Philox, "
          "copies deferred), %.3f (reference uniforms), device-resident %.3f; plaquette "
          "drop-in %.15f, deferred %.15f, device %.15f"
          % (1e3 * per_sweep["philox"], 1e3 * per_sweep["deferred"],
             1e3 * per_sweep["reference"], 1e3 * dev, plaq["philox", 420],
             plaq["deferred", 420], p_dev), flush=True)
    assert abs(p_dev - plaq["philox", 420]) <= 1e-12 * abs(p_dev)
    assert plaq["deferred", 420] == plaq["philox", 420]
    assert 0.0 < dev and min(per_sweep.values()) > 0.0
