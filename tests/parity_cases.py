"""Parity cases shared by the CPU oracle tests and the GPU tests.

Each case takes an engine factory -- ``oracle.OracleLattice`` (the CPU
restatement) or ``lceqcd_amd.Lattice`` (the HIP product); both expose the
reference routine names -- replays one golden fixture written by the
reference's own FP64 Fortran (tests/golden/make_golden.py), and returns
(name, got, expected, tolerance) tuples.
"""
import numpy as np

from helpers import golden


def _setup(make, fx, couplings=True):
    shape = tuple(int(x) for x in fx["shape"])
    L = make(*shape)
    if couplings and "couplings" in fx:
        L.set_couplings(*[float(x) for x in fx["couplings"]])
    if "u" in fx:
        L.put_links(fx["u"].reshape(3, 2, -1, 4))
    if "a" in fx:
        L.put_higgs(fx["a"].reshape(2, -1, 3))
    return L, shape


def case_staples(make, tol=1e-13):
    fx = golden("staples_6x4x8")
    L, shape = _setup(make, fx)
    got = np.zeros_like(fx["v"])
    for eo in (0, 1):
        for mu in (1, 2, 3):
            L.staples(eo, mu)
            got[mu - 1, eo] = L.get_staple()[mu - 1, eo]
    return [("staples v", got, fx["v"], tol)]


def case_a_force(make, tol=1e-13):
    fx = golden("a_force_6x4x8")
    L, shape = _setup(make, fx)
    out = []
    for eo in (0, 1):
        L.a_force(eo)
        out.append(("force eo=%d" % eo, L.get_force(), fx["force"][eo], tol))
    return out


def case_agauge(make, name, tol=1e-12):
    fx = golden(name)
    L, shape = _setup(make, fx)
    L.rng_feed(fx["uniforms"])
    out = []
    for k in range(len(fx["accept"])):
        acc = L.agauge(1)
        obs = L.observables()
        out.append(("accept[%d]" % k, acc, fx["accept"][k], tol))
        out.append(("plaq0[%d]" % k, obs["plaq0"], fx["plaq0"][k], tol))
    out.append(("u", L.get_links().ravel(), fx["u_out"], tol))
    out.append(("consumed", L.rng_state()[0], len(fx["uniforms"]), 0))
    return out


def case_higgs(make, name, tol=1e-12):
    fx = golden(name)
    L, shape = _setup(make, fx)
    L.set_imode(int(fx["imode"]))
    L.rng_feed(fx["uniforms"])
    delta = float(fx["delta"])
    out = []
    kind = name.split("_")[0]
    for k in range(2):
        if kind == "ahiggs1":
            delta, acc = L.ahiggs1(delta, 1)
        elif kind == "ahiggs2":
            acc = L.ahiggs2(1)
        else:
            acc = L.ahiggs3(1)
        obs = L.observables()
        got = [acc, delta, obs["a02"], obs["a04"], obs["a0kin"]]
        for j, lab in enumerate(("accept", "delta", "a02", "a04", "a0kin")):
            out.append(("%s[%d]" % (lab, k), got[j], fx["obs"][k][j], tol))
    out.append(("a", L.get_higgs().ravel(), fx["a_out"], tol))
    out.append(("consumed", L.rng_state()[0], len(fx["uniforms"]), 0))
    return out


def case_renorm(make, tol=1e-14):
    fx = golden("renorm_6x4x8")
    L, shape = _setup(make, fx)
    L.renorm()
    return [("u", L.get_links().ravel(), fx["u_out"], tol)]


def case_or_gauge_fixing(make, tol=1e-12):
    fx = golden("or_gauge_fixing_6x4x8")
    L, shape = _setup(make, fx)
    L.set_gauge_fixing(float(fx["lam"]), float(fx["alpha"]), 1e-9)
    L.put_links_fixed(fx["u"])
    L.put_higgs_fixed(fx["a"])
    out = []
    for k in range(3):
        da = L.or_gauge_fixing()
        L.u_f_renorm()
        out.append(("da[%d]" % k, da, fx["da"][k], tol))
    out.append(("u_f", L.get_links_fixed().ravel(), fx["u_f"], tol))
    out.append(("a_f", L.get_higgs_fixed().ravel(), fx["a_f"], tol))
    return out


def case_landau_gauge(make, name, tol=1e-10):
    fx = golden(name)
    L, shape = _setup(make, fx)
    L.set_gauge_fixing(float(fx["lam"]), float(fx["alpha"]), float(fx["limit_or"]))
    da, it = L.landau_gauge()
    # iteration counts can differ by one when da lands within rounding of
    # limit_or; the fields are compared only when the counts agree
    out = [("|iter_or1 - ref| <= 1", int(abs(it - int(fx["iter_or1"])) <= 1), 1, 0),
           ("da1 <= limit_or", int(da <= float(fx["limit_or"])), 1, 0)]
    if it == int(fx["iter_or1"]):
        out.append(("da1", da, float(fx["da1"]), 1e-6))
        out.append(("u_f", L.get_links_fixed().ravel(), fx["u_f"], tol))
        out.append(("a_f", L.get_higgs_fixed().ravel(), fx["a_f"], tol))
    return out


def case_sweep(make, name, tol=1e-11):
    fx = golden(name)
    L, shape = _setup(make, fx)
    L.rng_feed(fx["uniforms"])
    out = []
    for k in range(fx["obs"].shape[0]):
        a2 = L.ahiggs2(1)
        a3 = L.ahiggs3(1)
        a1 = L.agauge(1)
        L.renorm()
        obs = L.observables()
        got = [a2, a3, a1, obs["plaq0"], obs["a0kin"], obs["a02"], obs["a04"]]
        for j, lab in enumerate(("acc2", "acc3", "acc1", "plaq0", "a0kin", "a02", "a04")):
            out.append(("%s[%d]" % (lab, k), got[j], fx["obs"][k][j], tol))
    out.append(("u", L.get_links().ravel(), fx["u_out"], tol))
    out.append(("a", L.get_higgs().ravel(), fx["a_out"], tol))
    out.append(("consumed", L.rng_state()[0], len(fx["uniforms"]), 0))
    return out


def case_heat_kpi(make, tol=1e-13):
    fx = golden("heat_kpi_256")
    L = make(8, 8, 8)
    L.rng_feed(fx["uniforms"])
    v, nacch = L.heat_kpi(fx["b"], fx["a0_old"])
    return [("v", v, fx["v"], tol), ("nacch", nacch, int(fx["nacch"]), 0)]


def case_gauss(make, tol=1e-13):
    fx = golden("gauss_256")
    L = make(8, 8, 8)
    L.rng_feed(fx["uniforms"])
    x = L.gauss(fx["r"])
    return [("x", x, fx["x"], tol)]


AGAUGE = ["agauge_6x4x8_mixed", "agauge_4x4x4_inputfile", "agauge_8x8x8_production"]
HIGGS = ["ahiggs1_6x4x8_production", "ahiggs1_6x4x8_mixed", "ahiggs2_6x4x8_imode0",
         "ahiggs2_6x4x8_imode1", "ahiggs3_6x4x8_imode0", "ahiggs3_6x4x8_imode1",
         "ahiggs2_6x4x8_inputfile"]
LANDAU = ["landau_gauge_4x4x4_iso", "landau_gauge_4x4x4_alpha05"]
SWEEPS = ["sweep_4x4x4", "sweep_8x8x8"]


def all_cases():
    """(id, callable(make)) for every golden case."""
    cs = [("staples", case_staples), ("a_force", case_a_force), ("renorm", case_renorm),
          ("or_gauge_fixing", case_or_gauge_fixing), ("heat_kpi", case_heat_kpi),
          ("gauss", case_gauss)]
    cs += [(n, (lambda n: lambda make: case_agauge(make, n))(n)) for n in AGAUGE]
    cs += [(n, (lambda n: lambda make: case_higgs(make, n))(n)) for n in HIGGS]
    cs += [(n, (lambda n: lambda make: case_landau_gauge(make, n))(n)) for n in LANDAU]
    cs += [(n, (lambda n: lambda make: case_sweep(make, n))(n)) for n in SWEEPS]
    return cs


def check_all(results):
    from helpers import relerr
    bad = []
    for name, got, exp, tol in results:
        if tol == 0:
            if int(np.asarray(got)) != int(np.asarray(exp)):
                bad.append("%s: %s != %s" % (name, got, exp))
        else:
            e = relerr(got, exp)
            if not e <= tol:
                bad.append("%s: rel dev %.3e > %.1e" % (name, e, tol))
    assert not bad, "; ".join(bad)
