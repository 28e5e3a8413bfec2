"""The benchmark configurations as runnable flows on one device:

- :func:`run_plain` (C1): 16^3 cold start (unitconf, f/confget.f:295-307),
  streams seeded by RANLUX level 3 from the given seed (as
  f/measinit.f:214-217 seeds random_number from 314159 + rank), sweeps of 1
  KP + 4 OR + 1 Higgs Metropolis (initial step 0.4, self-tuned as
  f/ahiggs1.f:117-123), plaquette, Tr A^2 and the Polyakov loop
  (f/polyakov.f) every sweep;
- :func:`run_gauge_fixed` (C2): configurations saved every `spacing` sweeps in
  the reference's file format (f/confsave.f), each read back (f/confget.f
  oldconf), Landau gauge-fixed (f/landau_gauge.f, f/or_gauge_fixing1.f) with
  lambda = 1.7, alpha = 1 to da < 1e-12, then the z-correlators of
  f/correlator.f and f/scalarcorr.f;
- :func:`run_production` (C3): hot start, 1 KP + 4 OR + ahiggs1 per sweep,
  every `meas_every` sweeps Landau gauge fixing plus the correlators, as
  the measurement branch of f/main.f:58-79 does every imeas sweeps.

    python -m lceqcd_amd.runs plain|gauge_fixed|production [--quick]

Each returns (and the command prints) a JSON-serialisable record with the
timings and the measured series.  Couplings in the trace convention
(lceqcd_amd.workloads)."""
import argparse
import json
import os
import tempfile
import time

import numpy as np

from . import workloads as W
from .analysis import mean_error, screening_mass, tau_int
from .lattice import Lattice, SWEEP_KP_OR_METRO
from .ranlux import Ranlux


def ranlux_key(seed, luxury=3):
    """A Philox key from RANLUX(luxury, seed): 16 numbers int(10^6 u) as
    f/measinit.f:214-216 forms its seeds, key = s(1) + 2^32 s(2)."""
    R = Ranlux()
    R.rluxgo(luxury, int(seed), 0, 0)
    s = (np.float32(1000000.0) * np.asarray(R.ranlux(16), np.float32)).astype(np.int64)
    R.close()
    return int(s[0]) + (int(s[1]) << 32), s.tolist()


def _sync_time(L):
    L.sync()
    return time.perf_counter()


def run_plain(sweeps=None, seed=None, shape=None, beta_a=None, beta_2=None):
    """C1: cold start, per-sweep plaquette, Tr A^2 and Polyakov loop."""
    c = W.C1
    sweeps = c["sweeps"] if sweeps is None else sweeps
    shape = c["shape"] if shape is None else tuple(shape)
    seed = c["seed"] if seed is None else seed
    ba = c["beta_a"] if beta_a is None else beta_a
    b2 = c["beta_2"] if beta_2 is None else beta_2
    L = Lattice(*shape)
    L.set_couplings(*W.coupling(ba, b2))
    L.unitconf()
    key, seeds = ranlux_key(seed)
    L.rng_philox(key, 0)
    d = c["delta"]
    rows = []
    t0 = _sync_time(L)
    for it in range(sweeps):
        d = L.sweep(SWEEP_KP_OR_METRO, 1, n_or=4, delta=d, measflag=1)
        h = L.history(1)[0]
        psq, _ = L.polyakov()
        rows.append([h[0], W.tr_a2(h[2]), np.sqrt(psq), h[8], d])
    el = _sync_time(L) - t0
    rows = np.array(rows)
    L.close()
    half = sweeps // 2
    return {
        "config": "C1", "shape": list(shape), "beta_a": ba, "beta_2": b2, "sweeps": sweeps,
        "seed": seed, "ranlux_seeds": seeds, "philox_key": key,
        "seconds": el, "ms_per_sweep_with_measurements": 1e3 * el / max(sweeps, 1),
        "series": {"plaquette": rows[:, 0].tolist(), "tr_a2": rows[:, 1].tolist(),
                   "polyakov": rows[:, 2].tolist(), "higgs_accept": rows[:, 3].tolist(),
                   "delta": rows[:, 4].tolist()},
        "second_half_means": {k: list(mean_error(rows[half:, i]))
                              for i, k in enumerate(("plaquette", "tr_a2", "polyakov"))},
    }


def run_gauge_fixed(configs=None, spacing=20, therm=200, shape=None, limit_or=None, workdir=None):
    """C2: saved configurations, read back, Landau gauge-fixed, correlators."""
    c = W.C2
    configs = c["configs"] if configs is None else configs
    shape = c["shape"] if shape is None else tuple(shape)
    limit_or = c["limit_or"] if limit_or is None else limit_or
    L = Lattice(*shape)
    L.set_couplings(*W.coupling(c["beta_a"], c["beta_2"]))
    L.hotstart(c["seed"], 0.5)
    L.rng_philox(c["seed"], 0)
    d = L.sweep(SWEEP_KP_OR_METRO, therm, n_or=4, delta=0.4, measflag=0)
    own = workdir is None
    wd = tempfile.mkdtemp(prefix="lceq_c2_") if own else workdir
    files = []
    for k in range(configs):
        d = L.sweep(SWEEP_KP_OR_METRO, spacing, n_or=4, delta=d, measflag=0)
        fu, fa = os.path.join(wd, "u_conf_%03d" % k), os.path.join(wd, "a_conf_%03d" % k)
        L.confsave(fu, fa)
        files.append((fu, fa))
    G = Lattice(*shape)
    G.set_gauge_fixing(c["lam"], c["alpha"], limit_or, 10000000)
    out = []
    for fu, fa in files:
        G.oldconf(fu, fa)
        t0 = _sync_time(G)
        da, iters = G.landau_gauge()
        t1 = _sync_time(G)
        ca, cu, c3 = G.correlator()
        m, cs = G.scalarcorr()
        t2 = _sync_time(G)
        out.append({"iterations": iters, "da": da, "gauge_fixing_ms": 1e3 * (t1 - t0),
                    "us_per_iteration": 1e6 * (t1 - t0) / max(iters, 1),
                    "correlators_ms": 1e3 * (t2 - t1), "corr_a": ca.tolist(),
                    "corr_u": cu.tolist(), "corr_3": c3.tolist(), "scalar_mean": m,
                    "scalar_corr": cs.tolist()})
    if own:
        for fu, fa in files:
            os.remove(fu)
            os.remove(fa)
        os.rmdir(wd)
    L.close()
    G.close()
    it = np.array([o["iterations"] for o in out])
    return {"config": "C2", "shape": list(shape), "configs": configs, "spacing": spacing,
            "lambda": c["lam"], "alpha": c["alpha"], "limit_or": limit_or,
            "iterations_mean": float(it.mean()),
            "us_per_iteration_mean": float(np.mean([o["us_per_iteration"] for o in out])),
            "measurements": out}


def run_production(shape, beta_a, sweeps=None, meas_every=None, therm=0, seed=None,
                   limit_or=None, progress=None):
    """C3 at one lattice and coupling: sweeps with gauge fixing plus the
    correlators every meas_every sweeps (f/main.f:58-79's measurement)."""
    c = W.C3
    sweeps = c["sweeps"] if sweeps is None else sweeps
    meas_every = c["meas_every"] if meas_every is None else meas_every
    seed = c["seed"] if seed is None else seed
    limit_or = c["limit_or"] if limit_or is None else limit_or
    shape = tuple(shape)
    V = shape[0] * shape[1] * shape[2]
    L = Lattice(*shape)
    L.set_couplings(*W.coupling(beta_a, c["beta_2"]))
    L.set_gauge_fixing(c["lam"], c["alpha"], limit_or, 10000000)
    L.hotstart(seed, 0.5)
    L.rng_philox(seed, 0)
    d = L.sweep(SWEEP_KP_OR_METRO, therm, n_or=4, delta=0.4, measflag=0) if therm else 0.4
    t_sweep = t_gf = t_corr = 0.0
    iters, hist, corr_s, corr_a, corr_u = [], [], [], [], []
    done = 0
    while done < sweeps:
        k = min(meas_every, sweeps - done)
        t0 = _sync_time(L)
        d = L.sweep(SWEEP_KP_OR_METRO, k, n_or=4, delta=d, measflag=1)
        t1 = _sync_time(L)
        hist.append(L.history(k))
        done += k
        t_sweep += t1 - t0
        if k == meas_every:
            t1 = _sync_time(L)
            da, it = L.landau_gauge()
            t2 = _sync_time(L)
            ca, cu, c3 = L.correlator()
            m, cs = L.scalarcorr()
            t3 = _sync_time(L)
            t_gf += t2 - t1
            t_corr += t3 - t2
            iters.append(it)
            corr_s.append(cs)
            corr_a.append(ca)
            corr_u.append(cu)
            if progress:
                progress("%s beta_A %.2f: %d/%d sweeps, gauge fixing %d iterations"
                         % ("x".join(map(str, shape)), beta_a, done, sweeps, it))
    L.close()
    h = np.concatenate(hist)
    tra2 = W.tr_a2(h[:, 2])
    rec = {"config": "C3", "shape": list(shape), "beta_a": beta_a, "beta_2": c["beta_2"],
           "sweeps": sweeps, "meas_every": meas_every, "measurements": len(iters),
           "ms_per_sweep": 1e3 * t_sweep / max(sweeps, 1),
           "site_updates_per_s": V * sweeps / max(t_sweep, 1e-12),
           "gauge_fixing_s": t_gf, "gauge_fixing_iterations_mean":
               float(np.mean(iters)) if iters else None,
           "us_per_gauge_fixing_iteration": 1e6 * t_gf / max(sum(iters), 1),
           "correlators_ms_per_measurement": 1e3 * t_corr / max(len(iters), 1),
           "plaquette": list(mean_error(h[:, 0])), "tr_a2": list(mean_error(tra2)),
           "tau_int_tr_a2": float(tau_int(tra2)[0])}
    if corr_s:
        rec["corr_a_mean"] = np.mean(corr_a, axis=0).tolist()
        rec["corr_u_mean"] = np.mean(corr_u, axis=0).tolist()
        rec["scalar_corr_mean"] = np.mean(corr_s, axis=0).tolist()
    if len(corr_s) >= 4:
        m, e = screening_mass(np.array(corr_s), z=1, nblk=min(20, len(corr_s)))
        rec["scalar_screening_mass"] = [m, e]
    return rec


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("config", choices=["plain", "gauge_fixed", "production"])
    ap.add_argument("--quick", action="store_true", help="small sizes (a smoke run)")
    args = ap.parse_args()
    say = lambda m: print(m, flush=True)
    if args.config == "plain":
        rec = run_plain(sweeps=20 if args.quick else None)
    elif args.config == "gauge_fixed":
        rec = run_gauge_fixed(configs=2 if args.quick else None,
                              shape=(8, 8, 8) if args.quick else None)
    else:
        rec = []
        shapes = [(8, 8, 16)] if args.quick else W.C3["shapes"]
        for shape in shapes:
            for ba in W.C3["beta_a"]:
                rec.append(run_production(shape, ba, sweeps=100 if args.quick else None,
                                          progress=say))
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
