"""Throughput of the lattice sweep on one MI355X.

Workload (the metric's 64^3 production configuration): 3D SU(2) + adjoint
Higgs on a 64^3 lattice in FP64, hot start (Haar links, Higgs components
N(0, 0.5^2)), couplings beta_G = 9, beta_A = 0.33, beta_2 = -0.9, beta_4 = 0.1
(trace convention; reference normalisation beta_A/2, beta_2/2, beta_4/4).
One step = one sweep = 1 Kennedy-Pendleton heatbath link sweep + 4 link
overrelaxation sweeps (each 6 checkerboard x direction launches, with the
Higgs-hopping Metropolis) + 1 Higgs Metropolis sweep (2 parities, device-tuned
step), with plaquette / <rho> / <rho^2> / <a.F> accumulated every sweep.

    python bench.py [--steps K] [--warmup W] [--lattice 64,64,64] [--cpu-seconds S]
                    [--cpu-procs P]

cpu_baseline: the CPU restatement on P host cores (default 16), one z-slab
of the lattice per thread (an upper bound of a P-rank CPU run); the
single-threaded run of the whole lattice is config.cpu_port_one_thread.

Before the timed region the autotuner picks the launch shape (threads per
block, XCD-aware block order) from 10 sweeps per shape and restores the chain
(--shape T,X fixes it instead, e.g. for profiling).
Prints one JSON line.  value = site-updates/s (V sites per sweep).  The
roofline object also carries copy_ceiling: the rate of the library's plain
device copy (lceq_copy_bandwidth) moving the OR launch's own bytes, of a
1 GiB copy, and of 30 read streams with one write stream
(lceq_stream_bandwidth, 1 GB), timed in the same run; hbm_resident: the same OR launch timed on
a 160^3 lattice (490 MB, more than the 256 MB Infinity Cache holds, which the
64^3 field set does not exceed); and the source of its traffic figure.  The
learned object compares the learned Higgs proposal in batched form with the
plain sweep on the same replicas (4 x 48^3 at beta_A = 0.34 in the broken
phase and 4 x 48^3 near the transition, 20000 sweeps each, one network):
tau_int of Tr A^2, acceptance, independent samples/s and their ratio, also
with the collection and training time charged (--no-learned leaves it out).
The replicas96 object times the 96-replica scan (539 MB of fields) in the
library's default replica groups and as one group, 48 sweeps each
(--no-replicas leaves it out).  The reference_schedule object times the
reference's own schedule (ahiggs2, ahiggs3, agauge, renorm) at 64^3 on the
GPU and the reference's Fortran sweep of it under mpiexec on --cpu-procs host
cores with its halo exchanges (--no-cpu leaves it out, with cpu_baseline).

    python bench.py --workload replicas96 [--steps K] [--group-mb M] [--hist-out FILE]

measures the batched-replica scan instead: 96 independent lattices cycling
through 16^3, 24^3, 32^3, 40^3, 48^3 and 24^2x96, on the grid
beta_A in linspace(0.28, 0.40, 12) x beta_2 in linspace(-1.2, -0.6, 8)
(trace convention, beta_G = 9, beta_4 = 0.1), seeds 12345 + r, all swept by
the same launches (replayed hipGraphs of 8 sweeps, launch shape picked by
the autotuner, which leaves the chains unchanged), in replica groups of
about M MiB of fields (the library's default without --group-mb; 0 = one
group) so that a group's fields stay in the Infinity Cache through its
sweep; value = aggregate site-updates/s.  The same batch is also timed as one
group for comparison.  --hist-out writes per-replica <Tr A^2> histograms
(JSON).

    python bench.py --workload mlp48 | mlp96 [--sweeps K]   (K: 60000 at 48^3, 20000 at 96^3)

the learned Higgs proposal at 48^3 or 96^3 near the transition (trace
convention beta_G = 9, beta_A = 0.34, beta_2 = -0.9, beta_4 = 0.1): a network
trained on 2000 plain-sweep configurations, then K timed sweeps of the plain
sweep (1 KP + 4 OR + ahiggs1) and K of the learned one (1 KP + 4 OR + learned
Higgs update) from the same configuration; reports acceptance, tau_int of
Tr A^2 (Madras-Sokal) and effectively independent samples/s = sweeps/s /
(2 tau_int); value = the learned sweep's independent samples/s.

    python bench.py --workload refsched64 [--steps K]

the reference's own schedule (f/main.f:37-41) at 64^3 with the couplings of
its shipped inputfile (beta = 9, x = 0.1, y = 0.5, imode 0): one step =
ahiggs2 + ahiggs3 + agauge + renorm; value = site-updates/s of that sweep.
Then Landau gauge fixing (lambda = 1.7, alpha = 1, limit_or = 1e-10) from the
last configuration, timed per iteration, and the three z-correlator
measurements; each dominant launch is timed with HIP events.  cpu_baseline:
the reference's own routines (oracle/_ref/refharness64_64, built from its
sources with 8-byte reals) timed on the same configuration, when built.

    python bench.py --workload scan [--scan-lattice 24,24,24] [--therm T] [--steps K]

<Tr A^2> against beta_2 at beta_A = 0.28, 0.34, 0.40 (96 replicas, one batch):
locates the symmetric/broken transition.

    python bench.py --workload mlpbatch

only the batched learned-vs-plain comparison of the default line.

    python bench.py --workload plain | gaugefix32 | production

the benchmark configurations as runs (lceqcd_amd.runs): 16^3 cold start with
per-sweep plaquette, Tr A^2 and Polyakov loop; 10 saved 32^3 configurations
Landau-fixed to 1e-12 with the correlators; 64^3 and 32^2x128 at beta_A =
0.30, 0.33, 0.36, 5000 sweeps with gauge fixing and correlators every 50.

    python bench.py --workload kernels [--lattice L,L,L] [--steps K]

each routine K times on one lattice: the command the per-kernel rocprofv3
counter passes profile (a few hundred dispatches).

    python bench.py --workload ceilings
    python bench.py --workload sizes [--steps K]
    python bench.py --workload groups

the copy and multi-stream read ceilings; the production sweep on every
single-lattice size of the benchmark configurations (16^3 to 96^3) and on
128^3 and 160^3; replica batches of 378-755 MB swept in the default replica
groups and as one group; one JSON line each.

    python bench.py --calibrate-copy

copies a 512 MiB device buffer three times (a plain copy kernel of known
size): rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE over it give the factors that
turn those counters into bytes for the lattice kernels.
"""
import argparse
import json
import os
import sys
import subprocess
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "site-updates/s & HBM GB/s at 64³; indep. samples/s learned vs plain"
HBM_PEAK_GBPS = 8000.0
# algorithmic bytes per lattice site for one (parity, mu) link launch:
# U_mu(eo) read+write 32, U_mu(oe) read 16, U_nu both parities for two nu 64,
# Higgs both parities 24  ->  136 B per site of the lattice (272 B per link)
LINK_PASS_BYTES_PER_SITE = 136
HIGGS_PASS_BYTES_PER_SITE = 132
SWEEP_BYTES_PER_SITE = 6 * LINK_PASS_BYTES_PER_SITE * 5 + 2 * HIGGS_PASS_BYTES_PER_SITE
# beyond-L2 bytes of one 64^3 overrelaxation launch by threads per block
# (XCD-aware order), FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00 (factors from a
# 512 MiB device copy), measured with rocprofv3 --pmc
# (profiles/k_link_or_64cubed_hbm_traffic.md)
OR_TRAFFIC_64 = {512: 33.66e6 + 4.03e6, 256: 33.68e6 + 4.04e6, 128: 33.68e6 + 4.04e6,
                 64: 33.65e6 + 4.04e6}


# couplings (trace convention) of the batched learned-vs-plain comparison:
# the broken phase of the learned-proposal configuration, and one near the
# symmetric/broken transition
LEARNED_COUPLINGS = "0.34,-0.9;0.34,0.2"


def couplings(beta_a_trace=0.33):
    from lceqcd_amd import couplings_from_trace_convention
    return couplings_from_trace_convention(9.0, beta_a_trace, -0.9, 0.1)


def cpu_baseline(shape, seconds, n_or):
    """The CPU restatement (single thread) on the same workload, bounded."""
    from oracle.oracle import OracleLattice
    O = OracleLattice(*shape)
    O.set_couplings(*couplings())
    O.hotstart(12345, 0.5)
    O.philox(12345, 0)
    d = 0.4
    sweeps = 0
    t0 = time.perf_counter()
    while True:
        O.agauge(1)
        for _ in range(n_or):
            O.link_or(1)
        d, _ = O.ahiggs1(d, 1)
        sweeps += 1
        el = time.perf_counter() - t0
        if el >= seconds or sweeps >= 200:
            break
    V = shape[0] * shape[1] * shape[2]
    return {"value": V * sweeps / el, "unit": "site-updates/s", "cores": 1, "kind": "port",
            "sample": "%d sweeps (1 KP + %d OR + ahiggs1) of the %dx%dx%d workload, %.1f s, "
                      "oracle/lceq_oracle.c single-threaded" % (sweeps, n_or, *shape, el)}


def cpu_baseline_slabs(shape, seconds, n_or, procs):
    """The CPU restatement on `procs` host cores at once: thread p sweeps its
    own lx x ly x (lz/procs) z-slab lattice (the share of one rank in a
    z-decomposition), periodic in z instead of a halo exchange, so this is an
    upper bound of a `procs`-rank CPU run.  The oracle's C calls release the
    interpreter lock and keep no shared state.  The rate is that of the slowest
    slab (all ranks would advance in step).  None unless lz/procs is even."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle.oracle import OracleLattice
    lx, ly, lz = shape
    if lz % procs or (lz // procs) % 2:
        return None
    sub = (lx, ly, lz // procs)

    def run(p):
        O = OracleLattice(*sub)
        O.set_couplings(*couplings())
        O.hotstart(12345 + p, 0.5)
        O.philox(12345 + p, 0)
        d, sweeps, t0 = 0.4, 0, time.perf_counter()
        while True:
            O.agauge(1)
            for _ in range(n_or):
                O.link_or(1)
            d, _ = O.ahiggs1(d, 1)
            sweeps += 1
            el = time.perf_counter() - t0
            if el >= seconds:
                return sweeps, el
    with ThreadPoolExecutor(max_workers=procs) as ex:
        res = list(ex.map(run, range(procs)))
    rate = min(k / el for k, el in res)
    return {"value": lx * ly * lz * rate, "unit": "site-updates/s", "cores": procs, "kind": "port",
            "sample": "%d threads, each sweeping (1 KP + %d OR + ahiggs1) its own %dx%dx%d z-slab "
                      "of the %dx%dx%d workload for %.0f s (%d-%d sweeps) with oracle/lceq_oracle.c, "
                      "periodic in z instead of a halo exchange: an upper bound of a %d-rank CPU "
                      "run, at the slowest slab's rate"
                      % (procs, n_or, *sub, *shape, seconds, min(k for k, _ in res),
                         max(k for k, _ in res), procs)}



def replicas96(args):
    import numpy as np
    import lceqcd_amd
    from lceqcd_amd.workloads import C4_BETA_A, C4_BETA_2, C4_STRADDLE_BETA_2, c4_replicas
    nrep = 96
    ba = C4_BETA_A
    b2 = C4_STRADDLE_BETA_2 if args.grid == "straddle" else C4_BETA_2
    shapes, coup, seeds, _ = c4_replicas(nrep, ba, b2)
    B = lceqcd_amd.Batch(shapes)
    B.set_couplings(coup)
    B.hotstart(seeds, 0.5)
    B.rng_philox(seeds)
    if args.group_mb is not None:
        B.set_group_bytes(int(args.group_mb * (1 << 20)))
    B.sweep(args.warmup, n_or=args.n_or, delta=0.4, measflag=0)
    if args.shape:
        t, x = (int(v) for v in args.shape.split(","))
        B.set_shape(t, bool(x))
        tuned = {}
    else:
        tuned = B.autotune(n_or=args.n_or, trials=3)
    B.sync()
    t0 = time.perf_counter()
    B.sweep(args.steps, n_or=args.n_or, measflag=1)
    B.sync()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    V = B.total_sites
    hist = B.history(min(args.steps, 1024))
    # the same batch swept as one group, for comparison (the chains go on)
    group_bytes, first = B.groups()
    one_ms = None
    if len(first) > 1:
        B.set_group_bytes(0)
        n1 = min(args.steps, 48)
        B.sweep(8, n_or=args.n_or, measflag=0)
        B.sync()
        t2 = time.perf_counter()
        B.sweep(n1, n_or=args.n_or, measflag=0)
        B.sync()
        one_ms = (time.perf_counter() - t2) * 1e3 / n1
        B.set_group_bytes(group_bytes)
    tra2 = hist[:, :, 2] / 2.0  # Tr A^2 = |a|^2 / 2 per site (A = a.sigma/2)
    rec = {
        "metric": METRIC,
        "value": V / (ms * 1e-3),
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "96 replicas (16^3..48^3, 24^2x96) on a beta_A x beta_2 grid, 1 KP + %d OR "
                        "+ ahiggs1 per sweep, hipGraph of 8 sweeps, autotuned launch shape, "
                        "replica groups of %d MiB of fields" % (args.n_or, group_bytes >> 20),
            "grid": {"beta_a": [float(x) for x in ba], "beta_2": [float(x) for x in b2],
                     "kind": args.grid},
            "total_sites": V,
            "sweep_algorithmic_GBps": SWEEP_BYTES_PER_SITE * V / (ms * 1e-3) / 1e9,
            "sweep_algorithmic_frac_of_hbm": SWEEP_BYTES_PER_SITE * V / (ms * 1e-3) / 1e9
            / HBM_PEAK_GBPS,
            "tr_a2_mean_range": [float(tra2.mean(0).min()), float(tra2.mean(0).max())],
            "replica_groups": {
                "group_bytes": group_bytes, "groups": len(first), "first_replicas": first,
                "field_MB": 120 * V / 1e6,
                "one_group_ms_per_sweep": one_ms,
                "speedup_over_one_group": None if one_ms is None else one_ms / ms},
            "launch_shape": {"threads_per_block": B.get_shape()[0], "xcd_order": B.get_shape()[1]},
            "autotune_ms_per_sweep": {"%d%s" % (t, "x" if x else ""): v
                                      for (t, x), v in tuned.items()},
        },
    }
    if args.hist_out:
        edges = np.linspace(0, max(1e-9, float(tra2.max())), 41)
        with open(args.hist_out, "w") as f:
            json.dump({"edges": edges.tolist(), "replicas": [
                {"shape": shapes[r], "beta_a": ba[r % 12], "beta_2": b2[r // 12],
                 "hist": np.histogram(tra2[:, r], edges)[0].tolist(),
                 "mean": float(tra2[:, r].mean()), "std": float(tra2[:, r].std())}
                for r in range(nrep)]}, f)
    print(json.dumps(rec), flush=True)


def replica_groups_summary(n_or, steps=48):
    """The 96-replica scan (literal grid) swept in the library's default
    replica groups and as one group, for the default line: autotuned launch
    shape, `steps` timed sweeps each way (8-sweep graphs)."""
    import lceqcd_amd
    from lceqcd_amd.workloads import c4_replicas
    shapes, coup, seeds, _ = c4_replicas(96)
    B = lceqcd_amd.Batch(shapes)
    B.set_couplings(coup)
    B.hotstart(seeds, 0.5)
    B.rng_philox(seeds)
    B.sweep(8, n_or=n_or, delta=0.4, measflag=0)
    B.autotune(n_or=n_or, trials=2)

    def timed():
        B.sweep(8, n_or=n_or, measflag=0)
        B.sync()
        t0 = time.perf_counter()
        B.sweep(steps, n_or=n_or, measflag=0)
        B.sync()
        return (time.perf_counter() - t0) * 1e3 / steps

    ms = timed()
    group_bytes, first = B.groups()
    B.set_group_bytes(0)
    one_ms = timed()
    V = B.total_sites
    B.close()
    return {"workload": "96 replicas (16^3..48^3, 24^2x96), 1 KP + %d OR + ahiggs1, "
                        "hipGraph of 8 sweeps, autotuned launch shape" % n_or,
            "total_sites": V, "field_MB": 120 * V / 1e6, "steps": steps,
            "group_bytes": group_bytes, "groups": len(first),
            "ms_per_sweep": ms, "site_updates_per_s": V / (ms * 1e-3),
            "sweep_algorithmic_GBps": SWEEP_BYTES_PER_SITE * V / (ms * 1e-3) / 1e9,
            "one_group_ms_per_sweep": one_ms, "speedup_over_one_group": one_ms / ms}


GROUP_BATCHES = [("4x96^3", [(96, 96, 96)] * 4), ("2x128^3", [(128, 128, 128)] * 2),
                 ("3x128^3", [(128, 128, 128)] * 3), ("12x64^3", [(64, 64, 64)] * 12),
                 ("96x32^3", [(32, 32, 32)] * 96)]


GROUP_SIZES_MIB = [0, 128, 192, 256, 288, 384]


def groups(args):
    """Replica groups (256 threads, XCD order, 1 KP + n_or OR + ahiggs1, best
    of three runs of 24 sweeps per point): the 96-replica scan and 48 x 48^3
    at group sizes of 0 (one group) to 384 MiB of fields, then batches of
    several compositions larger than the Infinity Cache in the library's
    default groups and as one group.  One JSON line."""
    import lceqcd_amd
    from lceqcd_amd.workloads import c4_replicas, coupling

    def timed(B, nsw=24):
        best = 1e9
        for _ in range(3):
            B.sweep(8, n_or=args.n_or, measflag=0)
            B.sync()
            t0 = time.perf_counter()
            B.sweep(nsw, n_or=args.n_or, measflag=0)
            B.sync()
            best = min(best, (time.perf_counter() - t0) * 1e3 / nsw)
        return best

    sweep = {}
    c4_shapes, c4_coup, c4_seeds, _ = c4_replicas(96)
    big48 = [(48, 48, 48)] * 48
    for name, shapes, coup, seeds in (
            ("96 replicas", c4_shapes, c4_coup, c4_seeds),
            ("48x48^3", big48, [coupling(0.34, -0.9 + 0.02 * r) for r in range(48)],
             [9 + r for r in range(48)])):
        B = lceqcd_amd.Batch(shapes)
        B.set_couplings(coup)
        B.hotstart(seeds, 0.5)
        B.rng_philox(seeds)
        B.set_shape(256, True)
        B.sweep(8, n_or=args.n_or, delta=0.4, measflag=0)
        pts = {}
        for mib in GROUP_SIZES_MIB:
            B.set_group_bytes(mib << 20)
            ms = timed(B)
            pts[mib] = {"groups": len(B.groups()[1]), "ms_per_sweep": ms}
            _log("groups %s at %d MiB: %d groups %.4f ms" % (name, mib, pts[mib]["groups"], ms))
        sweep[name] = pts
        B.close()
    out = {}
    for name, shapes in GROUP_BATCHES:
        n = len(shapes)
        B = lceqcd_amd.Batch(shapes)
        B.set_couplings([coupling(0.34, -0.9 + 0.02 * r) for r in range(n)])
        B.hotstart([5 + r for r in range(n)], 0.5)
        B.rng_philox([5 + r for r in range(n)])
        B.set_shape(256, True)
        B.sweep(8, n_or=args.n_or, delta=0.4, measflag=0)
        g_ms = timed(B)
        group_bytes, first = B.groups()
        B.set_group_bytes(0)
        one_ms = timed(B)
        V = sum(x * y * z for x, y, z in shapes)
        out[name] = {"field_MB": 120 * V / 1e6, "group_bytes": group_bytes,
                     "first_replicas": first, "grouped_ms_per_sweep": g_ms,
                     "one_group_ms_per_sweep": one_ms, "speedup": one_ms / g_ms,
                     "grouped_site_updates_per_s": V / (g_ms * 1e-3)}
        _log("groups %s: %d groups %.4f ms, one group %.4f ms"
             % (name, len(first), g_ms, one_ms))
        B.close()
    print(json.dumps({"workload": "groups", "group_sizes": sweep, "batches": out}), flush=True)


def _batch_history(B, nsweep, step_fn, chunk=1024):
    """Run nsweep sweeps of a batch in chunks of at most `chunk` (the
    history ring's size) and return the (nsweep, nrep, 16) history."""
    import numpy as np
    out, left = [], nsweep
    while left > 0:
        k = min(left, chunk)
        step_fn(k)
        out.append(B.history(k))
        left -= k
    return np.concatenate(out)


def scan(args):
    """<Tr A^2> against beta_2 at three beta_A (trace convention, beta_G = 9,
    beta_4 = 0.1): one replica batch of 24^3 lattices, hot start, 1 KP + 4
    OR + ahiggs1 per sweep; the transition is where <Tr A^2> falls fastest
    and its variance per site (the susceptibility) peaks."""
    import numpy as np
    import lceqcd_amd
    from lceqcd_amd.workloads import coupling
    ba = [0.28, 0.34, 0.40]
    b2 = np.linspace(-0.6, 2.5, 32)
    grid = [(a, b) for a in ba for b in b2]
    shape = tuple(int(x) for x in args.scan_lattice.split(","))
    V = shape[0] * shape[1] * shape[2]
    B = lceqcd_amd.Batch([shape] * len(grid))
    B.set_couplings([coupling(a, b) for a, b in grid])
    seeds = [20000 + r for r in range(len(grid))]
    B.hotstart(seeds, 0.5)
    B.rng_philox(seeds)
    B.sweep(args.therm, n_or=4, delta=0.4, measflag=0)
    B.sync()
    t0 = time.perf_counter()
    h = _batch_history(B, args.steps, lambda k: B.sweep(k, n_or=4, measflag=1))
    el = time.perf_counter() - t0
    tra2 = h[:, :, 2] / 2.0
    res = {}
    for i, a in enumerate(ba):
        rows = []
        for j, b in enumerate(b2):
            x = tra2[:, i * len(b2) + j]
            rows.append({"beta_2": float(b), "tr_a2": float(x.mean()),
                         "chi": float(V * x.var()), "plaq": float(h[:, i * len(b2) + j, 0].mean())})
        m = np.array([r["tr_a2"] for r in rows])
        chi = np.array([r["chi"] for r in rows])
        k = int(np.argmin(np.diff(m)))
        res["%.2f" % a] = {"points": rows,
                           "steepest_drop_between": [float(b2[k]), float(b2[k + 1])],
                           "chi_peak_beta_2": float(b2[int(np.argmax(chi))])}
    print(json.dumps({"workload": "scan", "lattice": list(shape), "replicas": len(grid),
                      "therm": args.therm, "sweeps": args.steps, "seconds": el,
                      "site_updates_per_s": len(grid) * V * args.steps / el,
                      "beta_a": res}), flush=True)


REF_SWEEP_BYTES_PER_SITE = 2 * 2 * HIGGS_PASS_BYTES_PER_SITE + 6 * LINK_PASS_BYTES_PER_SITE + 192
GFIX_PASS_BYTES_PER_SITE = 216   # all links read + written, a_f of one parity read + written
SLICES_BYTES_PER_SITE = 144      # u_f 96 + a_f 24 + a 24


def cpu_baseline_reference(L, V, seconds):
    """The reference's own CPU routines (oracle/_ref: its ahiggs2, ahiggs3,
    agauge and renorm built with 8-byte reals, one rank, array capacity
    64^3) on the lattice's current configuration: about `seconds` of sweeps.
    None when that build is absent."""
    from oracle import refrun
    if not os.access(refrun.big_harness_path(64), os.X_OK):
        return None
    u, a = L.get_links(), L.get_higgs()
    from lceqcd_amd import couplings_from_xy
    ba, b2, b4 = couplings_from_xy(9.0, 0.1, 0.5)
    kw = dict(prec=64, beta=9.0, beta_a=ba, beta_2=b2, beta_4=b4, imode=0)
    t1, _ = refrun.time_sweeps(L.shape, u, a, 1, **kw)
    k = max(1, min(100, int(seconds / max(t1, 1e-3))))
    tk, _ = refrun.time_sweeps(L.shape, u, a, k, **kw)
    return {"value": V * k / tk, "unit": "site-updates/s", "cores": 1, "kind": "reference",
            "sample": "%d sweeps of f/main.f:37-41 (the reference's ahiggs2, ahiggs3, agauge, "
                      "renorm; 8-byte reals, one rank) on the same %dx%dx%d configuration, "
                      "%.1f s timed inside oracle/_ref/refharness64_64" % ((k,) + L.shape + (tk,))}


def cpu_baseline_reference_slabs(L, seconds, procs):
    """The reference's own routines on `procs` host cores at once: process p
    sweeps z-slab p of the lattice's current configuration (lz/procs planes,
    the share of one rank in a z-decomposition over `procs` ranks) with
    periodic wrap in z instead of a halo exchange, so this is an upper bound
    of the reference's `procs`-rank MPI throughput.  About `seconds` of
    sweeps; None when the reference build is absent or lz % procs != 0."""
    from concurrent.futures import ThreadPoolExecutor
    import numpy as np
    from oracle import refrun
    lx, ly, lz = L.shape
    if not os.access(refrun.big_harness_path(64), os.X_OK) or lz % procs or (lz // procs) % 2:
        return None
    u, a = L.get_links(), L.get_higgs()
    dz = lz // procs
    h2 = lx * ly // 2
    shape = (lx, ly, dz)
    slabs = [(np.ascontiguousarray(u[:, :, h2 * dz * p:h2 * dz * (p + 1), :]),
              np.ascontiguousarray(a[:, h2 * dz * p:h2 * dz * (p + 1), :])) for p in range(procs)]
    from lceqcd_amd import couplings_from_xy
    ba, b2, b4 = couplings_from_xy(9.0, 0.1, 0.5)
    kw = dict(prec=64, beta=9.0, beta_a=ba, beta_2=b2, beta_4=b4, imode=0)

    def run_all(k):
        with ThreadPoolExecutor(max_workers=procs) as ex:
            return list(ex.map(lambda sl: refrun.time_sweeps(shape, sl[0], sl[1], k, **kw)[0],
                               slabs))
    t1 = max(run_all(1))
    k = max(1, min(2000, int(seconds / max(t1, 1e-3))))
    tk = max(run_all(k))
    V = lx * ly * lz
    return {"value": V * k / tk, "unit": "site-updates/s", "cores": procs, "kind": "reference",
            "sample": "%d concurrent processes of the reference's own f/main.f:37-41 routines "
                      "(ahiggs2, ahiggs3, agauge, renorm; 8-byte reals), each sweeping one "
                      "%dx%dx%d z-slab of the same %dx%dx%d configuration %d times with periodic "
                      "wrap instead of the halo exchange: an upper bound of the reference's "
                      "%d-rank MPI throughput; slowest process %.1f s inside "
                      "oracle/_ref/refharness64_64" % ((procs,) + shape + L.shape + (k, procs, tk))}


MPI_GRIDS = {1: (1, 1, 1), 2: (1, 1, 2), 4: (1, 2, 2), 8: (2, 2, 2), 16: (2, 2, 4)}


def cpu_baseline_reference_mpi(shape, procs_list, seconds, timeout=600):
    """The reference's own sweep under MPI with its halo exchanges
    (oracle/_ref/refmpi32_64: its ahiggs2, ahiggs3, agauge, renorm and
    typeconstruct, adresses as shipped, 4-byte reals, the rank grid of
    f/topology.f), run with mpiexec on 1..16 host cores; about `seconds` of
    sweeps per rank count, each mpiexec run limited to `timeout` seconds.
    Returns one record per rank count, or None when the build or mpiexec is
    absent."""
    import resource
    import subprocess
    exe = os.path.join(ROOT, "oracle", "_ref", "refmpi32_64")
    mpiexec = "/opt/conda/bin/mpiexec"
    if not (os.access(exe, os.X_OK) and os.access(mpiexec, os.X_OK)):
        return None

    def stack():
        hard = resource.getrlimit(resource.RLIMIT_STACK)[1]
        resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))

    def run(procs, nwarm, nsweep):
        grid = MPI_GRIDS[procs]
        cmd = [mpiexec, "-n", str(procs), exe] + [str(v) for v in shape + grid] + [
            str(nwarm), str(nsweep)]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                           timeout=timeout, preexec_fn=stack, cwd=ROOT)
        line = [x for x in r.stdout.splitlines() if x.startswith("refmpi")]
        if r.returncode != 0 or not line:
            raise RuntimeError("refmpi failed: %s %s" % (r.stdout[-500:], r.stderr[-500:]))
        f = line[-1].split()
        return float(f[9]), float(f[10])

    V = shape[0] * shape[1] * shape[2]
    out = []
    for procs in procs_list:
        grid = MPI_GRIDS[procs]
        if any(n % (2 * g) for n, g in zip(shape, grid)):
            continue
        t1, _ = run(procs, 1, 2)
        k = max(3, min(2000, int(seconds / max(t1, 1e-3))))
        tk, plaq = run(procs, 2, k)
        out.append({"value": V / tk, "unit": "site-updates/s", "cores": procs,
                    "kind": "reference", "rank_grid": list(grid), "seconds_per_sweep": tk,
                    "plaq0": plaq,
                    "sample": "mpiexec -n %d oracle/_ref/refmpi32_64: the reference's own "
                              "f/main.f:37-41 sweep (ahiggs2, ahiggs3, agauge, renorm, 4-byte "
                              "reals as shipped) on %dx%dx%d over a %dx%dx%d rank grid with its "
                              "halo exchanges, %d timed sweeps, slowest rank"
                              % ((procs,) + tuple(shape) + tuple(grid) + (k,))})
    return out


def reference_schedule_summary(seconds, procs=16):
    """For the default line: the reference schedule (f/main.f:37-41:
    ahiggs2, ahiggs3, agauge, renorm) at 64^3 on the GPU, and the
    reference's own Fortran sweep of the same schedule under MPI on `procs`
    host cores with its halo exchanges (about `seconds` of sweeps), at the
    couplings refmpi.f sets."""
    import lceqcd_amd
    shape = (64, 64, 64)
    V = shape[0] * shape[1] * shape[2]
    L = lceqcd_amd.Lattice(*shape)
    L.set_couplings(9.0, 0.175, -0.45, 0.025)
    L.hotstart(314159, 0.5)
    L.rng_philox(314159, 0)
    L.sweep(lceqcd_amd.SWEEP_REFERENCE, 20, measflag=1)
    L.history(20)
    L.sync()
    n = 200
    t0 = time.perf_counter()
    L.sweep(lceqcd_amd.SWEEP_REFERENCE, n, measflag=1)
    L.sync()
    ms = (time.perf_counter() - t0) * 1e3 / n
    L.history(n)
    L.close()
    rec = {"workload": "64^3 reference schedule (ahiggs2, ahiggs3, agauge, renorm), hot start, "
                       "beta=9 beta_a=0.175 beta_2=-0.45 beta_4=0.025, %d sweeps" % n,
           "gpu_ms_per_sweep": ms, "gpu_site_updates_per_s": V / (ms * 1e-3)}
    try:
        mpi = cpu_baseline_reference_mpi(shape, [procs], seconds, timeout=max(60.0, 4 * seconds))
    except Exception as e:  # noqa: BLE001 - recorded, the line still prints
        rec["reference_mpi_error"] = str(e)[-300:]
        return rec
    if mpi:
        rec["reference_mpi"] = mpi[-1]
        rec["gpu_over_reference_mpi"] = rec["gpu_site_updates_per_s"] / mpi[-1]["value"]
    return rec


def refsched64(args):
    """The reference schedule, gauge fixing and correlators at 64^3."""
    import lceqcd_amd
    from lceqcd_amd import couplings_from_xy
    shape = tuple(int(x) for x in args.lattice.split(","))
    V = shape[0] * shape[1] * shape[2]
    L = lceqcd_amd.Lattice(*shape)
    L.set_couplings(9.0, *couplings_from_xy(9.0, 0.1, 0.5))
    L.set_imode(0)
    L.hotstart(314159, 0.5)
    L.rng_philox(314159, 0)
    L.sweep(lceqcd_amd.SWEEP_REFERENCE, args.warmup, measflag=1)
    L.history(args.warmup)
    if args.shape:
        t, x = (int(v) for v in args.shape.split(","))
        L.set_shape(t, bool(x))
    else:  # tuned on the production sweep; the chain is restored
        L.autotune(n_or=args.n_or, trials=10)
    L.sync()
    t0 = time.perf_counter()
    L.sweep(lceqcd_amd.SWEEP_REFERENCE, args.steps, measflag=1)
    L.sync()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    hist = L.history(args.steps)
    # agauge's KP launches and the ahiggs2/ahiggs3 launches, event-timed
    L.profile_begin(lceqcd_amd.KC_LINK_KP, 6 * args.steps)
    L.sweep(lceqcd_amd.SWEEP_REFERENCE, args.steps, measflag=1)
    kp_ms, kp_n = L.profile_end()
    L.history(args.steps)
    L.profile_begin(lceqcd_amd.KC_HIGGS, 4 * args.steps)
    L.sweep(lceqcd_amd.SWEEP_REFERENCE, args.steps, measflag=1)
    hg_ms, hg_n = L.profile_end()
    L.history(args.steps)
    # Landau gauge fixing at the inputfile's parameters (f/landau_gauge.f)
    L.set_gauge_fixing(1.7, 1.0, 1e-10)
    L.sync()
    t0 = time.perf_counter()
    da, iters = L.landau_gauge()
    gf_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    ncorr = 20
    for _ in range(ncorr):
        L.correlator()
        L.scalarcorr()
        L.polyakov()
    corr_ms = (time.perf_counter() - t0) * 1e3 / ncorr
    kp_us = kp_ms * 1e3 / max(kp_n, 1)
    kp_bytes = LINK_PASS_BYTES_PER_SITE * V
    # one iteration = two k_gfix launches (one per parity)
    gf_us = gf_ms * 1e3 / max(iters, 1)
    gf_gbps = 2 * GFIX_PASS_BYTES_PER_SITE * V / (gf_us * 1e-6) / 1e9
    rec = {
        "metric": METRIC,
        "value": V / (ms * 1e-3),
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "%dx%dx%d reference schedule ahiggs2 + ahiggs3 + agauge + renorm "
                        "(beta=9, x=0.1, y=0.5, imode 0), hot start; Landau gauge fixing "
                        "lambda=1.7 alpha=1 limit 1e-10; correlator + scalarcorr + polyakov"
                        % shape,
            "lattice": list(shape),
            "sweep_algorithmic_GBps": REF_SWEEP_BYTES_PER_SITE * V / (ms * 1e-3) / 1e9,
            "launch_shape": {"threads_per_block": L.get_shape()[0],
                             "xcd_order": L.get_shape()[1]},
            "higgs_launch_us": hg_ms * 1e3 / max(hg_n, 1),
            "landau_gauge": {"iterations": iters, "da": da, "ms": gf_ms,
                             "us_per_iteration": gf_us,
                             "bytes_per_iteration": 2 * GFIX_PASS_BYTES_PER_SITE * V,
                             "algorithmic_GBps": gf_gbps,
                             "frac_of_hbm": gf_gbps / HBM_PEAK_GBPS},
            "correlators_ms": corr_ms,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_link<true,false> (agauge: KP heatbath + hopping Metropolis)",
            "achieved": kp_bytes / (kp_us * 1e-6) / 1e9,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": kp_bytes / (kp_us * 1e-6) / 1e9 / HBM_PEAK_GBPS,
            "traffic": None,
            "bytes_per_launch": kp_bytes,
            "avg_launch_us": kp_us,
            "launches": kp_n,
        },
        "observables_last_sweep": {"plaq0": hist[-1, 0], "a02": hist[-1, 2],
                                   "a04": hist[-1, 3], "a0kin": hist[-1, 1],
                                   "link_accept": hist[-1, 7], "higgs_accept": hist[-1, 8]},
    }
    if not args.no_cpu:
        cb = cpu_baseline_reference(L, V, args.cpu_seconds)
        cm = cpu_baseline_reference_slabs(L, args.cpu_seconds, args.cpu_procs)
        try:
            mpi = cpu_baseline_reference_mpi(shape, [p for p in (1, 2, 4, 8, 16)
                                                     if p <= args.cpu_procs],
                                             args.cpu_seconds / 4)
        except (RuntimeError, OSError, subprocess.SubprocessError) as e:
            mpi = None
            rec["config"]["reference_mpi_error"] = str(e)[:500]
        if mpi:
            rec["cpu_baseline"] = mpi[-1]
            rec["config"]["reference_mpi_scaling"] = mpi
            if cm is not None:
                rec["config"]["reference_slabs_upper_bound"] = cm
        elif cm is not None:
            rec["cpu_baseline"] = cm
        elif cb is not None:
            rec["cpu_baseline"] = cb
        if cb is not None:
            rec["config"]["reference_one_rank"] = cb
    print(json.dumps(rec), flush=True)


def learned(args, L0):
    import numpy as np
    import lceqcd_amd
    from lceqcd_amd import couplings_from_trace_convention
    from lceqcd_amd.analysis import tau_int
    shape = (L0,) * 3
    V = L0 ** 3
    coup = couplings_from_trace_convention(9.0, 0.34, -0.9, 0.1)
    L = lceqcd_amd.Lattice(*shape)
    L.set_couplings(*coup)
    L.hotstart(12345, 0.5)
    L.rng_philox(12345, 0)
    d = L.sweep(lceqcd_amd.SWEEP_KP_OR_METRO, args.therm, n_or=4, delta=0.4, measflag=0)
    # training: 2000 configurations, a subsample of sites from each
    M = lceqcd_amd.Mlp(seed=1)
    stride = max(1, V // 2 // 3456)
    t0 = time.perf_counter()
    for c in range(args.train_configs):
        d = L.sweep(lceqcd_amd.SWEEP_KP_OR_METRO, 1, n_or=4, delta=d, measflag=0)
        M.collect(L, nsteps=8, stride=stride)
    t1 = time.perf_counter()
    loss = M.train(args.train_steps, batch=8192, lr=2e-3, seed=1)
    L.sync()
    t2 = time.perf_counter()
    u0, a0 = L.get_links(), L.get_higgs()

    def chain(learned_update):
        X = lceqcd_amd.Lattice(*shape)
        X.set_couplings(*coup)
        X.put_links(u0)
        X.put_higgs(a0)
        X.rng_philox(777 if learned_update else 778, 0)
        dd = d
        hist = []
        X.sync()
        ts = time.perf_counter()
        left = args.sweeps
        while left > 0:
            k = min(left, 4000)
            if learned_update:
                X.sweep_mlp(M, k, n_or=4, measflag=1)
            else:
                dd = X.sweep(lceqcd_amd.SWEEP_KP_OR_METRO, k, n_or=4, delta=dd, measflag=1)
            hist.append(X.history(k))
            left -= k
        X.sync()
        el = time.perf_counter() - ts
        h = np.concatenate(hist)
        tra2 = h[:, 2] / 2.0  # Tr A^2 = |a|^2 / 2 per site
        tau, dtau, w = tau_int(tra2)
        sps = args.sweeps / el
        return {"sweeps_per_s": sps, "site_updates_per_s": V * sps, "tau_int": tau,
                "tau_int_err": dtau, "window": int(w), "acceptance": float(h[:, 8].mean()),
                "tr_a2_mean": float(tra2.mean()), "indep_samples_per_s": sps / (2.0 * tau)}

    plain = chain(False)
    lrn = chain(True)
    ratio = lrn["indep_samples_per_s"] / plain["indep_samples_per_s"]
    rec = {
        "metric": METRIC,
        "value": lrn["indep_samples_per_s"],
        "unit": "independent samples/s of <Tr A^2>",
        "n_gpus": 1,
        "steps": args.sweeps,
        "warmup": args.therm,
        "ms_per_step": 1e3 / lrn["sweeps_per_s"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (network bf16 MFMA, fp32 accumulate)",
        "data": "synthetic",
        "config": {
            "workload": "%d^3 learned Higgs proposal vs plain Metropolis, beta_G=9 beta_A=0.34 "
                        "beta_2=-0.9 beta_4=0.1 (trace convention), 1 KP + 4 OR per sweep" % L0,
            "train": {"configs": args.train_configs, "records": M.records(),
                      "steps": args.train_steps, "batch": 8192,
                      "loss_first": float(loss[:20].mean()), "loss_last": float(loss[-20:].mean()),
                      "collect_s": t1 - t0, "train_s": t2 - t1},
            "plain": plain,
            "learned": lrn,
            "learned_over_plain": ratio,
        },
    }
    print(json.dumps(rec), flush=True)


def learned_batch(args, log=None):
    """The learned Higgs proposal in batched form against the plain sweep
    on the same replicas: R replicas of an L^3 lattice per coupling (two
    couplings: the broken phase beta_A = 0.34, beta_2 = -0.9, and one near
    the symmetric/broken transition), one network for all replicas (the
    couplings are inputs), trained on records the batch's own plain sweeps
    generate.  From one thermalised state per replica, S sweeps of the plain
    batch (1 KP + 4 OR + ahiggs1) and S of the learned batch (1 KP + 4 OR +
    learned update), both replayed from hipGraphs.  Per replica tau_int of
    Tr A^2 (Madras-Sokal); effectively independent samples/s = sweeps/s /
    (2 tau_int), summed over the replicas of a coupling; the amortised
    figure charges the learned chain with the collection and training time."""
    import numpy as np
    import lceqcd_amd
    from lceqcd_amd.analysis import tau_int
    from lceqcd_amd.workloads import coupling
    L0 = args.learned_lattice
    shape = (L0,) * 3
    V = L0 ** 3
    groups = [tuple(float(v) for v in g.split(",")) for g in args.learned_couplings.split(";")]
    R = args.learned_replicas
    grid = [g for g in groups for _ in range(R)]
    nrep = len(grid)
    B = lceqcd_amd.Batch([shape] * nrep)
    B.set_couplings([coupling(a, b) for a, b in grid])
    seeds = [31000 + r for r in range(nrep)]
    B.hotstart(seeds, 0.5)
    B.rng_philox(seeds)
    B.sweep(args.therm, n_or=4, delta=0.4, measflag=0)
    M = lceqcd_amd.Mlp(seed=1)
    stride = max(1, V // 2 // 3456)
    B.sync()
    t0 = time.perf_counter()
    for c in range(args.learned_train_configs):
        B.sweep(1, n_or=4, measflag=0, graph=False)
        B.mlp_collect(M, nsteps=8, stride=stride)
    t1 = time.perf_counter()
    loss = M.train(args.train_steps, batch=8192, lr=2e-3, seed=1)
    t2 = time.perf_counter()
    if log:
        log("learned batch: %d records collected in %.1f s, trained in %.1f s"
            % (M.records(), t1 - t0, t2 - t1))
    state = [(B.get_links(r), B.get_higgs(r)) for r in range(nrep)]
    out = {}
    for name in ("plain", "learned"):
        for r in range(nrep):
            B.put_links(r, state[r][0])
            B.put_higgs(r, state[r][1])
        B.rng_philox([41000 + 1000 * (name == "learned") + r for r in range(nrep)])
        if name == "plain":
            step = lambda k: B.sweep(k, n_or=4, measflag=1)
        else:
            step = lambda k: B.sweep_mlp(M, k, n_or=4, measflag=1)
        step(8)  # graph capture outside the timed chain
        B.history(8)
        B.sync()
        ts = time.perf_counter()
        h = _batch_history(B, args.learned_sweeps, step)
        el = time.perf_counter() - ts
        out[name] = (h, el)
        if log:
            log("learned batch: %s chain %d sweeps in %.1f s" % (name, args.learned_sweeps, el))
    res = {"lattice": list(shape), "replicas_per_coupling": R, "sweeps": args.learned_sweeps,
           "train": {"configs": args.learned_train_configs, "records": M.records(),
                     "steps": args.train_steps, "collect_s": t1 - t0, "train_s": t2 - t1,
                     "loss_first": float(loss[:20].mean()), "loss_last": float(loss[-20:].mean())},
           "couplings": []}
    for gi, (ba, b2) in enumerate(groups):
        reps = range(gi * R, (gi + 1) * R)
        g = {"beta_a": ba, "beta_2": b2}
        for name in ("plain", "learned"):
            h, el = out[name]
            sps = args.learned_sweeps / el  # sweeps/s of every replica (all advance together)
            taus, errs, acc, tra = [], [], [], []
            for r in reps:
                x = h[:, r, 2] / 2.0
                t, dt, _ = tau_int(x)
                taus.append(t)
                errs.append(dt)
                acc.append(float(h[:, r, 8].mean()))
                tra.append(float(x.mean()))
            indep = sum(sps / (2.0 * t) for t in taus)
            g[name] = {"batch_sweeps_per_s": sps, "site_updates_per_s": sps * V * nrep,
                       "tau_int": taus, "tau_int_err": errs, "acceptance": float(np.mean(acc)),
                       "tr_a2_mean": float(np.mean(tra)), "indep_samples_per_s": indep}
        g["learned_over_plain"] = g["learned"]["indep_samples_per_s"] / g["plain"]["indep_samples_per_s"]
        # training charged to this coupling's share of the learned chain
        share = R / nrep
        el_l = out["learned"][1]
        n_ind = g["learned"]["indep_samples_per_s"] * el_l
        g["learned_over_plain_amortised"] = (n_ind / (el_l + share * (t2 - t0))) \
            / g["plain"]["indep_samples_per_s"]
        res["couplings"].append(g)
    return res


def hbm_resident_or(shape, n_or, tb_x, sweeps=10, order=0):
    """The overrelaxation launch on a lattice too large for the 256 MB
    Infinity Cache (160^3: 490 MB of fields, 557 MB moved per launch), timed
    with HIP events as the 64^3 launch is, at the same launch shape and site
    order `order` (-1: the one the site-order tuner picks)."""
    import lceqcd_amd
    V = shape[0] * shape[1] * shape[2]
    L = lceqcd_amd.Lattice(*shape)
    L.set_couplings(*couplings())
    L.hotstart(12345, 0.5)
    L.rng_philox(12345, 0)
    L.set_shape(*tb_x)
    d = L.sweep(lceqcd_amd.SWEEP_KP_OR_METRO, 2, n_or=n_or, delta=0.4, measflag=0)
    if order < 0:
        orders = L.autotune_site_order(n_or=n_or, trials=3)
    else:
        L.set_site_order(order)
        orders = {}
    L.profile_begin(lceqcd_amd.KC_LINK_OR, 6 * n_or * sweeps)
    L.sweep(lceqcd_amd.SWEEP_KP_OR_METRO, sweeps, n_or=n_or, delta=d, measflag=1)
    ms, n = L.profile_end()
    L.history(sweeps)
    so = L.get_site_order()
    L.close()
    us = ms * 1e3 / max(n, 1)
    b = LINK_PASS_BYTES_PER_SITE * V
    gbps = b / (us * 1e-6) / 1e9
    return {"lattice": list(shape), "bytes_per_launch": b, "avg_launch_us": us, "launches": n,
            "achieved": gbps, "frac": gbps / HBM_PEAK_GBPS,
            "site_order": so, "site_order_ms_per_sweep": orders}


def kernels(args):
    """Each routine of the update and measurement path a few times on one
    lattice (default 64^3), for per-kernel profiles: rocprofv3 --pmc
    serialises every dispatch, so a profiled command must stay short.  The
    JSON line gives the host time per call (launch and synchronisation
    included); the kernel durations come from the profiler's trace."""
    import lceqcd_amd
    from lceqcd_amd import couplings_from_xy
    shape = tuple(int(x) for x in args.lattice.split(","))
    L = lceqcd_amd.Lattice(*shape)
    L.set_couplings(9.0, *couplings_from_xy(9.0, 0.1, 0.5))
    L.set_imode(0)
    L.hotstart(314159, 0.5)
    L.rng_philox(314159, 0)
    if args.shape:
        L.set_shape(*(int(v) for v in args.shape.split(",")))
    if args.order > 0:
        L.set_site_order(args.order)
    L.sweep(lceqcd_amd.SWEEP_REFERENCE, 2, measflag=1)
    L.history(2)
    L.set_gauge_fixing(1.7, 1.0, 1e-10)
    L.put_links_fixed(L.get_links())
    L.put_higgs_fixed(L.get_higgs())
    calls = [("ahiggs2", lambda: L.ahiggs2(1)), ("ahiggs3", lambda: L.ahiggs3(1)),
             ("ahiggs1", lambda: L.ahiggs1(0.4, 1)), ("agauge", lambda: L.agauge(1)),
             ("link_or", lambda: L.link_or(1)), ("renorm", L.renorm),
             ("or_gauge_fixing", L.or_gauge_fixing), ("u_f_renorm", L.u_f_renorm),
             ("correlator", L.correlator), ("scalarcorr", L.scalarcorr),
             ("polyakov", L.polyakov)]
    out = {}
    for name, fn in calls:
        fn()
        L.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        L.sync()
        out[name] = 1e3 * (time.perf_counter() - t0) / args.steps
    print(json.dumps({"workload": "kernels", "lattice": list(shape), "calls": args.steps,
                      "host_ms_per_call": out}), flush=True)


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def flows(args):
    """The plain-method, gauge-fixed and production configurations as runs
    (lceqcd_amd.runs), one JSON line each."""
    from lceqcd_amd import runs
    from lceqcd_amd.workloads import C3
    if args.workload == "plain":
        rec = runs.run_plain()
    elif args.workload == "gaugefix32":
        rec = runs.run_gauge_fixed()
    else:
        rec = [runs.run_production(shape, ba, progress=_log) for shape in C3["shapes"]
               for ba in C3["beta_a"]]
    print(json.dumps({"workload": args.workload, "result": rec}), flush=True)


def copy_rate(nbytes, reps):
    """Read+write GB/s of the library's plain device copy of nbytes."""
    import ctypes
    from lceqcd_amd import _lib
    us = ctypes.c_double()
    nbytes = max(16, int(nbytes) // 16 * 16)
    _lib.check(_lib.load().lceq_copy_bandwidth(nbytes, reps, ctypes.byref(us)))
    return 2 * nbytes / (us.value * 1e-6) / 1e9


def stream_rate(streams, elements, reps):
    """GB/s of the library's multi-stream probe: `streams` arrays of
    `elements` 16-byte values read at the same index, one written."""
    import ctypes
    from lceqcd_amd import _lib
    us = ctypes.c_double()
    _lib.check(_lib.load().lceq_stream_bandwidth(int(streams), int(elements), int(reps),
                                                 ctypes.byref(us)))
    return (streams + 1) * elements * 16 / (us.value * 1e-6) / 1e9


def copy_ceiling(launch_bytes, achieved):
    """The practical bandwidth beside the 8 TB/s peak: a plain copy that
    moves the launch's own bytes (half read, half written), one of 1 GiB,
    which no cache holds, and 30 read streams + 1 write stream of one 160^3
    parity class each (1.0 GB, no cache holds it either), the access shape of
    the link launch's ~30 operand arrays."""
    same = copy_rate(launch_bytes / 2, 200)
    return {"same_bytes_GBps": same, "hbm_resident_1GiB_GBps": copy_rate(1 << 30, 20),
            "hbm_resident_30_streams_GBps": stream_rate(30, 160 ** 3 // 2, 10),
            "achieved_frac_of_same_bytes_copy": achieved / same}


def ceilings():
    """The practical bandwidth ceilings the link launches are compared with:
    plain device copies of 35.65 MB (one 64^3 OR launch's bytes) and 1 GiB,
    and the multi-stream read probe with 8, 16 and 30 read streams of one
    64^3 and one 160^3 parity class each (one write stream).  One JSON line,
    GB/s."""
    out = {"copy_35.65MB": copy_rate(35651584 / 2, 200), "copy_1GiB": copy_rate(1 << 30, 20)}
    for L in (64, 160):
        n = L ** 3 // 2
        for k in (8, 16, 30):
            out["streams%d_%dcubed_class" % (k, L)] = stream_rate(k, n, 200 if L == 64 else 10)
    print(json.dumps({"workload": "ceilings", "GBps": out}), flush=True)


SIZES = [(16, 16, 16), (24, 24, 96), (32, 32, 32), (48, 48, 48), (64, 64, 64), (32, 32, 128),
         (96, 96, 96), (128, 128, 128), (160, 160, 160)]


def sizes(args):
    """The production sweep (1 KP + n_or OR + ahiggs1, hot start, the 64^3
    line's couplings) on every single-lattice size of the benchmark
    configurations, and on 128^3 and 160^3 (whose fields no cache holds):
    autotuned launch shape, then args.steps timed sweeps (scaled down to
    about 1.2e9 site-updates at the larger sizes) and the OR launch from HIP
    events.  One JSON line."""
    import lceqcd_amd
    out = []
    for shape in SIZES:
        V = shape[0] * shape[1] * shape[2]
        steps = max(10, min(args.steps, int(1.2e9 / (V * 1.0 + 1))))
        L = lceqcd_amd.Lattice(*shape)
        L.set_couplings(*couplings())
        L.hotstart(12345, 0.5)
        L.rng_philox(12345, 0)
        d = L.sweep(lceqcd_amd.SWEEP_KP_OR_METRO, 10, n_or=args.n_or, delta=0.4, measflag=0)
        L.autotune(n_or=args.n_or, trials=3 if V > 2e6 else 10)
        L.sync()
        t0 = time.perf_counter()
        d = L.sweep(lceqcd_amd.SWEEP_KP_OR_METRO, steps, n_or=args.n_or, delta=d, measflag=0)
        L.sync()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        L.profile_begin(lceqcd_amd.KC_LINK_OR, 6 * args.n_or * steps)
        d = L.sweep(lceqcd_amd.SWEEP_KP_OR_METRO, steps, n_or=args.n_or, delta=d, measflag=0)
        or_ms, or_n = L.profile_end()
        or_us = 1e3 * or_ms / max(or_n, 1)
        rec = {"lattice": list(shape), "sites": V, "fields_MB": 120 * V / 1e6, "steps": steps,
               "ms_per_sweep": ms, "site_updates_per_s": V / (ms * 1e-3),
               "sweep_algorithmic_GBps": SWEEP_BYTES_PER_SITE * V / (ms * 1e-3) / 1e9,
               "or_launch_us": or_us,
               "or_launch_GBps": LINK_PASS_BYTES_PER_SITE * V / (or_us * 1e-6) / 1e9,
               "launch_shape": list(L.get_shape())}
        _log("sizes %s: %.4f ms/sweep, %.3g site-updates/s, OR %.2f us"
             % ("x".join(map(str, shape)), ms, rec["site_updates_per_s"], or_us))
        out.append(rec)
        L.close()
    print(json.dumps({"workload": "sizes", "n_or": args.n_or, "peak_GBps": HBM_PEAK_GBPS,
                      "sizes": out}), flush=True)


def calibrate_copy():
    import torch
    n = 64 * 1024 * 1024  # 64 Mi doubles = 512 MiB per buffer
    src = torch.rand(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    print(json.dumps({"calibration": "device copy", "bytes_per_copy": n * 8, "copies": 3,
                      "read_plus_write_GBps": 3 * 2 * n * 8 / el / 1e9}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--lattice", default="64,64,64")
    ap.add_argument("--n-or", type=int, default=4)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-procs", type=int, default=16,
                    help="host cores of the multi-core CPU baselines (z-slabs of the lattice)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="single64",
                    choices=["single64", "replicas96", "mlp48", "mlp96", "refsched64", "scan",
                             "mlpbatch", "plain", "gaugefix32", "production", "kernels",
                             "ceilings", "sizes", "groups"])
    ap.add_argument("--sweeps", type=int, default=None,
                    help="sweeps per chain (learned workloads; default 60000 at 48^3, "
                         "20000 at 96^3: a few % statistical error on tau_int)")
    ap.add_argument("--train-configs", type=int, default=2000)
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--therm", type=int, default=300,
                    help="thermalisation sweeps before training (learned workloads)")
    ap.add_argument("--hist-out", default=None)
    ap.add_argument("--group-mb", type=float, default=None,
                    help="replicas96: replica group size in MiB of fields (0: one group; "
                         "default: the library's)")
    ap.add_argument("--scan-lattice", default="24,24,24")
    ap.add_argument("--order", type=int, default=0,
                    help="site order (strip rows, 0 = index order, -1 = the site-order "
                         "tuner's choice) of the measured lattice")
    ap.add_argument("--big-order", type=int, default=0,
                    help="site order of the 160^3 HBM-resident launch (-1: tuned)")
    ap.add_argument("--grid", default="literal", choices=["literal", "straddle"],
                    help="replicas96: the configuration's beta_2 grid (broken phase "
                         "throughout) or one that straddles the transition")
    ap.add_argument("--no-learned", action="store_true",
                    help="leave the batched learned-vs-plain object out of the default line")
    ap.add_argument("--no-replicas", action="store_true",
                    help="leave the 96-replica object out of the default line")
    ap.add_argument("--learned-lattice", type=int, default=48)
    ap.add_argument("--learned-replicas", type=int, default=4,
                    help="replicas per coupling of the batched learned comparison")
    ap.add_argument("--learned-couplings", default=LEARNED_COUPLINGS,
                    help="beta_A,beta_2 pairs (trace convention) separated by ';'")
    ap.add_argument("--learned-sweeps", type=int, default=20000)
    ap.add_argument("--learned-train-configs", type=int, default=250)
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--shape", default=None,
                    help="threads_per_block,xcd_order for the link and Higgs launches "
                         "(skips the autotuner; e.g. 256,1)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(1)
    if args.calibrate_copy:
        return calibrate_copy()
    if args.workload == "replicas96":
        return replicas96(args)
    if args.workload == "refsched64":
        return refsched64(args)
    if args.workload == "scan":
        return scan(args)
    if args.workload == "kernels":
        return kernels(args)
    if args.workload == "ceilings":
        return ceilings()
    if args.workload == "sizes":
        return sizes(args)
    if args.workload == "groups":
        return groups(args)
    if args.workload == "mlpbatch":
        print(json.dumps({"workload": "mlpbatch", "learned": learned_batch(args, log=_log)}),
              flush=True)
        return None
    if args.workload in ("plain", "gaugefix32", "production"):
        return flows(args)
    if args.sweeps is None:
        args.sweeps = 60000 if args.workload == "mlp48" else 20000
    if args.workload in ("mlp48", "mlp96"):
        return learned(args, int(args.workload[3:]))

    import lceqcd_amd
    shape = tuple(int(x) for x in args.lattice.split(","))
    V = shape[0] * shape[1] * shape[2]
    L = lceqcd_amd.Lattice(*shape)
    L.set_couplings(*couplings())
    L.hotstart(12345, 0.5)
    L.rng_philox(12345, 0)
    d = L.sweep(lceqcd_amd.SWEEP_KP_OR_METRO, args.warmup, n_or=args.n_or, delta=0.4, measflag=1)
    L.history(args.warmup)
    if args.shape:
        t, x = (int(v) for v in args.shape.split(","))
        L.set_shape(t, bool(x))
        tuned = {}
    else:
        tuned = L.autotune(n_or=args.n_or, trials=10)
    if args.order < 0:
        tuned_order = L.autotune_site_order(n_or=args.n_or, trials=10)
    else:
        L.set_site_order(args.order)
        tuned_order = {}
    L.sync()

    # timed region: K sweeps, synchronised on both sides
    t0 = time.perf_counter()
    d = L.sweep(lceqcd_amd.SWEEP_KP_OR_METRO, args.steps, n_or=args.n_or, delta=d, measflag=1)
    L.sync()
    t1 = time.perf_counter()
    hist = L.history(args.steps)
    ms = (t1 - t0) * 1e3 / args.steps

    # the same K sweeps again with HIP events around every launch of the
    # dominant kernel (the link-overrelaxation class, 6 * n_or launches per sweep)
    nlaunch = 6 * args.n_or * args.steps
    L.profile_begin(lceqcd_amd.KC_LINK_OR, nlaunch)
    d = L.sweep(lceqcd_amd.SWEEP_KP_OR_METRO, args.steps, n_or=args.n_or, delta=d, measflag=1)
    or_ms, or_n = L.profile_end()
    L.history(args.steps)
    L.profile_begin(lceqcd_amd.KC_LINK_KP, 6 * args.steps)
    d = L.sweep(lceqcd_amd.SWEEP_KP_OR_METRO, args.steps, n_or=args.n_or, delta=d, measflag=1)
    kp_ms, kp_n = L.profile_end()
    L.history(args.steps)
    avg_s = or_ms * 1e-3 / max(or_n, 1)
    launch_bytes = LINK_PASS_BYTES_PER_SITE * V
    achieved = launch_bytes / avg_s / 1e9

    rec = {
        "metric": METRIC,
        "value": V / (ms * 1e-3),
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "%dx%dx%d SU(2)+adjoint-Higgs sweep: 1 KP heatbath + %d link OR + 1 Higgs "
                        "Metropolis, hot start, beta_G=9 beta_A=0.33 beta_2=-0.9 beta_4=0.1 "
                        "(trace convention)" % (shape + (args.n_or,)),
            "lattice": list(shape),
            "sweep_algorithmic_GBps": SWEEP_BYTES_PER_SITE * V / (ms * 1e-3) / 1e9,
            "link_updates_per_s": 3 * V * (1 + args.n_or) / (ms * 1e-3),
            "launch_shape": {"threads_per_block": L.get_shape()[0], "xcd_order": L.get_shape()[1],
                             "site_order": L.get_site_order()},
            "autotune_ms_per_sweep": {"%d%s" % (t, "x" if x else ""): v
                                      for (t, x), v in tuned.items()},
            "site_order_ms_per_sweep": tuned_order,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_link<false,false> (link overrelaxation + hopping Metropolis)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": (OR_TRAFFIC_64.get(L.get_shape()[0]) if shape == (64, 64, 64)
                        and L.get_shape()[1] else None),
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE of this launch at this "
                              "launch shape (profiles/k_link_or_64cubed_hbm_traffic.md), not "
                              "measured in this run",
            "residency": "the 64^3 field set (31 MB) sits in the 256 MB Infinity Cache, so "
                         "frac is a fraction of the HBM peak, not of HBM traffic; the "
                         "hbm_resident object times the same launch on a 160^3 lattice "
                         "(490 MB) that no cache holds",
            "bytes_per_launch": launch_bytes,
            "avg_launch_us": avg_s * 1e6,
            "launches": or_n,
            "kp_avg_launch_us": kp_ms * 1e3 / max(kp_n, 1),
            "copy_ceiling": copy_ceiling(launch_bytes, achieved),
        },
        "observables_last_sweep": {"plaq0": hist[-1, 0], "a02": hist[-1, 2],
                                   "a04": hist[-1, 3], "a0kin": hist[-1, 1],
                                   "higgs_accept": hist[-1, 8], "delta": d},
    }
    big = (160, 160, 160)
    rec["roofline"]["hbm_resident"] = hbm_resident_or(big, args.n_or, L.get_shape(),
                                                      order=args.big_order)
    rec["roofline"]["hbm_resident"]["frac_of_1GiB_copy"] = (
        rec["roofline"]["hbm_resident"]["achieved"]
        / rec["roofline"]["copy_ceiling"]["hbm_resident_1GiB_GBps"])
    rec["roofline"]["hbm_resident"]["frac_of_30_streams"] = (
        rec["roofline"]["hbm_resident"]["achieved"]
        / rec["roofline"]["copy_ceiling"]["hbm_resident_30_streams_GBps"])
    if not args.no_learned:
        rec["learned"] = learned_batch(args, log=_log)
    if not args.no_replicas:
        rec["replicas96"] = replica_groups_summary(args.n_or)
        _log("replicas96: %.4f ms/sweep in %d groups, %.4f ms as one group"
             % (rec["replicas96"]["ms_per_sweep"], rec["replicas96"]["groups"],
                rec["replicas96"]["one_group_ms_per_sweep"]))
    if not args.no_cpu:
        rec["reference_schedule"] = reference_schedule_summary(args.cpu_seconds, args.cpu_procs)
        _log("reference schedule: GPU %.4f ms/sweep, reference under MPI %s"
             % (rec["reference_schedule"]["gpu_ms_per_sweep"],
                rec["reference_schedule"].get("reference_mpi", {}).get("seconds_per_sweep")))
        one = cpu_baseline(shape, args.cpu_seconds, args.n_or)
        many = cpu_baseline_slabs(shape, args.cpu_seconds, args.n_or, args.cpu_procs)
        rec["cpu_baseline"] = many if many is not None else one
        if many is not None:
            rec["config"]["cpu_port_one_thread"] = one
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
