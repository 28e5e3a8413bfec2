"""Run the reference's own Fortran hot path through oracle/_ref/refharness.

TEST INFRASTRUCTURE ONLY.  oracle/ref/Makefile builds the harness from the
reference sources under /root/reference/f (FP64 via -fdefault-real-8, and the
FP32 build as shipped).  Used by tests/golden/make_golden.py to write golden
vectors and by CPU tests that check the C restatement when the harness exists.
"""
import os
import subprocess
import tempfile

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(_HERE, "_ref")

CMD = dict(geom=1, staples=2, agauge=3, a_force=4, ahiggs1=5, ahiggs2=6,
           ahiggs3=7, renorm=8, or_gauge_fixing=9, landau_gauge=10, sweep=11,
           heat_kpi=12, gauss=13, confsave=14, oldconf=15, time=16)


def harness_path(prec=64):
    return os.path.join(REF_DIR, "refharness%d" % prec)


def available(prec=64):
    return os.access(harness_path(prec), os.X_OK)


def _unlimited_stack():
    import resource
    hard = resource.getrlimit(resource.RLIMIT_STACK)[1]
    resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))


def big_harness_path(prec=64):
    """The harness built with 64^3 array capacity (make -C oracle/ref big)."""
    return os.path.join(REF_DIR, "refharness%d_64" % prec)


def time_sweeps(shape, u, a, nsweep, *, prec=64, seed=314159, beta=9.0, beta_a=0.0,
                beta_2=1.0, beta_4=0.0, imode=0):
    """Seconds the reference's own routines take for nsweep sweeps of
    f/main.f:37-41 (ahiggs2, ahiggs3, agauge, renorm) from (u, a), timed
    inside the harness (command 16, no uniform capture); also the last
    plaq0.  Lattices above 24^3 need the 64^3-capacity build."""
    big = max(shape) > 24
    exe = big_harness_path(prec) if big else harness_path(prec)
    recs = run("time", shape, u, a, nrep=nsweep, seed=seed, beta=beta, beta_a=beta_a,
               beta_2=beta_2, beta_4=beta_4, imode=imode, prec=prec, exe=exe)
    return float(recs[0][0]), float(recs[1][0])


DROPIN_HARNESS = os.path.join(REF_DIR, "dropin_harness64")


def build():
    subprocess.run(["make", "-s", "-C", os.path.join(_HERE, "ref")],
                   check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)


def run(cmd, shape, u, a, *, arg1=0, arg2=0, imode=0, nrep=1, seed=12345,
        beta=9.0, beta_a=0.0, beta_2=1.0, beta_4=0.0, lam=1.0, alpha=1.0,
        limit_or=1e-9, delta=0.4, ex1=None, ex2=None, prec=64, exe=None, env=None,
        workdir=None):
    """Run one harness command; returns the list of output records.  exe
    selects another build of the harness (e.g. the drop-in one); workdir is
    the working directory (configuration files of confsave / oldconf)."""
    lx, ly, lz = shape
    n = lx * ly * lz // 2
    hdr = np.array([lx, ly, lz, CMD[cmd], arg1, arg2, imode, nrep, seed],
                   dtype=np.int32)
    par = np.array([beta, beta_a, beta_2, beta_4, lam, alpha, limit_or, delta],
                   dtype=np.float64)
    ex1 = np.zeros(n) if ex1 is None else np.asarray(ex1, dtype=np.float64)
    ex2 = np.zeros(n) if ex2 is None else np.asarray(ex2, dtype=np.float64)
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        fout = os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(hdr.tobytes())
            f.write(par.tobytes())
            f.write(np.ascontiguousarray(u, dtype=np.float64).tobytes())
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
            f.write(ex1.tobytes())
            f.write(ex2.tobytes())
        e = dict(os.environ)
        e.update(env or {})
        # the reference keeps large automatic arrays on the stack
        # (f/agauge.f:36-38), as its run scripts allow with an unlimited stack
        r = subprocess.run([exe or harness_path(prec), fin, fout], env=e, timeout=300,
                           stdout=subprocess.PIPE, stderr=subprocess.PIPE, cwd=workdir,
                           preexec_fn=_unlimited_stack)
        if r.returncode != 0:
            raise RuntimeError("harness failed (%d): %s %s" % (r.returncode, r.stdout[-2000:],
                                                              r.stderr[-2000:]))
        raw = open(fout, "rb").read()
    recs = []
    off = 0
    while off < len(raw):
        m = int(np.frombuffer(raw, dtype=np.int32, count=1, offset=off)[0])
        off += 4
        recs.append(np.frombuffer(raw, dtype=np.float64, count=m, offset=off).copy())
        off += 8 * m
    return recs
