"""Write tests/golden/trajectory25_*.npz: 25-sweep trajectories of the
reference's schedule (f/main.f:37-41: ahiggs2, ahiggs3, agauge, renorm)
run through the reference's own Fortran built here in FP64
(oracle/_ref/refharness64, harness command 11, which captures the uniform
stream the routines consume and checks that count).  Each fixture holds the
start configuration, the captured uniforms, the per-sweep acceptances and
observables and the final fields.

    make -C oracle/ref && python tests/golden/make_trajectory_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import refrun  # noqa: E402

NREP = 25
COUP = dict(beta=9.0, beta_a=0.2, beta_2=0.6, beta_4=0.05)
CASES = [((4, 4, 6), 0), ((6, 4, 4), 1)]


def rand_conf(shape, seed, asig):
    rng = np.random.default_rng(seed)
    lx, ly, lz = shape
    n = lx * ly * lz // 2
    u = rng.normal(size=(3, 2, n, 4))
    u /= np.linalg.norm(u, axis=-1, keepdims=True)
    a = rng.normal(size=(2, n, 3)) * asig
    return u, a


def reference_trajectory(shape, imode):
    u, a = rand_conf(shape, 31, 0.6)
    recs = refrun.run("sweep", shape, u, a, nrep=NREP, seed=4321, imode=imode, **COUP)
    obs = np.array([r[0] for r in recs[:7 * NREP]]).reshape(NREP, 7)
    u_out, a_out, chk, uni = recs[7 * NREP:7 * NREP + 4]
    assert chk[0] == 1.0  # the routines consumed exactly the captured stream
    return dict(u=u, a=a, obs=obs, u_out=np.asarray(u_out), a_out=np.asarray(a_out),
                uniforms=np.asarray(uni), imode=np.int32(imode), nrep=np.int32(NREP),
                couplings=np.array([COUP[k] for k in ("beta", "beta_a", "beta_2", "beta_4")]))


def name(shape, imode):
    return "trajectory25_%s_imode%d" % ("x".join(map(str, shape)), imode)


if __name__ == "__main__":
    if not refrun.available(64):
        sys.exit("oracle/_ref/refharness64 is not built (make -C oracle/ref)")
    for shape, imode in CASES:
        d = reference_trajectory(shape, imode)
        path = os.path.join(HERE, name(shape, imode) + ".npz")
        np.savez_compressed(path, **d)
        print(path, os.path.getsize(path), "bytes")
