"""Fixtures for the configuration file format (f/confsave.f, f/confget.f).

    python tests/golden/make_conf_golden.py

Runs the reference's own confsave and oldconf (oracle/_ref/refharness64,
built with 8-byte reals, and oracle/_ref/refharness32, the as-shipped
4-byte build) on one rank and stores, per build:
  u_bytes*, a_bytes*   the files confsave wrote for the input fields
  name_u, name_a       the file names confsave chose (beta = 9, x = 0.1,
                       y = 0.5, sweep 123)
  ref_u*, ref_a*       the fields oldconf read back from those files.
Inputs: helpers.rand_conf((6, 4, 8), 5, 1.0), so that both fields hold
non-trivial values in every component.
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))
from oracle import refrun  # noqa: E402
from helpers import rand_conf  # noqa: E402

SHAPE = (6, 4, 8)
SWEEP, BETA, X, Y = 123, 9.0, 0.1, 0.5


def main():
    u, a = rand_conf(SHAPE, 5, 1.0)
    out = {"u_in": u, "a_in": a, "shape": np.array(SHAPE)}
    for prec in (64, 32):
        with tempfile.TemporaryDirectory() as td:
            refrun.run("confsave", SHAPE, u, a, arg1=SWEEP, beta=BETA, lam=X, alpha=Y,
                       prec=prec, workdir=td)
            names = sorted(os.listdir(td))
            assert len(names) == 2, names
            name_a, name_u = names  # a_conf_... sorts before u_conf_...
            ub = open(os.path.join(td, name_u), "rb").read()
            ab = open(os.path.join(td, name_a), "rb").read()
            os.rename(os.path.join(td, name_u), os.path.join(td, "u_in"))
            os.rename(os.path.join(td, name_a), os.path.join(td, "a_in"))
            recs = refrun.run("oldconf", SHAPE, np.zeros_like(u), np.zeros_like(a), prec=prec,
                              workdir=td)
        out["u_bytes%d" % prec] = np.frombuffer(ub, np.uint8)
        out["a_bytes%d" % prec] = np.frombuffer(ab, np.uint8)
        out["ref_u%d" % prec] = recs[0].reshape(u.shape)
        out["ref_a%d" % prec] = recs[1].reshape(a.shape)
        out["name_u"] = np.array(name_u)
        out["name_a"] = np.array(name_a)
    np.savez_compressed(os.path.join(HERE, "conf_6x4x8.npz"), **out)


if __name__ == "__main__":
    main()
