"""Fixtures for the RANLUX restatement (f/luxury.f).

    python tests/golden/make_ranlux_golden.py

ranlux_luxtst.json: the expected output that f/luxtst.f:88-146 lists for its
own test program (parsed from that file's comment block).
ranlux_ref.npz: bit patterns of the reference's luxury.f, built single
precision by oracle/ref/Makefile into oracle/_ref/refranlux, for the streams
and seeding calls in CASES below.
"""
import json
import os
import re
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LUXTST = "/root/reference/f/luxtst.f"
REFRANLUX = os.path.join(ROOT, "oracle", "_ref", "refranlux")

# (luxury, seed, k1, lengths of the successive RANLUX calls)
CASES = [(0, 1, 0, (1, 23, 24, 25, 300)), (1, 12345, 0, (100, 7, 250)),
         (2, 2147483646, 0, (97, 97, 97)), (3, 314159, 0, (48, 400)),
         (4, 42, 0, (389, 11)), (24, 5, 0, (120,)), (2000, 777, 0, (50, 3)),
         (250, 99, 0, (240,)), (3, 1000, 223 * 3, (30,)), (4, 7674985, 161840, (200,))]
RANKS = list(range(16)) + [255, 4095]


def parse_luxtst():
    lines = open(LUXTST).read().splitlines()
    start = next(i for i, l in enumerate(lines) if "SHOULD BE" in l)
    block = [l.lstrip("!").strip() for l in lines[start:]]
    out = {"sequences": [], "restart": [], "rluxat": [], "skip_pairs": []}
    i = 0
    while i < len(block):
        l = block[i]
        m = re.match(r"RANLUX (.*?)(\d+)\s*-\s*(\d+):$", l)
        if m:
            out["sequences"].append({"label": l, "values": [float(x) for x in block[i + 1].split()]})
            i += 2
            continue
        if l.startswith("current RANLUX status saved"):
            vals = []
            for k in range(1, 6):
                vals += [int(x) for x in block[i + k].split()]
            out["restart"] = vals
            i += 6
            continue
        if l.startswith("RLUXAT values"):
            out["rluxat"].append([int(x) for x in l.split("=")[1].split()])
        if l.startswith("Next and 200th"):
            out["skip_pairs"].append([float(x) for x in l.split(":")[1].split()])
        i += 1
    return out


def run_ref(script):
    r = subprocess.run([REFRANLUX], input=script, capture_output=True, text=True, check=True)
    return np.array([int(l[1:]) for l in r.stdout.splitlines() if l.startswith("D")], np.int64)


def main():
    kat = parse_luxtst()
    assert len(kat["sequences"]) == 12 and len(kat["restart"]) == 25
    with open(os.path.join(HERE, "ranlux_luxtst.json"), "w") as f:
        json.dump(kat, f, indent=1)
    arrays = {}
    for c, (lux, seed, k1, lens) in enumerate(CASES):
        script = "1\n%d %d %d 0\n" % (lux, seed, k1)
        for n in lens:
            script += "2\n%d\n" % n
        script += "3\n5\n"
        v = run_ref(script)
        n = sum(lens)
        arrays["case%d_bits" % c] = v[:n].astype(np.int32)
        arrays["case%d_rluxut" % c] = v[n:n + 25].astype(np.int32)
        arrays["case%d_rluxat" % c] = v[n + 25:n + 29].astype(np.int32)
    arrays["cases"] = np.array([(lux, seed, k1) for lux, seed, k1, _ in CASES], np.int64)
    arrays["case_lengths"] = np.array([",".join(map(str, l)) for *_, l in CASES])
    arrays["measinit_ranks"] = np.array(RANKS, np.int32)
    arrays["measinit_seeds"] = np.stack(
        [run_ref("6\n%d\n" % r) for r in RANKS]).astype(np.int32)
    np.savez(os.path.join(HERE, "ranlux_ref.npz"), **arrays)


if __name__ == "__main__":
    main()
