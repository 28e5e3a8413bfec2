"""Build liblceq.so in-tree with hipcc for gfx950, and beside it the
test-only liblceq_oracleforms.so.

    python -m lceqcd_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = [os.path.join(HERE, "csrc", f) for f in ("lceq_kernels.hip", "lceq_api.hip", "lceq_batch.hip",
                                                   "lceq_ranlux.hip", "lceq_conf.hip",
                                                   "lceq_mlp.hip", "lceq_static.hip")]
OUT = os.path.join(HERE, "liblceq.so")
# Test-only build of the same sources: every operation rounded as written
# (no contraction) and 1/sqrt, cos(2 pi r) in the CPU restatement's forms
# (LCEQ_ORACLE_FORMS), for chains compared with the restatement bit for bit.
# Nothing in the package loads it unless asked (_lib.load_oracle_forms).
OUT_ORACLE_FORMS = os.path.join(HERE, "liblceq_oracleforms.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "--offload-arch=gfx950", "-fPIC", "-shared", "-std=c++17",
    # keep the compiler away from packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unknown-warning-option",
]
# a * b + c inside one expression may become one fused multiply-add
# (v_fma_f64, one rounding), at a difference from the CPU restatement (no FMA,
# as the reference's x86 build) of a few ulp per update.  Products are never
# fused across statements (=on, not =fast): the rounding of every expression
# is then fixed by the source, whatever kernel the site code is inlined into,
# so the batch kernels and the single-lattice kernels give the same bits.
PRODUCT_FLAGS = ["-ffp-contract=on"]
ORACLE_FORMS_FLAGS = ["-ffp-contract=off", "-DLCEQ_ORACLE_FORMS"]


FORTRAN = os.path.join(HERE, "fortran")
FC = os.environ.get("FC_LCEQ", "/opt/rocm/bin/amdflang")


def build_fortran():
    """The iso_c_binding module and the Fortran host driver (needs amdflang).
    Output: lceqcd_amd/fortran/lceq_main, linked against ../liblceq.so."""
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        cmd = [FC, "-O2", "-J", td, os.path.join(FORTRAN, "lceq_c.f90"),
               os.path.join(FORTRAN, "lceq_main.f90"), "-L", HERE, "-llceq",
               "-Wl,-rpath,$ORIGIN/..", "-o", os.path.join(FORTRAN, "lceq_main")]
        r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("amdflang failed building the Fortran driver")
    return os.path.join(FORTRAN, "lceq_main")


def build(verbose=False):
    """Compile every HIP source into lceqcd_amd/liblceq.so and the test-only
    liblceq_oracleforms.so (always rebuilds; the two compiles run side by
    side)."""
    jobs = []
    for out, extra in ((OUT, PRODUCT_FLAGS), (OUT_ORACLE_FORMS, ORACLE_FORMS_FLAGS)):
        tmp = out + ".tmp"
        cmd = [HIPCC] + FLAGS + extra + SRC + ["-o", tmp]
        jobs.append((out, tmp, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                                stderr=subprocess.PIPE, text=True)))
    failed = []
    for out, tmp, proc in jobs:
        so, se = proc.communicate()
        if proc.returncode != 0:
            sys.stderr.write(so + se)
            failed.append(os.path.basename(out))
            if os.path.exists(tmp):
                os.remove(tmp)
            continue
        if verbose:
            sys.stderr.write(se)
        os.replace(tmp, out)
    if failed:
        raise RuntimeError("hipcc failed building " + ", ".join(failed))
    if os.path.exists(FC):
        build_fortran()
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
