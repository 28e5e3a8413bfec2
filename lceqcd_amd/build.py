"""Build liblceq.so in-tree with hipcc for gfx950.

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
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "--offload-arch=gfx950", "-fPIC", "-shared", "-std=c++17",
    # keep the compiler away from packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # a * b + c may become one fused multiply-add (v_fma_f64, one rounding):
    # 5% faster sweeps than with every operation rounded as written, at a
    # difference from the CPU restatement (no FMA, as the reference's x86
    # build) of a few ulp per update
    "-ffp-contract=fast",
    "-Wall", "-Wno-unknown-warning-option",
]


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
    """Compile every HIP source into lceqcd_amd/liblceq.so (always rebuilds)."""
    tmp = OUT + ".tmp"
    cmd = [HIPCC] + FLAGS + SRC + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building liblceq.so")
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(tmp, OUT)
    if os.path.exists(FC):
        build_fortran()
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
