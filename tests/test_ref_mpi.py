"""The reference's sweep under MPI (oracle/_ref/refmpi32_64, oracle/ref/refmpi.f),
which bench.py --workload refsched64 times as the CPU baseline: the driver
runs on 1, 2 and 8 ranks with the rank grids the benchmark uses, reports
one line, and the halo exchanges of the reference routines keep the physics
of every decomposition the same (the plaquette after the same number of
sweeps from per-rank random starts agrees to within its fluctuations).
Skipped where the reference build or mpiexec is absent."""
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "refmpi32_64")
MPIEXEC = "/opt/conda/bin/mpiexec"

pytestmark = pytest.mark.skipif(not (os.access(EXE, os.X_OK) and os.access(MPIEXEC, os.X_OK)),
                                reason="refmpi32_64 or mpiexec absent (make -C oracle/ref mpi)")


def test_reference_sweep_under_mpi_decompositions():
    import bench
    res = bench.cpu_baseline_reference_mpi((16, 16, 16), [1, 2, 8], 0.5)
    assert [r["cores"] for r in res] == [1, 2, 8]
    assert [r["rank_grid"] for r in res] == [[1, 1, 1], [1, 1, 2], [2, 2, 2]]
    for r in res:
        print(r["cores"], r["rank_grid"], r["seconds_per_sweep"], r["plaq0"])
        assert r["seconds_per_sweep"] > 0.0 and r["value"] > 0.0
        assert r["kind"] == "reference" and "halo exchanges" in r["sample"]
        # broken-phase plaquette of the production couplings, all decompositions
        assert 0.85 < r["plaq0"] < 0.92


def test_rank_grid_must_divide_the_lattice():
    import bench
    # 16 ranks need a (2, 2, 4) grid; a 6^3 lattice cannot be split into even blocks
    assert bench.cpu_baseline_reference_mpi((6, 6, 6), [16], 0.1) == []
