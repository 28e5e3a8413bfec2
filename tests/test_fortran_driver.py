"""The Fortran host driver (lceqcd_amd/fortran/lceq_main.f90) over the C ABI:
builds with amdflang here; on the GPU it runs the f/main.f schedule on an 8^3
lattice, writes the reference's output files, and its action records equal
the CPU restatement's run of the same schedule and stream."""
import glob
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "lceqcd_amd", "fortran")
FC = "/opt/rocm/bin/amdflang"

# the reference's 20-line inputfile layout (runs/inputfile), at 8^3
INPUTFILE = """08 08 08      !nx/ny/nz global lattice size
01 01 01      !dimproc
001           !nproc
{init}            ! init
0             !rand
0             !save
000000        !confstart
 9.000        !beta
0.1000        !x
0.5000        !y
1.7000        !lambda overrelaxation parameter
1.000000      !gauge transformation parameter
0.100E-09     !limit_or
0.100E-04     !limit_it
001           !inorm
0             !nsweep_flag
240           !ntime
{nsweep}      !nsweeps
00001         !imeas
000           !imode
"""


@pytest.mark.skipif(not os.path.exists(FC), reason="amdflang not installed")
def test_driver_compiles(tmp_path):
    r = subprocess.run([FC, "-O2", "-J", str(tmp_path), os.path.join(FDIR, "lceq_c.f90"),
                        os.path.join(FDIR, "lceq_main.f90"), "-c"], cwd=tmp_path,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_driver_runs_reference_schedule(tmp_path, oracle_lib):
    exe = os.path.join(FDIR, "lceq_main")
    assert os.path.exists(exe), "lceq_main not built (python -m lceqcd_amd.build)"
    (tmp_path / "inputfile").write_text(INPUTFILE.format(init=0, nsweep="00002"))
    print("running lceq_main 8^3 nsweep=2 ntherm=0", flush=True)
    r = subprocess.run([exe, "inputfile", "4242", "0"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    act = glob.glob(str(tmp_path / "action_s08_be09000_x01000_y05000_u000000"))
    assert act
    lines = open(act[0]).read().splitlines()
    assert lines[0].startswith("# 08 08 08") and len(lines) == 5
    rec = np.array([[float(x) for x in l.split()[1:]] for l in lines[3:]])
    corr = open(glob.glob(str(tmp_path / "correl_s08_*"))[0]).read().splitlines()
    assert len(corr) == 3 + 2 * (1 + 8)
    job = open(glob.glob(str(tmp_path / "joblog_s08_*"))[0]).read()
    assert "Gauge  acceptance" in job and "number of iteration for or. gauge fixing" in job
    # the same schedule on the CPU restatement (couplings of f/paraminit.f:150-153)
    from lceqcd_amd import couplings_from_xy
    O = oracle_lib.OracleLattice(8, 8, 8)
    O.set_couplings(9.0, *couplings_from_xy(9.0, 0.1, 0.5))
    O.hotstart(4242, 0.5)
    O.philox(4242, 0)
    exp = []
    for it in range(2):
        O.ahiggs2(1)
        O.ahiggs3(1)
        O.agauge(1)
        O.renorm()
        ob = O.observables()
        exp.append([ob["plaq0"], ob["a0kin"], ob["a02"], ob["a04"]])
        O.set_gauge_fixing(1.7, 1.0, 1e-10)
        O.landau_gauge()
        O.correlator()
    np.testing.assert_allclose(rec, np.array(exp), rtol=1e-12)
    # the configuration saved at the end (f/finalize.f:43) is the final state,
    # in the reference's file format and upper-case name
    from lceqcd_amd.conf import conf_read
    u, a = conf_read((8, 8, 8), tmp_path / "u_conf_s08_BE09000_x01000_y05000_U000002",
                     tmp_path / "a_conf_s08_BE09000_x01000_y05000_U000002")
    np.testing.assert_allclose(u, O.get_links(), rtol=0, atol=1e-11)
    np.testing.assert_allclose(a, O.get_higgs(), rtol=0, atol=1e-11)


@pytest.mark.gpu
def test_driver_restarts_from_saved_configuration(tmp_path):
    """init = 3 reads the configuration a previous run saved (under the
    upper-case name, which the lower-case lookup of the reference misses) and
    a zero-sweep run saves it again unchanged."""
    exe = os.path.join(FDIR, "lceq_main")
    from lceqcd_amd.conf import conf_read, conf_write
    from helpers import rand_conf
    u, a = rand_conf((8, 8, 8), 9, 1.0)
    conf_write((8, 8, 8), u, a, tmp_path / "u_conf_s08_BE09000_x01000_y05000_U000000",
               tmp_path / "a_conf_s08_BE09000_x01000_y05000_U000000")
    (tmp_path / "inputfile").write_text(INPUTFILE.format(init=3, nsweep="00000"))
    print("running lceq_main 8^3 init=3 nsweep=0", flush=True)
    r = subprocess.run([exe, "inputfile", "1", "0"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    # nsweep = 0: saved under confstart + nsweep = 0 again, same fields
    u2, a2 = conf_read((8, 8, 8), tmp_path / "u_conf_s08_BE09000_x01000_y05000_U000000",
                       tmp_path / "a_conf_s08_BE09000_x01000_y05000_U000000")
    assert np.array_equal(u2, u) and np.array_equal(a2, a)


@pytest.mark.gpu
@pytest.mark.parametrize("corrflag", [2, 3])
def test_driver_polyakov_and_scalarcorr_outputs(tmp_path, oracle_lib, corrflag):
    """corrflag 2 (polyakov + scalarcorr, f/main.f:68-71) and 3 (gauge fixing,
    the correlators and polyakov, f/main.f:73-78): the polyak_ and corra2_
    files carry the reference's records (f/polyakov.f:110-117,
    f/scalarcorr.f:106-113) and their values equal the CPU restatement's run
    of the same schedule and stream."""
    exe = os.path.join(FDIR, "lceq_main")
    (tmp_path / "inputfile").write_text(INPUTFILE.format(init=0, nsweep="00002"))
    print("running lceq_main 8^3 nsweep=2 ntherm=0 corrflag", corrflag, flush=True)
    r = subprocess.run([exe, "inputfile", "4242", "0", "8", str(corrflag)], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr

    def records(prefix):
        lines = open(glob.glob(str(tmp_path / (prefix + "_s08_*")))[0]).read().splitlines()
        body = lines[3:]
        return [(int(body[k * 10]), [float(v) for v in body[k * 10 + 1:k * 10 + 10]])
                for k in range(len(body) // 10)]

    pol = records("polyak")
    ca2 = records("corra2")
    assert [p[0] for p in pol] == [1, 2]
    assert len(ca2) == (2 if corrflag == 2 else 0)
    from lceqcd_amd import couplings_from_xy
    O = oracle_lib.OracleLattice(8, 8, 8)
    O.set_couplings(9.0, *couplings_from_xy(9.0, 0.1, 0.5))
    O.hotstart(4242, 0.5)
    O.philox(4242, 0)
    for it in range(2):
        O.ahiggs2(1)
        O.ahiggs3(1)
        O.agauge(1)
        O.renorm()
        if corrflag == 3:
            O.set_gauge_fixing(1.7, 1.0, 1e-10)
            O.landau_gauge()
        m, c = O.polyakov()
        np.testing.assert_allclose(pol[it][1], [m] + list(c), rtol=1e-12, atol=1e-14)
        if corrflag == 2:
            m, c = O.scalarcorr()
            np.testing.assert_allclose(ca2[it][1], [m] + list(c), rtol=1e-12, atol=1e-14)


@pytest.mark.gpu
def test_driver_seeds_the_device_streams_from_ranlux(tmp_path, oracle_lib):
    """Without a numeric seed the driver seeds its device streams as
    f/measinit.f:214-217 seeds the reference's random_number on rank 0
    (lceq_rng_measinit): the joblog lists the 16 seeds, which equal the ones
    the reference's own f/luxury.f built here gives (tests/golden/
    ranlux_ref.npz), and the run equals the CPU restatement's run of the
    same schedule with the Philox key iseed(1) + 2^32 iseed(2)."""
    exe = os.path.join(FDIR, "lceq_main")
    (tmp_path / "inputfile").write_text(INPUTFILE.format(init=0, nsweep="00002"))
    print("running lceq_main 8^3 nsweep=2 ntherm=0 with RANLUX seeding", flush=True)
    r = subprocess.run([exe, "inputfile", "ranlux", "0"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    job = open(glob.glob(str(tmp_path / "joblog_s08_*"))[0]).read().splitlines()
    seeds = [int(v) for v in next(l for l in job if l.startswith("# iseed")).split()[2:]]
    key = int(next(l for l in job if l.startswith("# philox key")).split()[-1])
    from helpers import golden
    z = golden("ranlux_ref")
    want = z["measinit_seeds"][list(z["measinit_ranks"]).index(0)]
    assert seeds == [int(v) for v in want]
    assert key == seeds[0] + (seeds[1] << 32)
    lines = open(glob.glob(str(tmp_path / "action_s08_*"))[0]).read().splitlines()
    rec = np.array([[float(x) for x in l.split()[1:]] for l in lines[3:]])
    from lceqcd_amd import couplings_from_xy
    O = oracle_lib.OracleLattice(8, 8, 8)
    O.set_couplings(9.0, *couplings_from_xy(9.0, 0.1, 0.5))
    O.hotstart(key, 0.5)
    O.philox(key, 0)
    exp = []
    for it in range(2):
        O.ahiggs2(1)
        O.ahiggs3(1)
        O.agauge(1)
        O.renorm()
        ob = O.observables()
        exp.append([ob["plaq0"], ob["a0kin"], ob["a02"], ob["a04"]])
        O.set_gauge_fixing(1.7, 1.0, 1e-10)
        O.landau_gauge()
        O.correlator()
    np.testing.assert_allclose(rec, np.array(exp), rtol=1e-12)


@pytest.mark.gpu
def test_lattice_rng_measinit_matches_explicit_key():
    """lceq_rng_measinit(rank) returns the rank's seeds and keys the streams
    with measinit_key(seeds): the same chain as lceq_rng_philox(key, 0)."""
    import lceqcd_amd
    out = []
    for how in ("measinit", "explicit"):
        L = lceqcd_amd.Lattice(8, 8, 8)
        L.set_couplings(9.0, 0.175, -0.45, 0.025)
        if how == "measinit":
            s = L.rng_measinit(5)
            assert (s == lceqcd_amd.measinit_seeds(5)).all()
        else:
            L.rng_philox(lceqcd_amd.measinit_key(lceqcd_amd.measinit_seeds(5)), 0)
        L.unitconf()
        L.sweep(0, 3, n_or=4, delta=0.4, measflag=1)
        out.append((L.get_links(), L.get_higgs()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
