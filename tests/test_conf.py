"""Configuration files (f/confsave.f, f/confget.f oldconf) against the files
the reference itself wrote and read back (tests/golden/conf_6x4x8.npz, made
by tests/golden/make_conf_golden.py with the reference's confsave/oldconf in
both its 8-byte-real and as-shipped 4-byte builds), and against the
reference's reader run live on files written here when oracle/_ref is built.
The file functions are host code and run without a GPU; the device lattice
round trip is marked gpu."""
import os

import numpy as np
import pytest

from helpers import golden, rand_conf

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def conf():
    from lceqcd_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from lceqcd_amd import build
        build.build()
    from lceqcd_amd import conf as c
    return c


@pytest.mark.parametrize("prec", [64, 32])
def test_writer_bytes_equal_reference_confsave(conf, tmp_path, prec):
    z = golden("conf_6x4x8")
    shape = tuple(int(v) for v in z["shape"])
    up, ap = tmp_path / "u", tmp_path / "a"
    conf.conf_write(shape, z["u_in"], z["a_in"], up, ap, real_bytes=prec // 8)
    assert np.array_equal(np.fromfile(up, np.uint8), z["u_bytes%d" % prec])
    assert np.array_equal(np.fromfile(ap, np.uint8), z["a_bytes%d" % prec])


@pytest.mark.parametrize("prec", [64, 32])
def test_reader_on_reference_files(conf, tmp_path, prec):
    z = golden("conf_6x4x8")
    shape = tuple(int(v) for v in z["shape"])
    up, ap = tmp_path / "u", tmp_path / "a"
    z["u_bytes%d" % prec].tofile(up)
    z["a_bytes%d" % prec].tofile(ap)
    u, a = conf.conf_read(shape, up, ap, real_bytes=prec // 8)
    # what the reference's own oldconf read from the same files
    assert np.array_equal(u, z["ref_u%d" % prec])
    assert np.array_equal(a, z["ref_a%d" % prec])


def test_filenames_match_reference(conf):
    z = golden("conf_6x4x8")
    assert conf.conf_filename("u", 8, 9.0, 0.1, 0.5, 123) == str(z["name_u"])
    assert conf.conf_filename("a", 8, 9.0, 0.1, 0.5, 123) == str(z["name_a"])
    # f/measinit.f:246-250 spells the name oldconf opens in lower case
    assert conf.conf_filename("u", 8, 9.0, 0.1, 0.5, 123, saved=False) == \
        "u_conf_s08_be09000_x01000_y05000_u000123"


def test_filename_edit_descriptors(conf):
    # I5.5 of a negative value needs six characters: asterisks
    assert "_y*****_" in conf.conf_filename("u", 8, 9.0, 0.1, -0.45, 1)
    # I2.2 overflows from 100 on; NINT rounds half away from zero
    assert conf.conf_filename("a", 100, 9.0005, 0.00005, 0.0, 7).startswith(
        "a_conf_s**_BE09001_x00001_y00000_U000007")


@pytest.mark.parametrize("shape", [(2, 2, 2), (4, 6, 2), (8, 4, 6)])
@pytest.mark.parametrize("rb", [8, 4])
def test_round_trip(conf, tmp_path, shape, rb):
    u, a = rand_conf(shape, 11, 2.0)
    conf.conf_write(shape, u, a, tmp_path / "u", tmp_path / "a", real_bytes=rb)
    u2, a2 = conf.conf_read(shape, tmp_path / "u", tmp_path / "a", real_bytes=rb)
    cast = (lambda x: x) if rb == 8 else (lambda x: x.astype(np.float32).astype(np.float64))
    assert np.array_equal(u2, cast(u))
    assert np.array_equal(a2, cast(a))
    lx, ly, lz = shape
    assert os.path.getsize(tmp_path / "u") == 6 * lz * 16 * 2 * lx * ly // 2
    assert os.path.getsize(tmp_path / "a") == 2 * lz * 8 * 3 * lx * ly // 2


def test_single_field_and_errors(conf, tmp_path):
    from lceqcd_amd import LceqError
    shape = (4, 4, 4)
    u, a = rand_conf(shape, 3)
    conf.conf_write(shape, None, a, None, tmp_path / "a")
    assert not (tmp_path / "u").exists()
    _, a2 = conf.conf_read(shape, None, tmp_path / "a")
    assert np.array_equal(a2, a)
    with pytest.raises(LceqError, match="cannot open"):
        conf.conf_read(shape, tmp_path / "missing", None)
    # a file for a smaller lattice is too short for this one
    conf.conf_write((4, 4, 2), rand_conf((4, 4, 2), 1)[0], None, tmp_path / "small", None)
    with pytest.raises(LceqError, match="needs"):
        conf.conf_read(shape, tmp_path / "small", None)
    with pytest.raises(LceqError, match="real_bytes"):
        conf.conf_write(shape, u, None, tmp_path / "u", None, real_bytes=2)
    with pytest.raises(LceqError, match="even"):
        conf.conf_write((3, 4, 4), u, None, tmp_path / "u", None)


def test_find_conf_file_tolerates_case(conf, tmp_path):
    shape = (4, 4, 4)
    u, a = rand_conf(shape, 3)
    name = conf.conf_filename("u", 4, 9.0, 0.1, 0.5, 10)  # what confsave writes
    conf.conf_write(shape, u, None, tmp_path / name, None)
    assert conf.find_conf_file(str(tmp_path), "u", 4, 9.0, 0.1, 0.5, 10) == str(tmp_path / name)
    assert conf.find_conf_file(str(tmp_path), "a", 4, 9.0, 0.1, 0.5, 10) is None


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "refharness64")),
                    reason="oracle/_ref not built")
@pytest.mark.parametrize("prec", [64, 32])
def test_reference_oldconf_reads_our_files(conf, tmp_path, prec):
    from oracle import refrun
    shape = (4, 6, 4)
    u, a = rand_conf(shape, 21, 1.5)
    conf.conf_write(shape, u, a, tmp_path / "u_in", tmp_path / "a_in", real_bytes=prec // 8)
    recs = refrun.run("oldconf", shape, np.zeros_like(u), np.zeros_like(a), prec=prec,
                      workdir=str(tmp_path))
    cast = (lambda x: x) if prec == 64 else (lambda x: x.astype(np.float32).astype(np.float64))
    assert np.array_equal(recs[0].reshape(u.shape), cast(u))
    assert np.array_equal(recs[1].reshape(a.shape), cast(a))


@pytest.mark.gpu
@pytest.mark.parametrize("rb", [8, 4])
def test_lattice_confsave_oldconf(conf, tmp_path, rb):
    """confsave of a device lattice, oldconf into a fresh one: the fields
    come back (rounded to float for the 4-byte format) and a sweep from both
    gives the same result."""
    import lceqcd_amd
    shape = (8, 6, 4)
    print("lattice confsave/oldconf", shape, rb, flush=True)
    L = lceqcd_amd.Lattice(*shape)
    L.hotstart(5, 0.7)
    if rb == 4:  # start from float-representable fields so the copy is exact
        L.put_links(L.get_links().astype(np.float32).astype(np.float64))
        L.put_higgs(L.get_higgs().astype(np.float32).astype(np.float64))
    L.confsave(tmp_path / "u", tmp_path / "a", real_bytes=rb)
    M = lceqcd_amd.Lattice(*shape)
    M.oldconf(tmp_path / "u", tmp_path / "a", real_bytes=rb)
    assert np.array_equal(M.get_links(), L.get_links())
    assert np.array_equal(M.get_higgs(), L.get_higgs())
    for X in (L, M):
        X.set_couplings(9.0, 0.175, -0.45, 0.025)
        X.rng_philox(3, 0)
        X.sweep(0, 2, n_or=2, delta=0.4, measflag=0)
    assert np.array_equal(M.get_links(), L.get_links())
    L.close()
    M.close()
