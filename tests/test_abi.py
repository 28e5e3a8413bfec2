"""The C ABI library loads on a machine without a GPU and exports every entry
point include/lceq.h declares; argument validation happens before any HIP
call.  No kernel is launched here."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "lceq.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(lceq_[a-z0-9_]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def lib():
    from lceqcd_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from lceqcd_amd import build
        build.build()
    return _lib.load()


def test_every_declared_symbol_is_exported(lib):
    names = header_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from lceqcd_amd import _lib
    # status-returning entry points; the two value-returning ones are bound
    # apart with their own result types
    names = set(header_functions()) - {"lceq_last_error", "lceq_measinit_key"}
    assert names == set(_lib.SIGNATURES), names ^ set(_lib.SIGNATURES)


def test_shape_validation_before_device_use(lib):
    h = C.c_void_p()
    for shape in ((3, 4, 4), (4, 0, 4), (4, 4, 5), (1, 2, 2)):
        rc = lib.lceq_create(*shape, C.byref(h))
        assert rc == -1 and not h.value
        assert b"even" in lib.lceq_last_error()


def test_byte_offset_range_checked_before_device_use(lib):
    """Device loads address a site of a class array by an unsigned 32-bit byte
    offset s * 16 (and stores through a raw buffer of 2^32 - 1 bytes), so
    2^28 or more sites per parity must be refused before any allocation
    (816^3 passes the 32-bit link-count check: 3 V < 2^31)."""
    h = C.c_void_p()
    assert 816 ** 3 * 3 < 2 ** 31 and 816 ** 3 // 2 > 2 ** 28
    rc = lib.lceq_create(816, 816, 816, C.byref(h))
    assert rc == -1 and not h.value
    assert b"2^28" in lib.lceq_last_error()
    dims = (C.c_int * 6)(16, 16, 16, 816, 816, 816)
    b = C.c_void_p()
    rc = lib.lceq_batch_create(2, dims, C.byref(b))
    assert rc == -1 and not b.value
    assert b"replica 1" in lib.lceq_last_error() and b"2^28" in lib.lceq_last_error()
    # exactly 2^28 sites per parity (1024 x 1024 x 512) is refused as well
    rc = lib.lceq_create(1024, 1024, 512, C.byref(h))
    assert rc == -1 and not h.value and b"2^28" in lib.lceq_last_error()


def test_null_handle_is_rejected(lib):
    assert lib.lceq_agauge(None, 1, None) == -1
    assert lib.lceq_renorm(None) == -1
    assert lib.lceq_landau_gauge(None, None, None) == -1


def test_copy_bandwidth_arguments_checked_before_device_use(lib):
    us = C.c_double(-1.0)
    for nbytes, reps in ((0, 5), (24, 5), (1 << 20, 0)):
        assert lib.lceq_copy_bandwidth(nbytes, reps, C.byref(us)) == -1
    assert lib.lceq_copy_bandwidth(1 << 20, 5, None) == -1
    assert us.value == -1.0


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from lceqcd_amd import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "liblceq.so"))
    with pytest.raises(_lib.LceqError, match="not built"):
        _lib.load()


def test_coupling_maps():
    from lceqcd_amd import couplings_from_trace_convention, couplings_from_xy
    assert couplings_from_trace_convention(9.0, 0.35, -0.9, 0.1) == (9.0, 0.175, -0.45, 0.025)
    ba, b2, b4 = couplings_from_xy(9.0, 0.1, 0.5)
    assert ba == 9.0 and abs(b4 - 0.9) < 1e-15 and abs(b2 - 25.0856) < 1e-3


def test_stream_bandwidth_arguments_checked_before_device_use(lib):
    us = C.c_double(-1.0)
    for streams, n, reps in ((5, 1 << 16, 5), (30, 0, 5), (16, 1 << 16, 0), (32, 1 << 16, 5)):
        assert lib.lceq_stream_bandwidth(streams, n, reps, C.byref(us)) == -1
    assert lib.lceq_stream_bandwidth(30, 1 << 16, 5, None) == -1
    assert us.value == -1.0
