"""lceq_check_fields_static (the field guard with its scratch in a fixed
device array, for callers that cannot allocate) against lceq_check_fields,
through the C ABI: equal outputs and status on two calls in a row, the
second on a smaller lattice (the fixed array keeps the first call's
partials, which the second must not read), on clean fields and on fields
with planted non-unitary and non-finite values; and a lattice beyond the
fixed maximum is refused with an error."""
import ctypes as C

import numpy as np
import pytest

from helpers import rand_conf

pytestmark = pytest.mark.gpu


def _guard(lib, fn, L, tol):
    dev, bad = C.c_double(-1.0), C.c_int64(-1)
    f = getattr(lib, fn)
    f.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    f.restype = C.c_int
    rc = f(L.h, tol, C.byref(dev), C.byref(bad))
    return rc, dev.value, bad.value


def _lattice(shape, seed, plant):
    import lceqcd_amd
    L = lceqcd_amd.Lattice(*shape)
    u, a = rand_conf(shape, seed, 0.5)
    if plant == "scaled":
        u[1, 0, 3] *= 1.0 + 3e-6
    elif plant == "nan":
        u[2, 1, 0, 2] = np.nan
        a[0, 1, 1] = np.inf
    L.put_links(u)
    L.put_higgs(a)
    return L


@pytest.mark.parametrize("plant", ["none", "scaled", "nan"])
def test_static_guard_equals_allocating_guard(plant):
    from lceqcd_amd import _lib
    lib = _lib.load()
    # first a lattice with many blocks, then a smaller one
    for shape, seed in (((48, 48, 64), 1), ((8, 6, 4), 2)):
        print("static guard", plant, shape, flush=True)
        L = _lattice(shape, seed, plant)
        for tol in (-1.0, 1e-12):
            got = _guard(lib, "lceq_check_fields_static", L, tol)
            ref = _guard(lib, "lceq_check_fields", L, tol)
            print("  tol %g: static %s, allocating %s" % (tol, got, ref), flush=True)
            assert got == ref
        if plant == "none":
            assert ref[0] == 0 and ref[2] == 0
        elif plant == "scaled":
            assert ref[0] == -6 and ref[1] > 1e-6
        else:
            assert ref[0] == -6 and ref[2] == 2
        L.close()


def test_static_guard_refuses_lattices_beyond_the_array():
    import lceqcd_amd
    from lceqcd_amd import _lib
    lib = _lib.load()
    # 2^24 sites per parity fit; 256 x 256 x 520 has 2^24 + 2^16 x 4
    L = lceqcd_amd.Lattice(256, 256, 520)
    rc, dev, bad = _guard(lib, "lceq_check_fields_static", L, -1.0)
    assert rc == -1 and dev == -1.0 and bad == -1
    assert b"fixed maximum" in lib.lceq_last_error()
    L.close()
