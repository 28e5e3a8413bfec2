"""Configuration files in the reference's on-disk format (f/confsave.f,
f/confget.f oldconf) over liblceq.so.  Host code: no GPU needed for
conf_write / conf_read / conf_filename."""
import ctypes as C
import os

import numpy as np

from ._lib import check, load


def _dims(shape):
    d = (C.c_int * 3)(*[int(x) for x in shape])
    return d


def _path(p):
    return None if p is None else os.fsencode(p)


def conf_write(shape, u, a, u_path, a_path, real_bytes=8):
    """Write links u (shape (3, 2, N, 4), reference layout) and Higgs a
    (shape (2, N, 3)) as the reference's confsave does; either path may be
    None.  real_bytes = 8 for a build with 8-byte reals, 4 for the
    as-shipped single-precision build (half of every record is padding)."""
    dp = C.POINTER(C.c_double)
    uu = None if u is None else np.ascontiguousarray(u, np.float64)
    aa = None if a is None else np.ascontiguousarray(a, np.float64)
    check(load().lceq_conf_write(_dims(shape), None if uu is None else uu.ctypes.data_as(dp),
                                 None if aa is None else aa.ctypes.data_as(dp), _path(u_path),
                                 _path(a_path), int(real_bytes)))


def conf_read(shape, u_path, a_path, real_bytes=8):
    """Read the files back: returns (u, a) in the layout of conf_write (None
    for a field whose path is None)."""
    lx, ly, lz = shape
    n = lx * ly * lz // 2
    dp = C.POINTER(C.c_double)
    u = np.zeros((3, 2, n, 4)) if u_path is not None else None
    a = np.zeros((2, n, 3)) if a_path is not None else None
    check(load().lceq_conf_read(_dims(shape), None if u is None else u.ctypes.data_as(dp),
                                None if a is None else a.ctypes.data_as(dp), _path(u_path),
                                _path(a_path), int(real_bytes)))
    return u, a


def conf_filename(field, lz, beta, x, y, sweep, saved=True):
    """The name confsave writes (saved=True: '_BE', '_U', f/confsave.f:249-253)
    or the one oldconf opens (saved=False: '_be', '_u', f/measinit.f:246-250);
    field is 'u' or 'a'."""
    buf = C.create_string_buffer(64)
    check(load().lceq_conf_filename(1 if field == "a" else 0, 1 if saved else 0, int(lz),
                                    float(beta), float(x), float(y), int(sweep), buf, 64))
    return buf.value.decode()


def find_conf_file(directory, field, lz, beta, x, y, sweep):
    """The file oldconf would open, tolerating the reference's case mismatch:
    the lower-case name of f/measinit.f first, then the upper-case name that
    f/confsave.f actually writes.  Returns None when neither exists."""
    for saved in (False, True):
        p = os.path.join(directory, conf_filename(field, lz, beta, x, y, sweep, saved))
        if os.path.exists(p):
            return p
    return None
