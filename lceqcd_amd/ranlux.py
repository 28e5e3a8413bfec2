"""RANLUX, the reference's seeding generator (f/luxury.f), over liblceq.so.

Host code only: usable without a GPU.  Values are float32 as in the
reference build (REAL is single precision there).
"""
import ctypes as C

import numpy as np

from ._lib import check, load


class Ranlux:
    """One generator state, COMMON /LUX1/ + /LUX2/ of f/luxury.h.

    A fresh generator initialises itself with luxury 3 and seed 314159265 at
    its first :meth:`ranlux` call unless :meth:`rluxgo` or :meth:`rluxin` ran
    first (f/luxury.f:31-66)."""

    def __init__(self):
        self._lib = load()
        h = C.c_void_p()
        check(self._lib.lceq_ranlux_create(C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            self._lib.lceq_ranlux_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def ranlux(self, n):
        """RANLUX(RVEC, LENV): n uniforms in (0, 1), float32."""
        out = np.empty(n, np.float32)
        check(self._lib.lceq_ranlux(self._h, out.ctypes.data_as(C.POINTER(C.c_float)), int(n)))
        return out

    def rluxgo(self, lux, ins, k1=0, k2=0):
        """RLUXGO(LUX, INT, K1, K2); raises LceqError where the reference
        prints an error message (the fallback state is still set)."""
        check(self._lib.lceq_rluxgo(self._h, int(lux), int(ins), int(k1), int(k2)))

    def rluxut(self):
        """RLUXUT: the 25-integer restart vector."""
        v = (C.c_int * 25)()
        check(self._lib.lceq_rluxut(self._h, v))
        return np.array(v[:], np.int32)

    def rluxin(self, isdext):
        """RLUXIN: restart from a 25-integer vector."""
        v = np.ascontiguousarray(isdext, np.int32)
        if v.shape != (25,):
            raise ValueError("the restart vector has 25 integers")
        check(self._lib.lceq_rluxin(self._h, v.ctypes.data_as(C.POINTER(C.c_int))))

    def rluxat(self):
        """RLUXAT: (luxury, initial seed, k1, k2) -- the restart point for rluxgo."""
        a = [C.c_int() for _ in range(4)]
        check(self._lib.lceq_rluxat(self._h, *[C.byref(x) for x in a]))
        return tuple(x.value for x in a)


def measinit_seeds(rank=0):
    """The 16 integer seeds MPI rank `rank` derives in f/measinit.f:214-216
    (RLUXGO(3, 314159 + rank, 0, 0), 16 RANLUX numbers, int(10**6 * u))."""
    v = (C.c_int * 16)()
    check(load().lceq_measinit_seeds(int(rank), v))
    return np.array(v[:], np.int32)


def measinit_key(iseed):
    """The Philox key the library derives from 16 measinit seeds
    (lceq_measinit_key): iseed[0] + 2^32 iseed[1]."""
    v = np.ascontiguousarray(np.asarray(iseed, dtype=np.int32))
    if v.shape != (16,):
        raise ValueError("measinit gives 16 seeds")
    return int(load().lceq_measinit_key(v.ctypes.data_as(C.POINTER(C.c_int))))
