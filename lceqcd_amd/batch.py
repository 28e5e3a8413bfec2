"""Batched replicas: many independent lattices of uneven sizes and couplings
swept together by liblceq.so (include/lceq.h, lceq_batch_*).  Replica r
follows exactly the chain of a single :class:`lceqcd_amd.Lattice` with the
same seed under ``sweep(SWEEP_KP_OR_METRO, ...)``."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check


def _dp(x):
    return x.ctypes.data_as(C.POINTER(C.c_double))


class Batch:
    def __init__(self, shapes):
        self._L = _lib.load()
        self.shapes = [tuple(int(v) for v in s) for s in shapes]
        dims = np.ascontiguousarray(np.array(self.shapes, dtype=np.int32).ravel())
        h = C.c_void_p()
        check(self._L.lceq_batch_create(len(self.shapes), dims.ctypes.data_as(C.POINTER(C.c_int)),
                                        C.byref(h)))
        self.h = h
        self.nrep = len(self.shapes)
        self.sites = [s[0] * s[1] * s[2] for s in self.shapes]
        self.total_sites = sum(self.sites)

    def close(self):
        if getattr(self, "h", None):
            self._L.lceq_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _u64(self, seeds):
        a = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint64))
        assert a.shape == (self.nrep,)
        return a, a.ctypes.data_as(C.POINTER(C.c_uint64))

    def set_couplings(self, coup):
        """coup: (nrep, 4) beta, beta_a, beta_2, beta_4 (reference normalisation)."""
        c = np.ascontiguousarray(np.asarray(coup, dtype=np.float64).reshape(self.nrep, 4))
        check(self._L.lceq_batch_set_couplings(self.h, _dp(c)))

    def rng_philox(self, seeds):
        a, p = self._u64(seeds)
        check(self._L.lceq_batch_rng_philox(self.h, p))

    def hotstart(self, seeds, sigma):
        a, p = self._u64(seeds)
        check(self._L.lceq_batch_hotstart(self.h, p, sigma))

    def _field(self, r, arr, ncomp):
        n = self.sites[r] // 2
        shape = (3, 2, n, 4) if ncomp == 4 else (2, n, 3)
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != shape:
            raise ValueError("replica %d: expected shape %s, got %s" % (r, shape, a.shape))
        return a, n

    def put_links(self, r, u):
        u, n = self._field(r, u, 4)
        check(self._L.lceq_batch_put_links(self.h, r, _dp(u), n))

    def get_links(self, r):
        n = self.sites[r] // 2
        u = np.empty((3, 2, n, 4))
        check(self._L.lceq_batch_get_links(self.h, r, _dp(u), n))
        return u

    def put_higgs(self, r, a):
        a, n = self._field(r, a, 3)
        check(self._L.lceq_batch_put_higgs(self.h, r, _dp(a), n))

    def get_higgs(self, r):
        n = self.sites[r] // 2
        a = np.empty((2, n, 3))
        check(self._L.lceq_batch_get_higgs(self.h, r, _dp(a), n))
        return a

    def sweep(self, nsweep, n_or=4, delta=None, measflag=1, graph=True):
        d = None
        if delta is not None:
            d = np.ascontiguousarray(np.broadcast_to(np.asarray(delta, dtype=np.float64),
                                                     (self.nrep,)))
        check(self._L.lceq_batch_sweep(self.h, nsweep, n_or, None if d is None else _dp(d),
                                       measflag, 1 if graph else 0))

    def sweep_mlp(self, mlp, nsweep, n_or=4, measflag=1, graph=True):
        """nsweep sweeps of 1 KP + n_or OR + the learned Higgs proposal of
        ``mlp`` (a :class:`lceqcd_amd.Mlp`) for every replica."""
        check(self._L.lceq_batch_sweep_mlp(self.h, mlp.h, int(nsweep), int(n_or), int(measflag),
                                           1 if graph else 0))

    def mlp_collect(self, mlp, nsteps=8, stride=1):
        """Training records of ``mlp`` from every replica at its own couplings."""
        check(self._L.lceq_batch_mlp_collect(self.h, mlp.h, int(nsteps), int(stride)))

    def sync(self):
        check(self._L.lceq_batch_sync(self.h))

    def history(self, n):
        """(n, nrep, 16) rows of the last n sweeps (columns as Lattice.history)."""
        out = np.zeros((n, self.nrep, 16))
        check(self._L.lceq_batch_history(self.h, _dp(out), n))
        return out

    def observables(self):
        out = np.zeros((self.nrep, 16))
        check(self._L.lceq_batch_observables(self.h, _dp(out)))
        return out

    def set_shape(self, threads_per_block=256, xcd_order=True):
        """Launch shape of the update kernels (same fields for every shape)."""
        check(self._L.lceq_batch_set_shape(self.h, int(threads_per_block), 1 if xcd_order else 0))

    def get_shape(self):
        t, x = C.c_int(), C.c_int()
        check(self._L.lceq_batch_get_shape(self.h, C.byref(t), C.byref(x)))
        return t.value, bool(x.value)

    def set_group_bytes(self, group_bytes):
        """Sweep the batch in runs of consecutive replicas of about
        ``group_bytes`` of fields each (0: one group); the chains do not
        depend on it."""
        check(self._L.lceq_batch_set_group_bytes(self.h, int(group_bytes)))

    def groups(self):
        """(group_bytes, [first replica of each group])."""
        gb, ng = C.c_int64(), C.c_int()
        check(self._L.lceq_batch_get_groups(self.h, C.byref(gb), C.byref(ng), None, 0))
        first = (C.c_int * max(1, ng.value))()
        check(self._L.lceq_batch_get_groups(self.h, None, None, first, ng.value))
        return gb.value, [int(first[g]) for g in range(ng.value)]

    def autotune(self, n_or=4, trials=3):
        """Time every launch shape on this batch and keep the fastest (the
        chains are left where they were).  Returns {(threads, xcd): ms/sweep}."""
        ms = np.zeros(8)
        check(self._L.lceq_batch_autotune(self.h, int(n_or), int(trials), _dp(ms)))
        return {((64, 128, 256, 512)[i // 2], bool(i % 2)): float(ms[i]) for i in range(8)}

