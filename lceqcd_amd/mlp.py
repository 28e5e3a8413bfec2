"""Learned Higgs proposal (liblceq's lceq_mlp_*; no counterpart in the
reference).  A 10 -> 64 -> 64 -> 6 SiLU network per site proposes a Gaussian
move of a(x) in the frame of the hopping force F(x); an FP64
Metropolis-Hastings test with both proposal densities keeps the chain exact.
Training (maximum likelihood on moves of the plain Metropolis update) and
inference (bf16 MFMA) run on the GPU."""
import ctypes as C

import numpy as np

from ._lib import check, load

NPARAM = 5254
NSTATE = 5274
NIN, NOUT = 10, 6
FP = C.POINTER(C.c_float)


class Mlp:
    def __init__(self, seed=1):
        self._lib = load()
        h = C.c_void_p()
        check(self._lib.lceq_mlp_create(int(seed), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self._lib.lceq_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_state(self):
        st = np.zeros(NSTATE, np.float32)
        check(self._lib.lceq_mlp_get_state(self.h, st.ctypes.data_as(FP), NSTATE))
        return st

    def set_state(self, state):
        st = np.ascontiguousarray(state, np.float32)
        if st.shape != (NSTATE,):
            raise ValueError("the state has %d floats" % NSTATE)
        check(self._lib.lceq_mlp_set_state(self.h, st.ctypes.data_as(FP), NSTATE))

    def collect(self, lattice, nsteps=8, stride=1):
        """Training records from the lattice's current configuration."""
        check(self._lib.lceq_mlp_collect(lattice.h, self.h, int(nsteps), int(stride)))

    def records(self):
        n = C.c_int64()
        check(self._lib.lceq_mlp_records(self.h, C.byref(n)))
        return n.value

    def clear(self):
        check(self._lib.lceq_mlp_clear(self.h))

    def train(self, steps, batch=4096, lr=1e-3, seed=1, fit_norm=True):
        """Adam steps; returns the per-step mean negative log-likelihood."""
        loss = np.zeros(max(int(steps), 1))
        check(self._lib.lceq_mlp_train(self.h, int(steps), int(batch), float(lr), int(seed),
                                       1 if fit_norm else 0,
                                       loss.ctypes.data_as(C.POINTER(C.c_double))))
        return loss[:steps]

    def set_sites_per_wave(self, sites=0):
        """Launch shape of the learned Higgs update: 64 (default, also 0) or
        32 sites per wave (fields do not depend on it)."""
        check(self._lib.lceq_mlp_set_sites_per_wave(self.h, int(sites)))

    def eval(self, x):
        """The inference network (bf16 MFMA) on rows of the 10 raw inputs."""
        xx = np.ascontiguousarray(x, np.float32).reshape(-1, NIN)
        y = np.zeros((len(xx), NOUT), np.float32)
        check(self._lib.lceq_mlp_eval(self.h, xx.ctypes.data_as(FP), len(xx), y.ctypes.data_as(FP)))
        return y
