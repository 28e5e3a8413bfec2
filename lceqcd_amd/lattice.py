"""Host-side mirror of the reference's update and measurement routines.

Each method of :class:`Lattice` carries the name, argument meaning and
outputs of one Fortran subroutine of derek-teaney/lceqcd and runs it on the
device-resident lattice through liblceq.so (include/lceq.h).  Field arrays use
the reference COMMON layout (f/INCLUDE/gaugefield.inc:71-76) with leading
dimension N = V/2:

* links ``(3, 2, N, 4)`` = (mu, parity, site, [Re alpha, Im alpha, Re beta, Im beta])
* higgs ``(2, N, 3)``    = (parity, site, a(1..3))

Parity 0 is the reference's "even" (x+y+z even, 0-based).  There is no CPU
fallback: every call runs the HIP kernels or raises :class:`LceqError`.
"""
import ctypes as C
import os
import math

import numpy as np

from . import _lib
from ._lib import LceqError, check

SWEEP_KP_OR_METRO = 0
SWEEP_REFERENCE = 1
KC_LINK_KP, KC_LINK_OR, KC_HIGGS, KC_GFIX = 1, 2, 3, 4
MULTALL_KINDS = ("nn", "dn", "nd", "dd", "ad", "an")


def _dp(x):
    return x.ctypes.data_as(C.POINTER(C.c_double))


def couplings_from_xy(beta, x, y):
    """(beta_a, beta_2, beta_4) from (beta, x, y) as f/paraminit.f:150-153."""
    sigma = 3.1759114
    pi = 2.0 * math.acos(0.0)
    beta_a = beta
    beta_2 = beta_a * (3 + 8 * y / beta ** 2 - sigma * (4 + 5 * x) / (2 * pi * beta)
                       - ((20 * x - 10 * x ** 2) * (math.log(1.5 * beta) + 0.09) + 8.7
                          + 11.6 * x) / (2 * pi ** 2 * beta ** 2))
    beta_4 = x * beta
    return beta_a, beta_2, beta_4


def couplings_from_trace_convention(beta_g, beta_a, beta_2, beta_4):
    """Map S = beta_G sum(1 - Tr U_P/2) - beta_A sum Tr[A U A' U^+]
    + sum[beta_2 Tr A^2 + beta_4 (Tr A^2)^2] with A = A^a sigma^a/2 onto the
    reference action (f/agauge.f:142-169, f/ahiggs1.f:77-78), where a = A^a:
    Tr[A U A' U^+] = a.R(U)a'/2 and Tr A^2 = |a|^2/2."""
    return beta_g, beta_a / 2.0, beta_2 / 2.0, beta_4 / 4.0


class Lattice:
    """A whole lx*ly*lz lattice resident on the current HIP device."""

    def __init__(self, lx, ly, lz, lib=None):
        """``lib``: another build of the library with the same entry points
        (the test-only ``_lib.load_oracle_forms()``); default liblceq.so."""
        self._L = lib or _lib.load()
        h = C.c_void_p()
        self._check(self._L.lceq_create(lx, ly, lz, C.byref(h)))
        self.h = h
        self.shape = (lx, ly, lz)
        self.V = lx * ly * lz
        self.N = self.V // 2

    def _check(self, rc):
        check(rc, self._L)

    def close(self):
        if getattr(self, "h", None):
            self._L.lceq_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def sync(self):
        self._check(self._L.lceq_sync(self.h))

    # ------------------------------------------------------------ parameters
    def set_couplings(self, beta, beta_a, beta_2, beta_4):
        self._check(self._L.lceq_set_couplings(self.h, beta, beta_a, beta_2, beta_4))

    def set_gauge_fixing(self, lam, alpha, limit_or, max_iter=1000000):
        self._check(self._L.lceq_set_gauge_fixing(self.h, lam, alpha, limit_or, max_iter))

    def set_imode(self, imode):
        self._check(self._L.lceq_set_imode(self.h, imode))

    # ------------------------------------------------------ learned proposal
    def higgs_mlp(self, mlp):
        """One Higgs sweep with the learned proposal (exact MH); returns accept."""
        acc = C.c_double()
        self._check(self._L.lceq_higgs_mlp(self.h, mlp.h, C.byref(acc)))
        return acc.value

    def sweep_mlp(self, mlp, nsweep, n_or=4, measflag=1):
        """1 KP + n_or OR + the learned Higgs update per sweep."""
        self._check(self._L.lceq_sweep_mlp(self.h, mlp.h, int(nsweep), int(n_or), int(measflag)))

    # ------------------------------------------------------ configuration files
    def confsave(self, u_path, a_path, real_bytes=8):
        """f/confsave.f: write the lattice in the reference's file format."""
        self._check(self._L.lceq_confsave(self.h, None if u_path is None else os.fsencode(u_path),
                                    None if a_path is None else os.fsencode(a_path), real_bytes))

    def oldconf(self, u_path, a_path, real_bytes=8):
        """f/confget.f:318 oldconf: load a configuration file into the lattice."""
        self._check(self._L.lceq_oldconf(self.h, None if u_path is None else os.fsencode(u_path),
                                   None if a_path is None else os.fsencode(a_path), real_bytes))

    def set_shape(self, threads_per_block=256, xcd_order=True):
        """Launch shape of the link and Higgs updates (same fields)."""
        self._check(self._L.lceq_set_shape(self.h, int(threads_per_block), 1 if xcd_order else 0))

    def get_shape(self):
        t, x = C.c_int(), C.c_int()
        self._check(self._L.lceq_get_shape(self.h, C.byref(t), C.byref(x)))
        return t.value, bool(x.value)

    def autotune(self, n_or=4, trials=5):
        """Time every launch shape on production sweeps, keep the fastest;
        the chain is restored.  Returns {(threads, xcd): ms per sweep}."""
        ms = np.zeros(8)
        self._check(self._L.lceq_autotune(self.h, int(n_or), int(trials), _dp(ms)))
        return {((64, 128, 256, 512)[i // 2], bool(i % 2)): float(ms[i]) for i in range(8)}

    # -------------------------------------------------------------- uniforms
    def set_site_order(self, strip_rows=0):
        """Site order of the link and Higgs updates: 0 (index order) or strips
        of ``strip_rows`` y rows advancing through z (same fields)."""
        self._check(self._L.lceq_set_site_order(self.h, int(strip_rows)))

    def get_site_order(self):
        v = C.c_int()
        self._check(self._L.lceq_get_site_order(self.h, C.byref(v)))
        return v.value

    def autotune_site_order(self, n_or=4, trials=5):
        """Time every site order at the current launch shape and keep the
        fastest (the chain is left where it was).  Returns {strip_rows: ms}."""
        ms = np.zeros(4)
        self._check(self._L.lceq_autotune_site_order(self.h, int(n_or), int(trials), _dp(ms)))
        return {o: float(t) for o, t in zip((0, 4, 8, 16), ms) if t >= 0}

    def rng_philox(self, seed, pass_counter=0):
        self._check(self._L.lceq_rng_philox(self.h, seed, pass_counter))

    def rng_measinit(self, rank=0):
        """Seed the device streams as f/measinit.f:214-217 seeds the
        reference's random_number for MPI rank ``rank``; returns the 16
        seeds (the Philox key is ranlux.measinit_key of them)."""
        v = (C.c_int * 16)()
        self._check(self._L.lceq_rng_measinit(self.h, int(rank), v))
        return np.array(v[:], np.int32)

    def rng_feed(self, uniforms):
        u = np.ascontiguousarray(uniforms, dtype=np.float64)
        self._check(self._L.lceq_rng_feed(self.h, _dp(u), len(u)))

    def rng_state(self):
        n = C.c_int64()
        p = C.c_uint64()
        self._check(self._L.lceq_rng_state(self.h, C.byref(n), C.byref(p)))
        return n.value, p.value

    # ---------------------------------------------------------------- fields
    def _put(self, fn, arr, shape):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != shape:
            raise ValueError("expected shape %s, got %s" % (shape, a.shape))
        self._check(fn(self.h, _dp(a), self.N))

    def _get(self, fn, shape):
        a = np.empty(shape)
        self._check(fn(self.h, _dp(a), self.N))
        return a

    def put_links(self, u):
        self._put(self._L.lceq_put_links, u, (3, 2, self.N, 4))

    def get_links(self):
        return self._get(self._L.lceq_get_links, (3, 2, self.N, 4))

    def put_higgs(self, a):
        self._put(self._L.lceq_put_higgs, a, (2, self.N, 3))

    def get_higgs(self):
        return self._get(self._L.lceq_get_higgs, (2, self.N, 3))

    def put_links_fixed(self, u):
        self._put(self._L.lceq_put_links_fixed, u, (3, 2, self.N, 4))

    def get_links_fixed(self):
        return self._get(self._L.lceq_get_links_fixed, (3, 2, self.N, 4))

    def put_higgs_fixed(self, a):
        self._put(self._L.lceq_put_higgs_fixed, a, (2, self.N, 3))

    def get_higgs_fixed(self):
        return self._get(self._L.lceq_get_higgs_fixed, (2, self.N, 3))

    def get_staple(self):
        return self._get(self._L.lceq_get_staple, (3, 2, self.N, 4))

    def get_force(self):
        return self._get(self._L.lceq_get_force, (self.N, 3))

    def unitconf(self):
        self._check(self._L.lceq_unitconf(self.h))

    def testconf(self):
        self._check(self._L.lceq_testconf(self.h))

    def hotstart(self, seed, sigma):
        self._check(self._L.lceq_hotstart(self.h, seed, sigma))

    # ------------------------------------------------------ reference routines
    def staples(self, eo, mu):
        """f/staples.f:9 staples(eo, oe, mu); the staple is read by get_staple()."""
        self._check(self._L.lceq_staples(self.h, eo, mu))

    def heat_kpi(self, b, a0_old):
        """f/heat_kpi.f:9 on len(b) sites; returns (v (n,4), nacch)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        a0 = np.ascontiguousarray(a0_old, dtype=np.float64)
        v = np.zeros((len(b), 4))
        nacch = C.c_int()
        self._check(self._L.lceq_heat_kpi(self.h, len(b), _dp(b), _dp(v), _dp(a0), C.byref(nacch)))
        return v, nacch.value

    def agauge(self, measflag):
        """f/agauge.f:7 agauge(measflag, accept); returns accept."""
        acc = C.c_double()
        self._check(self._L.lceq_agauge(self.h, measflag, C.byref(acc)))
        return acc.value

    def link_or(self, measflag):
        """Link overrelaxation sweep (no reference counterpart); returns accept."""
        acc = C.c_double()
        self._check(self._L.lceq_link_or(self.h, measflag, C.byref(acc)))
        return acc.value

    def a_force(self, eo):
        """f/a_force.f:7 a_force(eo); the force is read by get_force()."""
        self._check(self._L.lceq_a_force(self.h, eo))

    def gauss(self, r):
        """f/gauss.f:10 gauss(X, R, NH); returns X."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        x = np.zeros(len(r))
        self._check(self._L.lceq_gauss(self.h, _dp(x), _dp(r), len(r)))
        return x

    def ahiggs1(self, delta, measflag):
        """f/ahiggs1.f:7 ahiggs1(delta, measflag, accept); returns (delta, accept)."""
        d = C.c_double(delta)
        acc = C.c_double()
        self._check(self._L.lceq_ahiggs1(self.h, C.byref(d), measflag, C.byref(acc)))
        return d.value, acc.value

    def ahiggs2(self, measflag):
        """f/ahiggs2.f:7 ahiggs2(measflag, accept); returns accept."""
        acc = C.c_double()
        self._check(self._L.lceq_ahiggs2(self.h, measflag, C.byref(acc)))
        return acc.value

    def ahiggs3(self, measflag):
        """f/ahiggs3.f:7 ahiggs3(measflag, accept); returns accept."""
        acc = C.c_double()
        self._check(self._L.lceq_ahiggs3(self.h, measflag, C.byref(acc)))
        return acc.value

    def renorm(self):
        """f/renorm.f:6"""
        self._check(self._L.lceq_renorm(self.h))

    def check_fields(self, tol=-1.0):
        """Field guard: (max unitarity deviation of the links, number of
        non-finite link and Higgs values); raises LceqError if any value is
        non-finite or, for tol >= 0, the deviation exceeds tol."""
        dev = C.c_double()
        bad = C.c_int64()
        self._check(self._L.lceq_check_fields(self.h, float(tol), C.byref(dev), C.byref(bad)))
        return dev.value, bad.value

    def u_f_renorm(self):
        """f/u_f_renorm.f:6"""
        self._check(self._L.lceq_u_f_renorm(self.h))

    def or_gauge_fixing(self):
        """f/or_gauge_fixing1.f:5 or_gauge_fixing(da1); returns da1."""
        da = C.c_double()
        self._check(self._L.lceq_or_gauge_fixing(self.h, C.byref(da)))
        return da.value

    def landau_gauge(self):
        """f/landau_gauge.f:6 landau_gauge(da1, iter_or1); returns (da1, iter_or1)."""
        da = C.c_double()
        it = C.c_int()
        self._check(self._L.lceq_landau_gauge(self.h, C.byref(da), C.byref(it)))
        return da.value, it.value

    def correlator(self):
        """f/correlator.f:11; returns (corr_a, corr_u, corr_3) over z."""
        lz = self.shape[2]
        ca, cu, c3 = np.zeros(lz), np.zeros(lz), np.zeros(lz)
        self._check(self._L.lceq_correlator(self.h, _dp(ca), _dp(cu), _dp(c3)))
        return ca, cu, c3

    def scalarcorr(self):
        """f/scalarcorr.f:10; returns (mean rho-slice, corr)."""
        m = C.c_double()
        c = np.zeros(self.shape[2])
        self._check(self._L.lceq_scalarcorr(self.h, C.byref(m), _dp(c)))
        return m.value, c

    def polyakov(self):
        """f/polyakov.f:10; returns (mean(P)^2, corr)."""
        m = C.c_double()
        c = np.zeros(self.shape[2])
        self._check(self._L.lceq_polyakov(self.h, C.byref(m), _dp(c)))
        return m.value, c

    def multall(self, kind, op1, op2, s1=None, s2=None):
        """f/matmul.f MULTALL<kind>; kind in nn dn nd dd ad an; 0-based site maps."""
        k = MULTALL_KINDS.index(kind)
        op1 = np.ascontiguousarray(op1, dtype=np.float64)
        op2 = np.ascontiguousarray(op2, dtype=np.float64)
        n = op2.shape[0] if s2 is None else len(s2)
        out = np.zeros((n, 4))
        ip = C.POINTER(C.c_int)
        a1 = None if s1 is None else np.ascontiguousarray(s1, dtype=np.int32)
        a2 = None if s2 is None else np.ascontiguousarray(s2, dtype=np.int32)
        self._check(self._L.lceq_multall(self.h, k, _dp(op1), _dp(op2), _dp(out), n,
                                   None if a1 is None else a1.ctypes.data_as(ip),
                                   None if a2 is None else a2.ctypes.data_as(ip)))
        return out

    def observables(self):
        """COMMON /aver_loc/ (f/INCLUDE/measurement.inc:30)."""
        out = np.zeros(7)
        self._check(self._L.lceq_observables(self.h, _dp(out)))
        return dict(plaq0=out[0], a0kin=out[1], a02=out[2], a04=out[3], a0f=out[4:7].copy())

    # ---------------------------------------------------------------- sweeps
    def sweep(self, schedule, nsweep, n_or=4, delta=0.4, measflag=1):
        """nsweep device-resident sweeps; returns the (device-tuned) delta."""
        d = C.c_double(delta)
        self._check(self._L.lceq_sweep(self.h, schedule, nsweep, n_or, C.byref(d), measflag))
        return d.value

    def history(self, n):
        out = np.zeros((n, 16))
        self._check(self._L.lceq_history(self.h, _dp(out), n))
        return out

    def profile_begin(self, kernel_class, max_launches):
        self._check(self._L.lceq_profile_begin(self.h, kernel_class, max_launches))

    def profile_end(self):
        ms = C.c_double()
        n = C.c_int()
        self._check(self._L.lceq_profile_end(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value
