"""ctypes front end of the CPU restatement (liblceq_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.  The
product package never imports it.

Field arrays are numpy views in the reference COMMON layout
(f/INCLUDE/gaugefield.inc:71-76) with leading dimension N = V/2:
links (3, 2, N, 4) = (mu, parity, site, [Re alpha, Im alpha, Re beta, Im beta]),
higgs (2, N, 3).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

D = C.c_double
DP = C.POINTER(C.c_double)
I = C.c_int
I64 = C.c_int64
U64 = C.c_uint64
VP = C.c_void_p


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liblceq_oracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        sig = {
            "oc_create": (VP, [I, I, I]),
            "oc_destroy": (None, [VP]),
            "oc_error": (I, [VP]),
            "oc_set_couplings": (None, [VP, D, D, D, D]),
            "oc_set_gauge_fixing": (None, [VP, D, D, D, I]),
            "oc_set_imode": (None, [VP, I]),
            "oc_set_renorm_on_write": (None, [VP, I]),
            "oc_rng_feed": (None, [VP, DP, I64]),
            "oc_rng_consumed": (I64, [VP]),
            "oc_rng_philox": (None, [VP, U64, U64]),
            "oc_pass_counter": (U64, [VP]),
            "oc_site_tables": (None, [VP, VP, VP]),
            "oc_site_tables_closed_form": (None, [VP, VP, VP]),
            "oc_links": (DP, [VP]),
            "oc_higgs": (DP, [VP]),
            "oc_links_fixed": (DP, [VP]),
            "oc_higgs_fixed": (DP, [VP]),
            "oc_staple": (DP, [VP]),
            "oc_force": (DP, [VP]),
            "oc_unitconf": (None, [VP]),
            "oc_testconf": (None, [VP]),
            "oc_hotstart": (None, [VP, U64, D]),
            "oc_multall": (None, [I, DP, DP, DP, I, VP, VP]),
            "oc_staples": (None, [VP, I, I]),
            "oc_heat_kpi": (None, [VP, I, DP, DP, DP, C.POINTER(I64)]),
            "oc_agauge": (None, [VP, I, DP]),
            "oc_link_or": (None, [VP, I, DP]),
            "oc_a_force": (None, [VP, I]),
            "oc_gauss": (None, [VP, DP, DP, I]),
            "oc_ahiggs1": (None, [VP, DP, I, DP]),
            "oc_ahiggs2": (None, [VP, I, DP]),
            "oc_ahiggs3": (None, [VP, I, DP]),
            "oc_higgs_mlp": (None, [VP, C.POINTER(C.c_float), DP]),
            "oc_renorm": (None, [VP]),
            "oc_u_f_renorm": (None, [VP]),
            "oc_or_gauge_fixing": (None, [VP, DP]),
            "oc_landau_gauge": (None, [VP, DP, C.POINTER(I)]),
            "oc_correlator": (None, [VP, DP, DP, DP]),
            "oc_scalarcorr": (None, [VP, DP, DP]),
            "oc_polyakov": (None, [VP, DP, DP]),
            "oc_observables": (None, [VP, DP]),
            "oc_philox4x32_10": (None, [VP, VP, VP]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _dp(x):
    return x.ctypes.data_as(DP)


def philox4x32_10(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().oc_philox4x32_10(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    return out


def multall(kind, op1, op2, s1=None, s2=None):
    """kind in 'nn','dn','nd','dd','ad','an' (f/matmul.f)."""
    k = ["nn", "dn", "nd", "dd", "ad", "an"].index(kind)
    op1 = np.ascontiguousarray(op1, dtype=np.float64)
    op2 = np.ascontiguousarray(op2, dtype=np.float64)
    n = op2.shape[0] if s2 is None else len(s2)
    out = np.zeros((n, 4))
    a1 = None if s1 is None else np.ascontiguousarray(s1, dtype=np.int32)
    a2 = None if s2 is None else np.ascontiguousarray(s2, dtype=np.int32)
    lib().oc_multall(k, _dp(op1), _dp(op2), _dp(out), n,
                     None if a1 is None else a1.ctypes.data,
                     None if a2 is None else a2.ctypes.data)
    return out


class OracleLattice:
    """One single-rank lattice of the CPU restatement."""

    def __init__(self, lx, ly, lz):
        self._L = lib()
        self.h = self._L.oc_create(lx, ly, lz)
        if not self.h:
            raise ValueError("lattice extents must be even and >= 2")
        self.shape = (lx, ly, lz)
        self.V = lx * ly * lz
        self.N = self.V // 2
        self._fed = None

    def close(self):
        if self.h:
            self._L.oc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # ---- views
    def _view(self, fn, shape):
        p = fn(self.h)
        return np.ctypeslib.as_array(p, shape=shape)

    @property
    def u(self):
        return self._view(self._L.oc_links, (3, 2, self.N, 4))

    @property
    def a(self):
        return self._view(self._L.oc_higgs, (2, self.N, 3))

    @property
    def u_f(self):
        return self._view(self._L.oc_links_fixed, (3, 2, self.N, 4))

    @property
    def a_f(self):
        return self._view(self._L.oc_higgs_fixed, (2, self.N, 3))

    @property
    def v(self):
        return self._view(self._L.oc_staple, (3, 2, self.N, 4))

    @property
    def force(self):
        return self._view(self._L.oc_force, (self.N, 3))

    # ---- setup
    def set_couplings(self, beta, beta_a, beta_2, beta_4):
        self._L.oc_set_couplings(self.h, beta, beta_a, beta_2, beta_4)

    def set_gauge_fixing(self, lam, alpha, limit_or, max_iter=1000000):
        self._L.oc_set_gauge_fixing(self.h, lam, alpha, limit_or, max_iter)

    def set_imode(self, imode):
        self._L.oc_set_imode(self.h, imode)

    def set_renorm_on_write(self, flag):
        self._L.oc_set_renorm_on_write(self.h, int(flag))

    def feed(self, uniforms):
        self._fed = np.ascontiguousarray(uniforms, dtype=np.float64)
        self._L.oc_rng_feed(self.h, _dp(self._fed), len(self._fed))

    def consumed(self):
        return self._L.oc_rng_consumed(self.h)

    def philox(self, seed, pass_counter=0):
        self._L.oc_rng_philox(self.h, seed, pass_counter)

    def pass_counter(self):
        return self._L.oc_pass_counter(self.h)

    def error(self):
        return self._L.oc_error(self.h)

    def site_tables(self, closed_form=False):
        up = np.zeros((2, 3, self.N), dtype=np.int32)
        dn = np.zeros((2, 3, self.N), dtype=np.int32)
        f = self._L.oc_site_tables_closed_form if closed_form else self._L.oc_site_tables
        f(self.h, up.ctypes.data, dn.ctypes.data)
        return up, dn

    def unitconf(self):
        self._L.oc_unitconf(self.h)

    def testconf(self):
        self._L.oc_testconf(self.h)

    def hotstart(self, seed, sigma):
        self._L.oc_hotstart(self.h, seed, sigma)

    # ---- reference routines
    def staples(self, eo, mu):
        self._L.oc_staples(self.h, eo, mu - 1)

    def heat_kpi(self, b, a0_old):
        b = np.ascontiguousarray(b, dtype=np.float64)
        a0 = np.ascontiguousarray(a0_old, dtype=np.float64)
        v = np.zeros((len(b), 4))
        nacch = I64(0)
        self._L.oc_heat_kpi(self.h, len(b), _dp(b), _dp(v), _dp(a0), C.byref(nacch))
        return v, nacch.value

    def agauge(self, measflag):
        acc = D(0)
        self._L.oc_agauge(self.h, measflag, C.byref(acc))
        return acc.value

    def link_or(self, measflag):
        acc = D(0)
        self._L.oc_link_or(self.h, measflag, C.byref(acc))
        return acc.value

    def a_force(self, eo):
        self._L.oc_a_force(self.h, eo)

    def gauss(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        x = np.zeros(len(r))
        self._L.oc_gauss(self.h, _dp(x), _dp(r), len(r))
        return x

    def ahiggs1(self, delta, measflag):
        d = D(delta)
        acc = D(0)
        self._L.oc_ahiggs1(self.h, C.byref(d), measflag, C.byref(acc))
        return d.value, acc.value

    def ahiggs2(self, measflag):
        acc = D(0)
        self._L.oc_ahiggs2(self.h, measflag, C.byref(acc))
        return acc.value

    def ahiggs3(self, measflag):
        acc = D(0)
        self._L.oc_ahiggs3(self.h, measflag, C.byref(acc))
        return acc.value

    def higgs_mlp(self, state):
        """The learned Higgs proposal with network state (5274 float32)."""
        st = np.ascontiguousarray(state, dtype=np.float32)
        acc = D(0)
        self._L.oc_higgs_mlp(self.h, st.ctypes.data_as(C.POINTER(C.c_float)), C.byref(acc))
        return acc.value

    def renorm(self):
        self._L.oc_renorm(self.h)

    def u_f_renorm(self):
        self._L.oc_u_f_renorm(self.h)

    def or_gauge_fixing(self):
        da = D(0)
        self._L.oc_or_gauge_fixing(self.h, C.byref(da))
        return da.value

    def landau_gauge(self):
        da = D(0)
        it = I(0)
        self._L.oc_landau_gauge(self.h, C.byref(da), C.byref(it))
        return da.value, it.value

    def correlator(self):
        lz = self.shape[2]
        ca, cu, c3 = np.zeros(lz), np.zeros(lz), np.zeros(lz)
        self._L.oc_correlator(self.h, _dp(ca), _dp(cu), _dp(c3))
        return ca, cu, c3

    def scalarcorr(self):
        lz = self.shape[2]
        m = D(0)
        c = np.zeros(lz)
        self._L.oc_scalarcorr(self.h, C.byref(m), _dp(c))
        return m.value, c

    def polyakov(self):
        lz = self.shape[2]
        m = D(0)
        c = np.zeros(lz)
        self._L.oc_polyakov(self.h, C.byref(m), _dp(c))
        return m.value, c

    def observables(self):
        out = np.zeros(7)
        self._L.oc_observables(self.h, _dp(out))
        return dict(plaq0=out[0], a0kin=out[1], a02=out[2], a04=out[3],
                    a0f=out[4:7].copy())

    # ---- the product's method names (lceqcd_amd.Lattice), so that one parity
    # test body drives either implementation
    def rng_feed(self, uniforms):
        self.feed(uniforms)

    def rng_philox(self, seed, pass_counter=0):
        self.philox(seed, pass_counter)

    def rng_state(self):
        return self.consumed(), self.pass_counter()

    def put_links(self, u):
        self.u[:] = u

    def get_links(self):
        return self.u.copy()

    def put_higgs(self, a):
        self.a[:] = a

    def get_higgs(self):
        return self.a.copy()

    def put_links_fixed(self, u):
        self.u_f[:] = u

    def get_links_fixed(self):
        return self.u_f.copy()

    def put_higgs_fixed(self, a):
        self.a_f[:] = a

    def get_higgs_fixed(self):
        return self.a_f.copy()

    def get_staple(self):
        return self.v.copy()

    def get_force(self):
        return self.force.copy()
