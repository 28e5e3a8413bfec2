"""ctypes binding of liblceq.so (include/lceq.h)."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblceq.so")
# test-only build: no contraction, 1/sqrt and cos(2 pi r) as the CPU
# restatement writes them (lceqcd_amd/build.py)
ORACLE_FORMS_PATH = os.path.join(HERE, "liblceq_oracleforms.so")

D, I, I64, U64, VP = C.c_double, C.c_int, C.c_int64, C.c_uint64, C.c_void_p
DP, IP = C.POINTER(C.c_double), C.POINTER(C.c_int)

# name -> argtypes (all return int status)
SIGNATURES = {
    "lceq_create": [I, I, I, C.POINTER(VP)],
    "lceq_destroy": [VP],
    "lceq_sync": [VP],
    "lceq_set_couplings": [VP, D, D, D, D],
    "lceq_set_gauge_fixing": [VP, D, D, D, I],
    "lceq_set_imode": [VP, I],
    "lceq_set_shape": [VP, I, I],
    "lceq_get_shape": [VP, IP, IP],
    "lceq_autotune": [VP, I, I, DP],
    "lceq_set_site_order": [VP, I],
    "lceq_get_site_order": [VP, IP],
    "lceq_autotune_site_order": [VP, I, I, DP],
    "lceq_copy_bandwidth": [I64, I, DP],
    "lceq_stream_bandwidth": [I, I64, I, DP],
    "lceq_rng_philox": [VP, U64, U64],
    "lceq_rng_feed": [VP, DP, I64],
    "lceq_rng_state": [VP, C.POINTER(I64), C.POINTER(U64)],
    "lceq_put_links": [VP, DP, I64],
    "lceq_get_links": [VP, DP, I64],
    "lceq_put_higgs": [VP, DP, I64],
    "lceq_get_higgs": [VP, DP, I64],
    "lceq_put_links_fixed": [VP, DP, I64],
    "lceq_get_links_fixed": [VP, DP, I64],
    "lceq_put_higgs_fixed": [VP, DP, I64],
    "lceq_get_higgs_fixed": [VP, DP, I64],
    "lceq_get_staple": [VP, DP, I64],
    "lceq_get_force": [VP, DP, I64],
    "lceq_unitconf": [VP],
    "lceq_testconf": [VP],
    "lceq_hotstart": [VP, U64, D],
    "lceq_staples": [VP, I, I],
    "lceq_heat_kpi": [VP, I, DP, DP, DP, IP],
    "lceq_agauge": [VP, I, DP],
    "lceq_link_or": [VP, I, DP],
    "lceq_a_force": [VP, I],
    "lceq_gauss": [VP, DP, DP, I],
    "lceq_ahiggs1": [VP, DP, I, DP],
    "lceq_ahiggs2": [VP, I, DP],
    "lceq_ahiggs3": [VP, I, DP],
    "lceq_renorm": [VP],
    "lceq_u_f_renorm": [VP],
    "lceq_or_gauge_fixing": [VP, DP],
    "lceq_landau_gauge": [VP, DP, IP],
    "lceq_correlator": [VP, DP, DP, DP],
    "lceq_scalarcorr": [VP, DP, DP],
    "lceq_polyakov": [VP, DP, DP],
    "lceq_multall": [VP, I, DP, DP, DP, I, IP, IP],
    "lceq_observables": [VP, DP],
    "lceq_sweep": [VP, I, I, I, DP, I],
    "lceq_history": [VP, DP, I],
    "lceq_profile_begin": [VP, I, I],
    "lceq_profile_end": [VP, DP, IP],
    "lceq_batch_create": [I, IP, C.POINTER(VP)],
    "lceq_batch_destroy": [VP],
    "lceq_batch_info": [VP, IP, C.POINTER(I64)],
    "lceq_batch_set_couplings": [VP, DP],
    "lceq_batch_rng_philox": [VP, C.POINTER(U64)],
    "lceq_batch_hotstart": [VP, C.POINTER(U64), D],
    "lceq_batch_put_links": [VP, I, DP, I64],
    "lceq_batch_get_links": [VP, I, DP, I64],
    "lceq_batch_put_higgs": [VP, I, DP, I64],
    "lceq_batch_get_higgs": [VP, I, DP, I64],
    "lceq_batch_sweep": [VP, I, I, DP, I, I],
    "lceq_batch_sync": [VP],
    "lceq_batch_history": [VP, DP, I],
    "lceq_batch_observables": [VP, DP],
    "lceq_batch_set_shape": [VP, I, I],
    "lceq_batch_get_shape": [VP, IP, IP],
    "lceq_batch_set_group_bytes": [VP, I64],
    "lceq_batch_get_groups": [VP, C.POINTER(I64), IP, IP, I],
    "lceq_batch_autotune": [VP, I, I, DP],
    "lceq_batch_sweep_mlp": [VP, VP, I, I, I, I],
    "lceq_batch_mlp_collect": [VP, VP, I, I],
    "lceq_dims": [VP, IP],
    "lceq_mlp_create": [U64, C.POINTER(VP)],
    "lceq_mlp_destroy": [VP],
    "lceq_mlp_get_state": [VP, C.POINTER(C.c_float), I],
    "lceq_mlp_set_state": [VP, C.POINTER(C.c_float), I],
    "lceq_mlp_collect": [VP, VP, I, I],
    "lceq_mlp_records": [VP, C.POINTER(I64)],
    "lceq_mlp_clear": [VP],
    "lceq_mlp_train": [VP, I, I, D, U64, I, DP],
    "lceq_mlp_eval": [VP, C.POINTER(C.c_float), I, C.POINTER(C.c_float)],
    "lceq_mlp_set_sites_per_wave": [VP, I],
    "lceq_check_fields": [VP, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_int64)],
    "lceq_check_fields_static": [VP, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_int64)],
    "lceq_higgs_mlp": [VP, VP, DP],
    "lceq_sweep_mlp": [VP, VP, I, I, I],
    "lceq_conf_write": [IP, DP, DP, C.c_char_p, C.c_char_p, I],
    "lceq_conf_read": [IP, DP, DP, C.c_char_p, C.c_char_p, I],
    "lceq_confsave": [VP, C.c_char_p, C.c_char_p, I],
    "lceq_oldconf": [VP, C.c_char_p, C.c_char_p, I],
    "lceq_conf_filename": [I, I, I, D, D, D, I, C.c_char_p, I],
    "lceq_ranlux_create": [C.POINTER(VP)],
    "lceq_ranlux_destroy": [VP],
    "lceq_ranlux": [VP, C.POINTER(C.c_float), I],
    "lceq_rluxgo": [VP, I, I, I, I],
    "lceq_rluxut": [VP, IP],
    "lceq_rluxin": [VP, IP],
    "lceq_rluxat": [VP, IP, IP, IP, IP],
    "lceq_measinit_seeds": [I, IP],
    "lceq_rng_measinit": [VP, I, IP],
}

_lib = None


class LceqError(RuntimeError):
    pass


def _open(path):
    if not os.path.exists(path):
        raise LceqError(
            "%s is not built; run `python -m lceqcd_amd.build` "
            "(there is no CPU fallback)" % os.path.basename(path))
    lib = C.CDLL(path)
    for name, args in SIGNATURES.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = I
    lib.lceq_last_error.argtypes = []
    lib.lceq_last_error.restype = C.c_char_p
    lib.lceq_measinit_key.argtypes = [IP]
    lib.lceq_measinit_key.restype = U64
    return lib


def load():
    """Load the in-tree liblceq.so; fails loudly when it has not been built."""
    global _lib
    if _lib is None:
        _lib = _open(LIB_PATH)
    return _lib


_oracle_forms = None


def load_oracle_forms():
    """The test-only build liblceq_oracleforms.so (same entry points; every
    operation rounded as written, 1/sqrt and cos(2 pi r) in the CPU
    restatement's forms).  Pass it as ``Lattice(..., lib=...)``."""
    global _oracle_forms
    if _oracle_forms is None:
        _oracle_forms = _open(ORACLE_FORMS_PATH)
    return _oracle_forms


def check(rc, lib=None):
    if rc != 0:
        msg = (lib or load()).lceq_last_error().decode(errors="replace")
        raise LceqError("liblceq status %d: %s" % (rc, msg))
