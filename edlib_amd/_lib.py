"""ctypes binding of libedhip.so (the C ABI declared in include/edhip.h).

The library is built in-tree (edlib_amd/libedhip.so).  There is no CPU
fallback: if the library is missing, loading raises immediately.
"""
import ctypes
import os
import re

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "libedhip.so")
HEADER = os.path.join(ROOT, "include", "edhip.h")

EDH_OK = 0
EDH_MEM_HOST = 0
EDH_MEM_DEVICE = 1
EDH_COPY_H2D, EDH_COPY_D2H, EDH_COPY_D2D = 0, 1, 2
EDH_OP_SZ, EDH_OP_N = 0, 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p
_lib = None


class EdhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("edhip error %d: %s" % (code, msg))
        self.code = code


_SIGS = {
    "edh_last_error": (ctypes.c_char_p, []),
    "edh_version": (ctypes.c_int, []),
    "edh_model_hubbard": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, _dp, _dp, _dp, _dp, _dp]),
    "edh_model_siam": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp,
                                      ctypes.c_double]),
    "edh_model_free": (ctypes.c_int, [_vp]),
    "edh_model_sites": (ctypes.c_int, [_vp, _ip, _ip]),
    "edh_model_transitions": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _ip, _dp]),
    "edh_model_diagonal": (ctypes.c_int, [_vp, ctypes.c_longlong, _dp]),
    "edh_sector_csr": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _i64p, _i64p, _ip, _ip, _dp]),
    "edh_sector_diagonal": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp]),
    "edh_model_from_terms": (ctypes.c_int, [ctypes.POINTER(_vp), _vp]),
    "edh_model_apply": (ctypes.c_int, [_vp, ctypes.c_longlong, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong), _dp,
                                       _ip]),
    "edh_device_memory": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_size_t),
                                         ctypes.POINTER(ctypes.c_size_t)]),
    "edh_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int]),
    "edh_destroy": (ctypes.c_int, [_vp]),
    "edh_set_model": (ctypes.c_int, [_vp, _vp]),
    "edh_set_sector": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _i64p]),
    "edh_constant_shift": (ctypes.c_int, [_vp, ctypes.c_double]),
    "edh_sector_info": (ctypes.c_int, [_vp, _i64p, _i64p, _i64p]),
    "edh_alloc": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "edh_free": (ctypes.c_int, [_vp, _vp]),
    "edh_memcpy": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t, ctypes.c_int]),
    "edh_synchronize": (ctypes.c_int, [_vp]),
    "edh_av": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int64, ctypes.c_int, ctypes.c_int]),
    "edh_static_observables": (ctypes.c_int, [_vp, _vp, ctypes.c_int64, ctypes.c_int, ctypes.c_int, _dp]),
    "edh_batch_static_observables": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]),
    "edh_get_hv_variant": (ctypes.c_int, [_vp, _ip]),
    "edh_set_hv_variant": (ctypes.c_int, [_vp, ctypes.c_int]),
    "edh_av_time": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]),
    "edh_vv": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int64, _dp, ctypes.c_int]),
    "edh_vv_noalloc": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int64, _vp]),
    "edh_lanczos": (ctypes.c_int, [_vp, _vp, ctypes.c_int64, ctypes.c_int, _dp, _dp, _ip, ctypes.c_int]),
    "edh_lanczos_time": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _dp, _dp, _ip, _dp]),
    "edh_adag": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp, _i64p,
                                ctypes.c_int]),
    "edh_local_op": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _vp, _vp,
                                    ctypes.c_int]),
    "edh_cf_eval": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_double, _dp, _dp, _dp, _dp]),
    "edh_eigsh": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_uint64,
                                 _dp, _dp, _ip, _ip]),
    "edh_batch_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int]),
    "edh_batch_destroy": (ctypes.c_int, [_vp]),
    "edh_batch_add_sector": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, _ip, _i64p]),
    "edh_batch_remove_sector": (ctypes.c_int, [_vp, ctypes.c_int]),
    "edh_batch_submit_eig": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                            ctypes.c_int, ctypes.c_uint64, _ip]),
    "edh_batch_submit_lanczos": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int, _ip]),
    "edh_batch_static_observables_many": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _ip, ctypes.c_int, _dp]),
    "edh_batch_add_sectors": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _ip, _ip, _ip, ctypes.POINTER(ctypes.c_int64)]),
    "edh_batch_submit_gf": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _ip, _ip, _ip, _ip, ctypes.c_int, _ip]),
    "edh_batch_submit_gf2": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _ip, _ip, _ip, _ip, _ip, ctypes.c_int, _ip]),
    "edh_batch_lanczos_norm2": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _dp]),
    "edh_batch_step": (ctypes.c_int, [_vp, ctypes.c_int, _ip]),
    "edh_batch_job_info": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _ip, _i64p, _ip, _ip]),
    "edh_batch_job_states": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _ip]),
    "edh_batch_eig_bounds": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _dp]),
    "edh_batch_eig_result": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _ip]),
    "edh_batch_lanczos_result": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _ip]),
    "edh_batch_free_job": (ctypes.c_int, [_vp, ctypes.c_int]),
    "edh_batch_cancel": (ctypes.c_int, [_vp, ctypes.c_int]),
    "edh_batch_memory": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_size_t)]),
    "edh_batch_set_graphs": (ctypes.c_int, [_vp, ctypes.c_int]),
    "edh_batch_sweeps": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "edh_batch_autotune": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _ip]),
    "edh_batch_time_hv": (ctypes.c_int, [_vp, ctypes.c_int, _ip, ctypes.c_int, ctypes.c_uint64, _dp, _i64p, _i64p]),
    "edh_batch_cf_many": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _dp, _ip, _dp, _dp, ctypes.c_int, _dp, _dp, _dp]),
    "edh_batch_cf_many_z": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _dp, _ip, _dp, _dp, ctypes.c_int, _dp, _dp, _dp,
                                           _dp]),
    "edh_batch_cf_eval": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_double, _dp, _dp, _dp, _dp]),
}


def header_symbols():
    """Names of every function declared in include/edhip.h."""
    with open(HEADER) as f:
        txt = f.read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(edh_\w+)\s*\(", txt, re.M)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libedhip.so is not built (run __graft_entry__.build()); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code):
    if code != EDH_OK:
        raise EdhError(code, lib().edh_last_error().decode())
    return code
