"""ctypes binding of libnlfem_amd.so (declared in include/nlfem_amd.h).

The library is built in-tree by ``nlfem_amd.build`` (or ``__graft_entry__.build()``).  There is
no fallback: if the library is missing or fails to load, every call raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnlfem_amd.so")
_LIB = None

_dp = C.POINTER(C.c_double)
_lp = C.POINTER(C.c_long)
_llp = C.POINTER(C.c_longlong)
_ip = C.POINTER(C.c_int)
_cp = C.c_char_p
_fp = C.POINTER(C.c_float)

# exported symbols and their signatures: (restype, argtypes)
SIGNATURES = {
    "nlfem_last_error": (_cp, []),
    "nlfem_par_assemble": (C.c_int, [_cp, _cp, _cp, C.c_int, C.c_int, _lp, _lp, _dp, _lp, C.c_int, C.c_int, C.c_int,
                                     C.c_int, _dp, C.c_int, _dp, _dp, C.c_int, _dp, C.c_double, C.c_int, C.c_int, _cp,
                                     _cp, _cp, _cp, C.c_int, C.c_int, C.c_int, _lp, _lp, C.c_long, _dp, C.c_int, _dp,
                                     C.c_double, C.c_double, C.c_int, C.c_int]),
    "nlfem_par_evaluateMass": (C.c_int, [_dp, _dp, _lp, _lp, _dp, _lp, C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int,
                                         C.c_int, C.c_int, C.c_int, C.c_int]),
    "nlfem_method_retriangulate": (C.c_int, [_dp, _dp, C.c_double, _dp, C.c_int]),
    "nlfem_toRef": (C.c_int, [_dp, _dp, _dp]),
    "nlfem_toPhys": (C.c_int, [_dp, _dp, C.c_int, _dp]),
    "nlfem_solve2x2": (C.c_int, [_dp, _dp, _dp]),
    "nlfem_retriangulate_batch": (C.c_int, [C.c_int, _dp, _dp, C.c_double, C.c_int, _dp, _ip]),
    "nlfem_plan_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, _lp, _lp, _dp, _lp, C.c_int, C.c_int, _dp,
                                    C.c_int, _dp, _dp, C.c_int, _dp, C.c_double, C.c_int, _cp, _cp, _cp, C.c_int,
                                    C.c_int, _dp, C.c_int, _dp, C.c_double, C.c_int, C.c_int]),
    "nlfem_plan_assemble": (C.c_int, [C.c_void_p]),
    "nlfem_plan_csr_info": (C.c_int, [C.c_void_p, _llp, _llp]),
    "nlfem_plan_csr_copy": (C.c_int, [C.c_void_p, _llp, _ip, _dp]),
    "nlfem_plan_csr_device": (C.c_int, [C.c_void_p, C.POINTER(_llp), C.POINTER(_ip), C.POINTER(_dp)]),
    "nlfem_plan_neighbours_total": (C.c_int, [C.c_void_p, _llp]),
    "nlfem_plan_neighbours": (C.c_int, [C.c_void_p, _llp, _ip, C.POINTER(C.c_ubyte)]),
    "nlfem_plan_stats": (C.c_int, [C.c_void_p, _dp, C.c_int]),
    "nlfem_plan_destroy": (None, [C.c_void_p]),
    "nlfem_load_vector": (C.c_int, [C.c_int, _lp, _lp, _dp, _lp, C.c_int, C.c_int, _dp, C.c_int, _dp, C.c_int, _cp,
                                    C.c_int, _dp, _lp, _lp, C.c_long]),
    "nlfem_plan_set_zeta": (C.c_int, [C.c_void_p, _lp, _lp, C.c_long]),
    "nlfem_plan_set_estimates": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double]),
    "nlfem_plan_force_paths": (C.c_int, [C.c_void_p, C.c_int]),
    "nlfem_plan_set_cut_items": (C.c_int, [C.c_void_p, C.c_int]),
    "nlfem_plan_autotune": (C.c_int, [C.c_void_p, _ip, _dp]),
    "nlfem_plan_set_surrogate": (C.c_int, [C.c_void_p, C.c_int, _fp, _fp, C.c_float, C.c_double, C.c_double]),
    "nlfem_plan_surrogate_samples": (C.c_int, [C.c_void_p, C.c_longlong, C.c_uint, C.c_uint, C.c_double, C.c_double,
                                               _fp, _fp, _llp]),
    "nlfem_mlp_info": (C.c_int, [_ip, _ip, _ip, _ip, _ip]),
    "nlfem_mlp_forward": (C.c_int, [_fp, _fp, _fp, C.c_longlong, _fp]),
    "nlfem_mlp_forward_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_longlong, C.c_void_p,
                                            C.c_void_p]),
    "nlfem_mlp_grad": (C.c_int, [_fp, _fp, _fp, _fp, C.c_longlong, _fp, _fp]),
    "nlfem_mlp_train": (C.c_int, [_fp, _fp, _fp, _fp, C.c_longlong, C.c_int, C.c_int, C.c_float, C.c_uint, _fp]),
}


class NlfemError(RuntimeError):
    pass


def lib():
    """Load libnlfem_amd.so once; raise if it is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise NlfemError(f"{LIB_PATH} is not built: run `python -m nlfem_amd.build` (HIP extension required)")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def check(rc):
    if rc != 0:
        raise NlfemError(f"nlfem_amd error {rc}: {lib().nlfem_last_error().decode()}")
    return rc


def dp(a):
    return a.ctypes.data_as(_dp)


def lp(a):
    return a.ctypes.data_as(_lp)
