"""ctypes bindings of ``libstopeps_hip.so`` (declarations in include/stopeps_hip.h)."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libstopeps_hip.so")

_lib = None

c_int_p = ctypes.POINTER(ctypes.c_int)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_float_p = ctypes.POINTER(ctypes.c_float)
c_void_p = ctypes.c_void_p

_SIGNATURES = {
    "spp_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(c_void_p)]),
    "spp_destroy": (ctypes.c_int, [c_void_p]),
    "spp_last_error": (ctypes.c_char_p, [c_void_p]),
    "spp_synchronize": (ctypes.c_int, [c_void_p]),
    "spp_set_precision": (ctypes.c_int, [c_void_p, ctypes.c_int]),
    "spp_set_option": (ctypes.c_int, [c_void_p, ctypes.c_char_p, ctypes.c_double]),
    "spp_get_counters": (ctypes.c_int, [c_void_p, c_double_p, ctypes.c_int]),
    "spp_reset_counters": (ctypes.c_int, [c_void_p]),
    "spp_state_upload": (ctypes.c_int, [c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, c_int_p, c_int_p,
                                        ctypes.POINTER(c_double_p), ctypes.c_double, ctypes.c_double,
                                        ctypes.POINTER(c_void_p)]),
    "spp_state_free": (ctypes.c_int, [c_void_p]),
    "spp_sampling_cache_build": (ctypes.c_int, [c_void_p, c_void_p, ctypes.c_int, c_double_p,
                                                ctypes.POINTER(c_void_p)]),
    "spp_sampling_cache_build_kind": (ctypes.c_int, [c_void_p, c_void_p, ctypes.c_int, c_double_p, ctypes.c_int,
                                                     ctypes.POINTER(c_void_p)]),
    "spp_sampling_cache_free": (ctypes.c_int, [c_void_p]),
    "spp_sampling_cache_info": (ctypes.c_int, [c_void_p, c_double_p]),
    "spp_born_sample_batch": (ctypes.c_int, [c_void_p, c_void_p, ctypes.c_int, c_double_p, c_int_p, c_int_p,
                                             c_double_p, c_double_p, c_double_p, c_double_p, c_double_p,
                                             c_int_p, c_int_p]),
    "spp_born_sample_batch_repairs": (ctypes.c_int, [c_void_p, c_void_p, ctypes.c_int, c_double_p, c_int_p,
                                                     c_int_p, c_double_p, c_double_p, c_double_p, c_double_p,
                                                     c_double_p, c_int_p, c_int_p, c_double_p]),
    "spp_measure_energy_batch": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_int, ctypes.c_int, c_int_p,
                                                c_double_p, c_double_p, c_double_p, ctypes.c_int, ctypes.c_double,
                                                ctypes.c_double, c_double_p, c_double_p, c_double_p, c_double_p,
                                                c_double_p, c_int_p, c_int_p, c_double_p, c_double_p]),
    "spp_measure_energy_states": (ctypes.c_int, [c_void_p, ctypes.c_int, ctypes.POINTER(c_void_p), ctypes.c_int,
                                                 c_double_p, c_double_p, c_double_p, ctypes.c_int, ctypes.c_double,
                                                 ctypes.c_double, c_double_p, c_double_p, c_double_p, c_double_p,
                                                 c_double_p, c_int_p, c_int_p, c_double_p, c_double_p]),
    "spp_lognorm_batch": (ctypes.c_int, [c_void_p, ctypes.c_int, ctypes.POINTER(c_void_p), ctypes.c_int,
                                         c_double_p, c_double_p, c_double_p, c_int_p]),
    "spp_lognorm_batch_noalloc": (ctypes.c_int, [c_void_p, ctypes.c_int, ctypes.POINTER(c_void_p), ctypes.c_int,
                                                 c_double_p, c_double_p, c_double_p, c_int_p]),
    "spp_lognorm_grad_batch": (ctypes.c_int, [c_void_p, ctypes.c_int, ctypes.POINTER(c_void_p), ctypes.c_int,
                                              c_double_p, c_double_p, c_double_p, c_double_p, c_int_p]),
    "spp_selftest_gemm": (ctypes.c_int, [c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, c_double_p,
                                         c_double_p, c_double_p]),
    "spp_selftest_fu_bond": (ctypes.c_int, [c_int_p, c_double_p, c_double_p, c_double_p, c_double_p,
                                            c_double_p, c_double_p]),
    "spp_state_set_live": (ctypes.c_int, [c_void_p, c_void_p, c_int_p]),
    "spp_project_batch": (ctypes.c_int, [c_void_p, c_void_p, ctypes.c_int, c_int_p, c_double_p, c_int_p,
                                         ctypes.POINTER(c_void_p), c_double_p, c_double_p]),
    "spp_state_download": (ctypes.c_int, [c_void_p, c_void_p, ctypes.POINTER(c_double_p), c_double_p, c_double_p,
                                          c_int_p]),
    "spp_fu_evolution_step": (ctypes.c_int, [c_void_p, ctypes.c_int, ctypes.POINTER(c_void_p), ctypes.c_int,
                                             ctypes.c_int, ctypes.c_double, ctypes.c_int, c_double_p, c_double_p,
                                             c_double_p, c_double_p, c_double_p, c_double_p, c_double_p, c_int_p]),
    "spp_fu_set_repair": (ctypes.c_int, [c_void_p, c_void_p, c_void_p]),
    "spp_selftest_jacobi": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, c_double_p, ctypes.c_int,
                                           c_double_p, c_int_p]),
    "spp_selftest_fu_pair": (ctypes.c_int, [ctypes.c_int, c_int_p, c_int_p, c_int_p, c_int_p, c_double_p, c_double_p,
                                            c_int_p, c_int_p, c_double_p, c_double_p, c_double_p, c_double_p,
                                            c_double_p, c_double_p, c_int_p]),
    "spp_selftest_copy": (ctypes.c_int, [c_void_p, ctypes.c_longlong]),
    "spp_selftest_psd": (ctypes.c_int, [c_void_p, ctypes.c_int, ctypes.c_int, c_int_p, c_double_p, c_double_p,
                                        c_double_p, c_double_p, c_int_p]),
    "spp_selftest_ql": (ctypes.c_int, [ctypes.c_int, c_double_p, c_double_p, c_double_p, ctypes.c_int,
                                       ctypes.c_int]),
    "spp_selftest_spectrum": (ctypes.c_int, [c_void_p, ctypes.c_int, ctypes.c_int, c_int_p, c_double_p, c_double_p]),
    "spp_selftest_kron": (ctypes.c_int, [c_void_p, c_int_p, c_int_p, ctypes.c_int, ctypes.c_int, c_double_p,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         c_float_p, c_float_p, ctypes.c_int]),
}

EXPORTED = tuple(_SIGNATURES)


class HIPLibraryError(RuntimeError):
    """An API or HIP error reported by the library."""


def load():
    """Load the in-tree library; fails loudly when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HIPLibraryError(f"HIP extension {LIB_PATH} is not built; run stopeps.jl_amd/build.py")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(ctx, rc):
    if rc != 0:
        msg = load().spp_last_error(ctx).decode() if ctx else "error"
        raise HIPLibraryError(msg)


def dptr(a):
    return a.ctypes.data_as(c_double_p) if a is not None else None


def iptr(a):
    return a.ctypes.data_as(c_int_p) if a is not None else None
