"""ctypes binding of libxcquad.so (C-ABI declared in include/xcquad.h).

The library is the only compute path: if it is missing or fails to load, import of this module raises.  There is no CPU
fallback.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "libxcquad.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_llp = ctypes.POINTER(ctypes.c_longlong)
_vp = ctypes.c_void_p

# . name: (restype, argtypes) for every entry point of include/xcquad.h.
SIGNATURES = {
    "xq_grid_build": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _ip, _dp, ctypes.POINTER(_vp)]),
    "xq_grid_counts": (ctypes.c_int, [_vp, _ip, _ip]),
    "xq_grid_points": (ctypes.c_int, [_vp, _dp, _ip, _ip]),
    "xq_estimated_points": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _ip, _llp]),
    "xq_basis_set": (ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _ip]),
    "xq_integrate": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _dp, _dp, _dp, _dp, _dp]),
    "xq_integrate_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _dp, _dp, _dp, _dp, _dp]),
    "xq_noalloc_scratch": (ctypes.c_int, [_vp, _llp, _llp]),
    "xq_integrate_uks": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "xq_gradients": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _dp, _dp]),
    "xq_gradients_uks": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _dp, _dp, _dp]),
    "xq_estimate_bytes": (ctypes.c_int, [ctypes.c_longlong, ctypes.c_int, ctypes.c_double, ctypes.c_int, _llp]),
    "xq_device_bytes": (ctypes.c_int, [_vp, _llp]),
    "xq_device_memory": (ctypes.c_int, [_llp, _llp]),
    "xq_free": (None, [_vp]),
    "xq_last_error": (ctypes.c_char_p, []),
    "xq_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "xq_integrate_batch": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp), _ip, _ip, ctypes.POINTER(_dp), _dp, _dp,
                                          ctypes.POINTER(_dp), _dp]),
    "xq_batch_free": (None, [_vp]),
    "xq_test_fail_allocations": (None, [ctypes.c_longlong]),
}

# . The SCF wrapper's integrals (include/qcscf.h), in the same library.
SCF_SIGNATURES = {
    "qs_basis_create": (ctypes.c_int, [ctypes.c_int, _ip, _dp, ctypes.c_int, _dp, ctypes.c_int, ctypes.POINTER(_vp)]),
    "qs_basis_size": (ctypes.c_int, [_vp, _ip]),
    "qs_basis_free": (None, [_vp]),
    "qs_one_electron": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "qs_fit_create": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(_vp)]),
    "qs_fit_matrix": (ctypes.c_int, [_vp, _dp]),
    "qs_fit_b": (ctypes.c_int, [_vp, _dp, _dp]),
    "qs_fit_fock": (ctypes.c_int, [_vp, _dp, _dp]),
    "qs_fit_stats": (ctypes.c_int, [_vp, _llp, _llp, _llp]),
    "qs_fit_estimate_bytes": (ctypes.c_int, [ctypes.c_int, _ip, _dp, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp,
                                             ctypes.c_int, _llp]),
    "qs_fit_free": (None, [_vp]),
    "qs_dense_create": (ctypes.c_int, [ctypes.c_int, _dp, _dp, ctypes.POINTER(_vp)]),
    "qs_dense_aufbau": (ctypes.c_int, [_vp, _dp, ctypes.c_int, _dp, _dp]),
    "qs_dense_commutator": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "qs_dense_free": (None, [_vp]),
}


def load(path=LIBRARY):
    """Load the library and declare every entry point; raises OSError if it cannot be loaded."""
    if not os.path.exists(path):
        raise OSError("libxcquad.so not found at {:s}: build it with `python -m pdynamo3_amd.build`.".format(path))
    lib = ctypes.CDLL(path)
    for name, (res, args) in list(SIGNATURES.items()) + list(SCF_SIGNATURES.items()):
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = load()


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)


def last_error():
    msg = lib.xq_last_error()
    return msg.decode() if msg else ""


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)
