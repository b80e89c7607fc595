"""ctypes binding of the native library ``libapyib_amd.so`` (C ABI in
``include/apyib_amd.h``).

The library is built in-tree by ``apyib_amd/build.py``.  There is no
fallback: if the library is missing, every GPU entry point raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libapyib_amd.so")

APY_OK = 0
APY_E_ARG = -1
APY_E_HIP = -2
APY_E_WORKSPACE = -3
APY_E_SINGULAR = -4

_ERR = {
    APY_E_ARG: "invalid argument",
    APY_E_HIP: "HIP runtime error",
    APY_E_WORKSPACE: "workspace too small",
    APY_E_SINGULAR: "singular linear system",
}

APY_MK, APY_KM = 0, 1
APY_KN, APY_NK = 0, 1

c_i64 = ctypes.c_int64
c_int = ctypes.c_int
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p
c_size = ctypes.c_size_t


class GemmDesc(ctypes.Structure):
    _fields_ = [
        ("m", c_i64), ("n", c_i64), ("k", c_i64),
        ("a_layout", c_int), ("b_layout", c_int),
        ("conj_a", c_int), ("conj_b", c_int),
        ("alpha_re", c_dbl), ("alpha_im", c_dbl),
        ("beta_re", c_dbl), ("beta_im", c_dbl),
        ("a", c_vp), ("lda", c_i64), ("stride_a", c_i64),
        ("b", c_vp), ("ldb", c_i64), ("stride_b", c_i64),
        ("c", c_vp), ("ldc", c_i64), ("stride_c", c_i64),
        ("batch", c_i64),
        ("reduce_batch", c_int),
        ("split_k", c_int),
    ]


# name -> (restype, argtypes); every symbol here is declared in include/apyib_amd.h
SIGNATURES = {
    "apy_version": (c_int, []),
    "apy_zgemm_workspace": (c_size, [ctypes.POINTER(GemmDesc)]),
    "apy_zgemm": (c_int, [ctypes.POINTER(GemmDesc), c_vp, c_size, c_vp]),
}

_lib = None


class NativeError(RuntimeError):
    pass


def load():
    """Load the native library (once).  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeError(
            f"apyib_amd native library not found at {LIB_PATH}; "
            "build it with `python -m apyib_amd.build` (hipcc, gfx950)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status, what):
    if status != APY_OK:
        raise NativeError(f"{what} failed: {_ERR.get(status, status)} (status {status})")
