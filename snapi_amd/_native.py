"""ctypes binding of the HIP library (snapi_amd/libsnapi_gpu.so).

The library is required: there is no CPU fallback in the product path, and
loading fails with an error that says how to build it.
"""
import ctypes as C
import os

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsnapi_gpu.so")
_lib = None


class SnapiError(RuntimeError):
    def __init__(self, func, code):
        super().__init__("%s failed: %s (%d)" % (func, abi.STATUS_TEXT.get(code, "unknown"), code))
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("HIP library %s is missing: build it with `python -m snapi_amd.build` "
                               "(or __graft_entry__.build())" % LIB_PATH)
        # torch ships its own HIP runtime under the same soname; loaded first,
        # it is the one the library binds to, so that torch device tensors and
        # the library share one runtime.  Loaded after the library, torch would
        # bring up a second runtime that finds no device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.c_int
        LT, RY = C.POINTER(abi.SnapiLineTable), C.POINTER(abi.SnapiRay)
        L.snapi_version.restype = ip
        L.snapi_device_count.restype = ip
        L.snapi_synth_stokes.restype = ip
        L.snapi_synth_stokes.argtypes = [ip, LT, RY, ip, ip, vp, ip, dp, vp]
        L.snapi_synth_stokes_dev.restype = ip
        L.snapi_synth_stokes_dev.argtypes = [ip, LT, RY, ip, ip, vp, ip, vp, vp, vp]
        for name, args in (
                ("snapi_stokes_responses", [ip, LT, RY, ip, ip, vp, vp, ip, dp, vp, vp]),
                ("snapi_stokes_adjoint", [ip, LT, RY, ip, ip, vp, vp, ip, dp, vp, dp, vp]),
                ("snapi_lm_step", [ip, ip, ip, ip, vp, vp, dp, dp, dp, dp]),
                ("snapi_formal_scalar", [ip, ip, dp, dp, dp, dp, C.c_double, dp, dp])):
            if hasattr(L, name):
                f = getattr(L, name)
                f.restype = ip
                f.argtypes = args
        _lib = L
    return _lib


def check(func, rc):
    if rc != abi.SNAPI_OK:
        raise SnapiError(func, rc)
