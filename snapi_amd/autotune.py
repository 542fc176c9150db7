"""Launch-shape autotuner of the library (snapi_autotune_*): the synthesis
kernels and the float ALI accumulation in three occupancy bounds each, the
node-Jacobian tangent kernels in three workgroup sizes, timed on the first
call of each problem shape and the fastest kept for the process.  Every
shape gives the same results."""
import ctypes as C
import json

from ._native import lib, check

SYNTH, TANGENT, ALI = 0, 1, 2


def _decl():
    L = lib()
    if getattr(L, "_tune_declared", False):
        return L
    L.snapi_autotune_mode.argtypes = [C.c_int]
    L.snapi_autotune_force.argtypes = [C.c_int, C.c_int]
    L.snapi_autotune_clear.argtypes = []
    L.snapi_autotune_report.argtypes = [C.c_char_p, C.c_size_t]
    for f in ("snapi_autotune_mode", "snapi_autotune_force", "snapi_autotune_clear", "snapi_autotune_report"):
        getattr(L, f).restype = C.c_int
    L._tune_declared = True
    return L


def mode(m=-1):
    """Sets tuning on (1) or off (0); -1 only queries.  Returns the previous mode."""
    r = _decl().snapi_autotune_mode(int(m))
    if r < 0:
        check("snapi_autotune_mode", r)
    return r


def force(kernel, candidate):
    """Launches candidate `candidate` of `kernel` (SYNTH, TANGENT or ALI)
    from now on; -1 returns to the tuned choice."""
    check("snapi_autotune_force", _decl().snapi_autotune_force(int(kernel), int(candidate)))


def clear():
    check("snapi_autotune_clear", _decl().snapi_autotune_clear())


def report():
    """The choices made so far: a list of dicts (device, kernel, precision,
    nd, nline, ncomp, nlambda, extra, choice, ms per candidate)."""
    buf = C.create_string_buffer(1 << 16)
    check("snapi_autotune_report", _decl().snapi_autotune_report(buf, len(buf)))
    return [json.loads(x) for x in buf.value.decode().splitlines() if x.strip()]
