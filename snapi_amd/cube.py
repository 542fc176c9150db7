"""Cube I/O (snapi_ana_*, snapi_cube_to_state_dev, snapi_stokes_to_cube_dev):
ANA f0 files as the reference reads and writes them (cmlib/ana_read.cc,
ana_write.cc), the 12 x NX x NY x NZ model cube of read_spinor3d
(atmos/atmos_fio.cc:79-127) to column states, and column spectra to the
job's (nx, ny, 4, nlambda) output cube (master/job.cc:590-597)."""
import ctypes as C

import numpy as np

from . import abi
from ._native import lib, check

_DT = {0: np.int8, 1: np.int16, 2: np.int32, 3: np.float32, 4: np.float64, 5: np.int64}
_TYPE = {np.dtype(v): k for k, v in _DT.items()}


def _decl():
    L = lib()
    if getattr(L, "_cube_declared", False):
        return L
    ip = C.POINTER(C.c_int)
    L.snapi_ana_info.argtypes = [C.c_char_p, ip, ip, ip, ip]
    L.snapi_ana_read.argtypes = [C.c_char_p, C.c_void_p, C.c_size_t]
    L.snapi_ana_write.argtypes = [C.c_char_p, C.c_void_p, C.c_int, C.c_int, ip, C.c_char_p]
    L.snapi_cube_to_state_dev.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                          C.c_double, C.c_void_p, C.c_void_p]
    L.snapi_stokes_to_cube_dev.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    for f in ("snapi_ana_info", "snapi_ana_read", "snapi_ana_write", "snapi_cube_to_state_dev",
              "snapi_stokes_to_cube_dev"):
        getattr(L, f).restype = C.c_int
    L._cube_declared = True
    return L


def ana_info(path):
    """(type, dims with the fastest first, compressed)."""
    L = _decl()
    t, nd, comp = C.c_int(), C.c_int(), C.c_int()
    dims = (C.c_int * 16)()
    check("snapi_ana_info", L.snapi_ana_info(path.encode(), C.byref(t), C.byref(nd), dims, C.byref(comp)))
    return t.value, [dims[i] for i in range(nd.value)], bool(comp.value)


def ana_read(path):
    """The file's array in C order: shape = its dims reversed."""
    L = _decl()
    t, dims, _ = ana_info(path)
    a = np.empty(tuple(reversed(dims)), dtype=_DT[t])
    check("snapi_ana_read", L.snapi_ana_read(path.encode(), a.ctypes.data, a.nbytes))
    return a


def ana_write(path, array, header=None):
    """Writes a C-order array (dims = its shape reversed) uncompressed."""
    L = _decl()
    a = np.ascontiguousarray(array)
    if a.dtype not in _TYPE:
        raise ValueError("ANA has no type for %s" % a.dtype)
    dims = (C.c_int * a.ndim)(*reversed(a.shape))
    check("snapi_ana_write", L.snapi_ana_write(path.encode(), a.ctypes.data, _TYPE[a.dtype], a.ndim, dims,
                                               None if header is None else header.encode()))


def read_spinor3d(path):
    """The model cube [12][nx][ny][nz] in double (read_file + transpose as
    atmosphere::read_spinor3d does)."""
    a = ana_read(path)
    if a.ndim != 4 or a.shape[0] != 12:
        raise ValueError("%s: not a 12-parameter 4-D cube (shape %s)" % (path, a.shape))
    if a.dtype not in (np.float32, np.float64):
        raise ValueError("%s: data type is not floating point" % path)
    return np.ascontiguousarray(a, dtype=np.float64)


def cube_to_state_dev(cube_t, nfield, state_t, theta=np.pi, phi=0.0, stream=None):
    """cube_t [12][nx][ny][nz] float64 device tensor -> state_t [nx*ny][nfield][nz]
    (float64: FP64, float32: mixed)."""
    import torch
    L = _decl()
    _, nx, ny, nz = cube_t.shape
    prec = abi.SNAPI_FP64 if state_t.dtype == torch.float64 else abi.SNAPI_MIXED
    if tuple(state_t.shape) != (nx * ny, nfield, nz) or cube_t.dtype != torch.float64:
        raise ValueError("cube_to_state_dev: shapes or types do not match")
    sp = None if stream is None else C.c_void_p(stream.cuda_stream)
    check("snapi_cube_to_state_dev", L.snapi_cube_to_state_dev(prec, cube_t.data_ptr(), nx, ny, nz, nfield, theta,
                                                               phi, state_t.data_ptr(), sp))


def stokes_to_cube_dev(stokes_t, nx, ny, cube_t, stream=None):
    """stokes_t [nx*ny][nlambda][4] -> cube_t [nx][ny][4][nlambda] float64."""
    import torch
    L = _decl()
    nl = stokes_t.shape[1]
    prec = abi.SNAPI_FP64 if stokes_t.dtype == torch.float64 else abi.SNAPI_MIXED
    sp = None if stream is None else C.c_void_p(stream.cuda_stream)
    check("snapi_stokes_to_cube_dev", L.snapi_stokes_to_cube_dev(prec, stokes_t.data_ptr(), nx, ny, nl,
                                                                 cube_t.data_ptr(), sp))
