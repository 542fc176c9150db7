"""The two numerical kernels of the NLTE population responses
(snapi_formal_responses, snapi_response_solve):

* atmos_ppbez::formal_with_responses (atmos/plpar/atmos_ppbez_rts.cc:349-778):
  scalar formal solution of one ray and the response of every intensity to
  every opacity and emissivity;
* the Crout solves of atom::compute_nlte_population_responses
  (atmos/atmol/atom/atom_responses.cc:171-200) that give the level responses
  from the assembled (nd*nmap)^2 system and its right-hand sides.
"""
import ctypes as C

import numpy as np

from . import abi
from ._native import lib, check

_dp = C.POINTER(C.c_double)


def _declare(L):
    if getattr(L, "_nlteresp_declared", False):
        return
    L.snapi_formal_responses.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_double, _dp, _dp, _dp,
                                         _dp]
    L.snapi_formal_responses.restype = C.c_int
    L.snapi_formal_responses_dev.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 4 + [C.c_double] + \
        [C.c_void_p] * 5
    L.snapi_formal_responses_dev.restype = C.c_int
    L.snapi_response_solve.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
    L.snapi_response_solve.restype = C.c_int
    L._nlteresp_declared = True


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def formal_responses(grid, op, em, theta, boundary=-1.0):
    """grid [ncol][nd]; op, em [ncol][nprob][nd]; theta [nprob].
    Returns (I, L [ncol][nprob][nd], Rop, Rem [ncol][nprob][nd][nd])."""
    L = lib()
    _declare(L)
    g, o, e, th = _c(grid), _c(op), _c(em), _c(theta).ravel()
    if g.ndim == 1:
        g = g[None]
    if o.ndim == 2:
        o, e = o[None], e[None]
    ncol, nd = g.shape
    nprob = th.size
    if o.shape != (ncol, nprob, nd) or e.shape != o.shape:
        raise ValueError("op/em must be [ncol][nprob][nd] = %s, got %s" % ((ncol, nprob, nd), o.shape))
    I = np.zeros((ncol, nprob, nd))
    Ls = np.zeros_like(I)
    Rop = np.zeros((ncol, nprob, nd, nd))
    Rem = np.zeros_like(Rop)
    rc = L.snapi_formal_responses(ncol, nprob, nd, abi.ptr(g), abi.ptr(o), abi.ptr(e), abi.ptr(th), float(boundary),
                                  abi.ptr(I), abi.ptr(Ls), abi.ptr(Rop), abi.ptr(Rem))
    check("snapi_formal_responses", rc)
    return I, Ls, Rop, Rem


def formal_responses_dev(grid_t, op_t, em_t, theta_t, I_t, L_t, Rop_t, Rem_t, boundary=-1.0, stream=None):
    """snapi_formal_responses_dev on contiguous float64 torch device tensors."""
    import torch
    L = lib()
    _declare(L)
    ncol, nd = grid_t.shape
    nprob = theta_t.numel()
    want = [(grid_t, (ncol, nd)), (op_t, (ncol, nprob, nd)), (em_t, (ncol, nprob, nd)), (theta_t, (nprob,)),
            (I_t, (ncol, nprob, nd)), (L_t, (ncol, nprob, nd)), (Rop_t, (ncol, nprob, nd, nd)),
            (Rem_t, (ncol, nprob, nd, nd))]
    for t, shp in want:
        if t.dtype != torch.float64 or tuple(t.shape) != shp or not t.is_contiguous() or not t.is_cuda:
            raise ValueError("formal_responses_dev: expected a contiguous float64 device tensor of shape %s, got %s %s"
                             % (shp, t.dtype, tuple(t.shape)))
    sp = None if stream is None else C.c_void_p(stream.cuda_stream)
    rc = L.snapi_formal_responses_dev(ncol, nprob, nd, grid_t.data_ptr(), op_t.data_ptr(), em_t.data_ptr(),
                                      theta_t.data_ptr(), float(boundary), I_t.data_ptr(), L_t.data_ptr(),
                                      Rop_t.data_ptr(), Rem_t.data_ptr(), sp)
    check("snapi_formal_responses_dev", rc)


def response_solve(M, B):
    """M [nsys][n][n] (or [n][n]); B [nsys][nrhs][n] (or [nrhs][n]).
    Returns X with B's shape: X[r] solves M x = B[r]."""
    L = lib()
    _declare(L)
    M, B = _c(M), _c(B)
    one = M.ndim == 2
    if one:
        M, B = M[None], B[None]
    nsys, n, n2 = M.shape
    if n != n2 or B.shape[0] != nsys or B.shape[2] != n:
        raise ValueError("M must be [nsys][n][n] and B [nsys][nrhs][n]")
    X = np.zeros_like(B)
    rc = L.snapi_response_solve(nsys, n, B.shape[1], abi.ptr(M), abi.ptr(B), abi.ptr(X))
    check("snapi_response_solve", rc)
    return X[0] if one else X
