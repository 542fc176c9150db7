"""Batched LM inversion (snapi_invert) and the node model it uses."""
import ctypes as C

import numpy as np

from . import abi
from ._native import lib, check

PARAM = {"T": 1, "vt": 2, "vlos": 3, "B": 4, "theta": 5, "phi": 6}


class SnapiNodeModel(C.Structure):
    _fields_ = [("nnode", C.c_int), ("param", C.c_int * abi.MAX_NODES), ("logtau", C.c_double * abi.MAX_NODES)]


def node_model(nodes):
    """nodes: list of (parameter name, log tau) in the reference's order."""
    m = SnapiNodeModel()
    m.nnode = len(nodes)
    for i, (p, lt) in enumerate(nodes):
        m.param[i] = PARAM[p]
        m.logtau[i] = lt
    return m


def invert(table, ray, model, state, lam_vac, obs, nodes0, mask=None, wstokes=(1.0, 1.0, 1.0, 1.0), lm_start=1e3,
           max_iter=30, nslot=0, precision=abi.SNAPI_FP64, jacobian=0):
    """Returns (nodes [npix][nnode], chi2 [npix], iterations [npix], total pixel-iterations)."""
    dt = np.float64 if precision == abi.SNAPI_FP64 else np.float32
    st = np.ascontiguousarray(state, dtype=dt)
    npix, nf, nd = st.shape
    lam = np.ascontiguousarray(lam_vac, dtype=np.float64)
    ob = np.ascontiguousarray(obs, dtype=dt).reshape(npix, len(lam), 4)
    mk = None if mask is None else np.ascontiguousarray(mask, dtype=dt).reshape(npix, len(lam))
    nodes = np.ascontiguousarray(nodes0, dtype=np.float64).reshape(npix, model.nnode).copy()
    chi2 = np.zeros(npix)
    iters = np.zeros(npix, dtype=np.int32)
    total = C.c_longlong(0)
    ws = np.ascontiguousarray(wstokes, dtype=np.float64)
    L = lib()
    f = L.snapi_invert
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_int, C.POINTER(abi.SnapiLineTable), C.POINTER(abi.SnapiRay),
                  C.POINTER(SnapiNodeModel), C.c_int, C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_double),
                  C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.c_double, C.c_int, C.c_int,
                  C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_longlong)]
    rc = f(precision, jacobian, C.byref(table), C.byref(ray), C.byref(model), npix, nd, abi.vptr(st), len(lam),
           abi.ptr(lam), abi.vptr(ob), abi.vptr(mk) if mk is not None else None, abi.ptr(ws), lm_start, max_iter, nslot,
           abi.ptr(nodes), abi.ptr(chi2), iters.ctypes.data_as(C.POINTER(C.c_int)), C.byref(total))
    check("snapi_invert", rc)
    return nodes, chi2, iters, total.value
