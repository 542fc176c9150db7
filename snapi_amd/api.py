"""Python entry points over the C ABI (host arrays in, host arrays out, or
torch device tensors for the *_dev variants)."""
import ctypes as C

import numpy as np

from . import abi
from ._native import lib, check

_DT = {abi.SNAPI_FP64: np.float64, abi.SNAPI_MIXED: np.float32}


def synth_stokes(table, ray, state, lam_vac, precision=abi.SNAPI_FP64):
    """Emergent Stokes [ncol][nlambda][4] of a batch of columns
    (atmosphere::obs_stokes, atmos/atmos_obs.cc:16-87)."""
    dt = _DT[precision]
    st = np.ascontiguousarray(state, dtype=dt)
    if st.ndim == 2:
        st = st[None]
    ncol, nf, nd = st.shape
    if nf != abi.nfield(table.nline):
        raise ValueError("state has %d fields, line table needs %d" % (nf, abi.nfield(table.nline)))
    lam = np.ascontiguousarray(lam_vac, dtype=np.float64)
    out = np.zeros((ncol, len(lam), 4), dtype=dt)
    rc = lib().snapi_synth_stokes(precision, C.byref(table), C.byref(ray), ncol, nd, abi.vptr(st), len(lam),
                                  abi.ptr(lam), abi.vptr(out))
    check("snapi_synth_stokes", rc)
    return out


def synth_stokes_dev(table, ray, state_t, lam_t, out_t, precision=abi.SNAPI_FP64, stream=None):
    """Device variant on torch tensors already resident in HBM; launches on
    `stream` (a torch.cuda.Stream or None for the current stream)."""
    import torch
    ncol, nf, nd = state_t.shape
    nl = lam_t.shape[0]
    dt = torch.float64 if precision == abi.SNAPI_FP64 else torch.float32
    for t, d, shp in ((state_t, dt, (ncol, abi.nfield(table.nline), nd)), (lam_t, torch.float64, (nl,)),
                      (out_t, dt, (ncol, nl, 4))):
        if t.dtype != d or tuple(t.shape) != shp or not t.is_contiguous() or not t.is_cuda:
            raise ValueError("synth_stokes_dev: expected a contiguous device tensor %s of shape %s, got %s %s"
                             % (d, shp, t.dtype, tuple(t.shape)))
    s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
    rc = lib().snapi_synth_stokes_dev(precision, C.byref(table), C.byref(ray), ncol, nd,
                                      C.c_void_p(state_t.data_ptr()), lam_t.shape[0], C.c_void_p(lam_t.data_ptr()),
                                      C.c_void_p(out_t.data_ptr()), C.c_void_p(s))
    check("snapi_synth_stokes_dev", rc)
