"""Python entry points for the LM normal-equation step and the scalar formal
solver (host arrays in and out)."""
import ctypes as C

import numpy as np

from . import abi
from ._native import lib, check


def lm_step(J, r, lm, precision=abi.SNAPI_FP64):
    """J [npix][nrow][npar], r [npix][nrow], lm [npix] -> (JTJ, rhs, corr),
    the damped solve of atmosphere::stokes_lm_fit (atmos/atmos_fit.cc:129-160)."""
    dt = np.float64 if precision == abi.SNAPI_FP64 else np.float32
    J = np.ascontiguousarray(J, dtype=dt)
    r = np.ascontiguousarray(r, dtype=dt)
    lm = np.ascontiguousarray(lm, dtype=np.float64)
    npix, nrow, npar = J.shape
    JTJ = np.zeros((npix, npar, npar))
    rhs = np.zeros((npix, npar))
    corr = np.zeros((npix, npar))
    rc = lib().snapi_lm_step(precision, npix, nrow, npar, abi.vptr(J), abi.vptr(r), abi.ptr(lm), abi.ptr(JTJ),
                             abi.ptr(rhs), abi.ptr(corr))
    check("snapi_lm_step", rc)
    return JTJ, rhs, corr


def formal_scalar(grid, op, em, ct, boundary=-1.0):
    """Batched scalar formal solution with Lambda*: arrays [nprob][nd],
    ct [nprob] = cos(theta) per problem."""
    grid = np.ascontiguousarray(grid, dtype=np.float64)
    op = np.ascontiguousarray(op, dtype=np.float64)
    em = np.ascontiguousarray(em, dtype=np.float64)
    ct = np.ascontiguousarray(ct, dtype=np.float64)
    nprob, nd = op.shape
    I = np.zeros((nprob, nd))
    L = np.zeros((nprob, nd))
    rc = lib().snapi_formal_scalar(nprob, nd, abi.ptr(grid), abi.ptr(op), abi.ptr(em), abi.ptr(ct), boundary,
                                   abi.ptr(I), abi.ptr(L))
    check("snapi_formal_scalar", rc)
    return I, L


def stokes_responses(table, ray, state, dstate, lam_vac, precision=abi.SNAPI_FP64):
    """Emergent Stokes [ncol][nlambda][4] and per-depth responses
    [ncol][7][nd][nlambda][4] (atmosphere::obs_stokes_responses,
    atmos/atmos_rad_synthesis.cc:166-399), by the adjoint sweep."""
    dt = np.float64 if precision == abi.SNAPI_FP64 else np.float32
    st = np.ascontiguousarray(state, dtype=dt)
    ds = np.ascontiguousarray(dstate, dtype=dt)
    if st.ndim == 2:
        st, ds = st[None], ds[None]
    ncol, nf, nd = st.shape
    lam = np.ascontiguousarray(lam_vac, dtype=np.float64)
    stokes = np.zeros((ncol, len(lam), 4), dtype=dt)
    resp = np.zeros((ncol, 7, nd, len(lam), 4), dtype=dt)
    rc = lib().snapi_stokes_responses(precision, C.byref(table), C.byref(ray), ncol, nd, abi.vptr(st),
                                      abi.vptr(ds), len(lam), abi.ptr(lam), abi.vptr(stokes), abi.vptr(resp))
    check("snapi_stokes_responses", rc)
    return stokes, resp


def stokes_responses_atm(table, ray, phi, bvec, state, dstate, lam_vac, precision=abi.SNAPI_FP64):
    """Per-depth responses for any ray (ray.theta, phi) in the atmosphere's
    frame, as atmosphere::obs_stokes_responses returns them: the state holds
    the ray-frame projections, bvec [ncol][3][nd] the atmosphere-frame field
    (Bx, By, Bz).  Returns (stokes, resp) as stokes_responses."""
    dt = np.float64 if precision == abi.SNAPI_FP64 else np.float32
    st = np.ascontiguousarray(state, dtype=dt)
    ds = np.ascontiguousarray(dstate, dtype=dt)
    bv = np.ascontiguousarray(bvec, dtype=np.float64)
    if st.ndim == 2:
        st, ds, bv = st[None], ds[None], bv[None]
    ncol, nf, nd = st.shape
    if bv.shape != (ncol, 3, nd):
        raise ValueError("bvec must be [ncol][3][nd]")
    lam = np.ascontiguousarray(lam_vac, dtype=np.float64)
    stokes = np.zeros((ncol, len(lam), 4), dtype=dt)
    resp = np.zeros((ncol, 7, nd, len(lam), 4), dtype=dt)
    rc = lib().snapi_stokes_responses_atm(precision, C.byref(table), C.byref(ray), C.c_double(phi), abi.ptr(bv),
                                          ncol, nd, abi.vptr(st), abi.vptr(ds), len(lam), abi.ptr(lam),
                                          abi.vptr(stokes), abi.vptr(resp))
    check("snapi_stokes_responses_atm", rc)
    return stokes, resp


def transform_responses(resp, bvec, theta, phi):
    """atmosphere::transform_responses on responses [ncol][7][nd][nlambda][4]
    (float64: FP64 kernel, float32: the same in the mixed kernel); returns
    the transformed copy."""
    r = np.array(resp, order="C", copy=True)
    if r.dtype not in (np.float64, np.float32):
        raise ValueError("responses must be float64 or float32")
    prec = abi.SNAPI_FP64 if r.dtype == np.float64 else abi.SNAPI_MIXED
    ncol, _, nd, nl, _ = r.shape
    bv = np.ascontiguousarray(bvec, dtype=np.float64).reshape(ncol, 3, nd)
    rc = lib().snapi_transform_responses(prec, ncol, nd, nl, abi.ptr(bv), C.c_double(theta), C.c_double(phi),
                                         abi.vptr(r))
    check("snapi_transform_responses", rc)
    return r


def stokes_adjoint(table, ray, state, dstate, lam_vac, obs, wstokes=(1.0, 1.0, 1.0, 1.0),
                   precision=abi.SNAPI_FP64):
    """Gradient [ncol][7][nd] of chi^2/2 = sum w_s (S - S_obs)^2 / 2 with
    respect to the per-depth parameters, in one reverse sweep."""
    dt = np.float64 if precision == abi.SNAPI_FP64 else np.float32
    st = np.ascontiguousarray(state, dtype=dt)
    ds = np.ascontiguousarray(dstate, dtype=dt)
    if st.ndim == 2:
        st, ds = st[None], ds[None]
    ncol, nf, nd = st.shape
    lam = np.ascontiguousarray(lam_vac, dtype=np.float64)
    ob = np.ascontiguousarray(obs, dtype=dt).reshape(ncol, len(lam), 4)
    ws = np.ascontiguousarray(wstokes, dtype=np.float64)
    grad = np.zeros((ncol, 7, nd), dtype=dt)
    rc = lib().snapi_stokes_adjoint(precision, C.byref(table), C.byref(ray), ncol, nd, abi.vptr(st), abi.vptr(ds),
                                    len(lam), abi.ptr(lam), abi.vptr(ob), abi.ptr(ws), abi.vptr(grad))
    check("snapi_stokes_adjoint", rc)
    return grad


def optical_depth_scale(height, op, theta=0.0, referent=False):
    """tau [ncol][nd] of snapi_optical_depth_scale (atmos_ppbez::optical_depth_scale,
    plpar/atmos_ppbez_rts.cc:221-266); referent: -tau, compute_tau_referent."""
    h = np.ascontiguousarray(height, dtype=np.float64)
    o = np.ascontiguousarray(op, dtype=np.float64)
    if h.ndim == 1:
        h, o = h[None], o[None]
    if h.shape != o.shape:
        raise ValueError("height and opacity shapes differ: %s %s" % (h.shape, o.shape))
    tau = np.zeros_like(h)
    f = lib().snapi_optical_depth_scale
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double, C.c_int,
                  C.POINTER(C.c_double)]
    check("snapi_optical_depth_scale", f(h.shape[0], h.shape[1], abi.ptr(h), abi.ptr(o), float(theta),
                                         1 if referent else 0, abi.ptr(tau)))
    return tau

