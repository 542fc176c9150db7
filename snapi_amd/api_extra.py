"""Python entry points for the LM normal-equation step and the scalar formal
solver (host arrays in and out)."""
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
