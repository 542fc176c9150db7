"""Instrument effects: the Gaussian spectral convolution
(snapi_spectral_convolve[_dev]) and the sampled-PSF convolution
(snapi_psf_convolve[_dev]) the reference's fit applies to the spectrum and
to the node Jacobian (convolve_spectra_with_gauss / _psf,
convolve_response_with_gauss / _psf, cmlib/mathtools.cc:1350-1522;
atmos/atmos_fit.cc:104-115), and the scattered-light offset
(snapi_add_scattered_light[_dev], observable::add_scattered_light,
cmlib/obs.cc:412-428)."""
import ctypes as C

import numpy as np

from . import abi
from ._native import lib, check


def spectral_convolve(data, lam_vac, width, precision=abi.SNAPI_FP64):
    """data [nseries][nlambda][ncomp] (a Stokes array [npix][nlambda][4], or
    a Jacobian [npix][nlambda*4][nnode] viewed as [npix][nlambda][4*nnode]);
    width [cm].  Returns the convolved copy."""
    dt = np.float64 if precision == abi.SNAPI_FP64 else np.float32
    d = np.array(data, dtype=dt, order="C", copy=True)
    if d.ndim == 2:
        d = d[None]
    ns, nl, nc = d.shape
    lam = np.ascontiguousarray(lam_vac, dtype=np.float64)
    if lam.shape != (nl,):
        raise ValueError("lam must have %d wavelengths" % nl)
    f = lib().snapi_spectral_convolve
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), C.c_double, C.c_void_p]
    check("snapi_spectral_convolve", f(precision, ns, nl, nc, abi.ptr(lam), float(width), abi.vptr(d)))
    return d


def spectral_convolve_dev(data_t, lam_vac, lam_t, width, stream=None):
    """In place on a contiguous device tensor [nseries][nlambda][ncomp]
    (float64 or float32); lam_vac the host copy of lam_t."""
    import torch
    prec = abi.SNAPI_FP64 if data_t.dtype == torch.float64 else abi.SNAPI_MIXED
    ns, nl, nc = data_t.shape
    lam = np.ascontiguousarray(lam_vac, dtype=np.float64)
    f = lib().snapi_spectral_convolve_dev
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), C.c_void_p, C.c_double, C.c_void_p,
                  C.c_void_p]
    sp = None if stream is None else C.c_void_p(stream.cuda_stream)
    check("snapi_spectral_convolve_dev", f(prec, ns, nl, nc, abi.ptr(lam), lam_t.data_ptr(), float(width),
                                           data_t.data_ptr(), sp))


def psf_convolve(data, psf, precision=abi.SNAPI_FP64):
    """data [nseries][nlambda][ncomp] as for spectral_convolve; psf the
    samples on the wavelength step.  Returns the convolved copy."""
    dt = np.float64 if precision == abi.SNAPI_FP64 else np.float32
    d = np.array(data, dtype=dt, order="C", copy=True)
    if d.ndim == 2:
        d = d[None]
    ns, nl, nc = d.shape
    p = np.ascontiguousarray(psf, dtype=np.float64).reshape(-1)
    f = lib().snapi_psf_convolve
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), C.c_void_p]
    check("snapi_psf_convolve", f(precision, ns, nl, nc, p.size, abi.ptr(p), abi.vptr(d)))
    return d


def psf_convolve_dev(data_t, psf_t, stream=None):
    """In place on a contiguous device tensor [nseries][nlambda][ncomp]
    (float64 or float32); psf_t the samples as a float64 device tensor."""
    import torch
    prec = abi.SNAPI_FP64 if data_t.dtype == torch.float64 else abi.SNAPI_MIXED
    ns, nl, nc = data_t.shape
    f = lib().snapi_psf_convolve_dev
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    sp = None if stream is None else C.c_void_p(stream.cuda_stream)
    check("snapi_psf_convolve_dev", f(prec, ns, nl, nc, psf_t.numel(), psf_t.data_ptr(), data_t.data_ptr(), sp))


def add_scattered_light(stokes, fraction, continuum_level, precision=abi.SNAPI_FP64):
    """stokes [nseries][nlambda][4]: returns the copy with Stokes I raised by
    fraction * continuum_level."""
    dt = np.float64 if precision == abi.SNAPI_FP64 else np.float32
    d = np.array(stokes, dtype=dt, order="C", copy=True)
    if d.ndim == 2:
        d = d[None]
    ns, nl, _ = d.shape
    f = lib().snapi_add_scattered_light
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_void_p]
    check("snapi_add_scattered_light", f(precision, ns, nl, float(fraction), float(continuum_level), abi.vptr(d)))
    return d


def add_scattered_light_dev(stokes_t, fraction, continuum_level, stream=None):
    """In place on a contiguous device tensor [nseries][nlambda][4]."""
    import torch
    prec = abi.SNAPI_FP64 if stokes_t.dtype == torch.float64 else abi.SNAPI_MIXED
    ns, nl, _ = stokes_t.shape
    f = lib().snapi_add_scattered_light_dev
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_void_p, C.c_void_p]
    sp = None if stream is None else C.c_void_p(stream.cuda_stream)
    check("snapi_add_scattered_light_dev", f(prec, ns, nl, float(fraction), float(continuum_level),
                                             stokes_t.data_ptr(), sp))
