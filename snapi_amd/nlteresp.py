"""The two numerical kernels of the NLTE population responses
(snapi_formal_responses, snapi_response_solve):

* atmos_ppbez::formal_with_responses (atmos/plpar/atmos_ppbez_rts.cc:349-778):
  scalar formal solution of one ray and the response of every intensity to
  every opacity and emissivity;
* the Crout solves of atom::compute_nlte_population_responses
  (atmos/atmol/atom/atom_responses.cc:171-200) that give the level responses
  from the assembled (nd*nmap)^2 system and its right-hand sides;

and the whole system of one NLTE atom (snapi_nlte_response_system[_dev]):
atmosphere::compute_nlte_population_responses (atmos/atmos_responses.cc:
13-157) with atom::compute_profile_norm, atom::add_response_contributions_new
(atom_responses.cc:668-1128) and the rate part of
atom::compute_nlte_population_responses (atom_responses.cc:26-170).
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
    L.snapi_nlte_response_system.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double] + \
        [_dp] * 14
    L.snapi_nlte_response_system.restype = C.c_int
    L.snapi_nlte_response_system_dev.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                                 C.c_double, _dp, _dp] + [C.c_void_p] * 13
    L.snapi_nlte_response_system_dev.restype = C.c_int
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


_SYSTEM_INPUTS = ("sigma", "depth", "pops", "damp", "coll", "vr", "opem")


def _system_shapes(atom, ncol, nd, nray, nlam):
    nm, nt = atom.nmap, atom.ntr
    return {"sigma": (nt, nlam), "depth": (ncol, abi.SNAPI_NLTE_NFIELD, nd), "pops": (ncol, nm, nd),
            "damp": (ncol, 2, nt, nd), "coll": (ncol, 3, nm, nm, nd), "vr": (ncol, nray, nd),
            "opem": (ncol, nray, nlam, 8, nd)}


def response_system(atom, theta, wray, lam, lamw, sigma, depth, pops, damp, coll, vr, opem, tau_grid=1,
                    boundary=-1.0, solve=False):
    """The response system of ncol columns for one NLTE atom
    (abi.make_nlte_atom).  theta, wray [nray]; lam, lamw [nlam]; sigma
    [ntr][nlam]; per column depth [SNAPI_NLTE_NFIELD][nd], pops [nmap][nd],
    damp [2][ntr][nd], coll [3][nmap][nmap][nd], vr [nray][nd], opem
    [nray][nlam][8][nd] (a leading column axis, or none for one column).
    Returns (M [ncol][n][n], beta [ncol][3][nd][n]) and, with solve, the level
    responses [ncol][3][n][nd]; without a column axis in, none out."""
    L = lib()
    _declare(L)
    one = np.ndim(depth) == 2
    arrs = [_c(x) for x in (sigma, depth, pops, damp, coll, vr, opem)]
    if one:
        arrs = [arrs[0]] + [a[None] for a in arrs[1:]]
    th, wr, lm, lw = (_c(x).ravel() for x in (theta, wray, lam, lamw))
    ncol, nd = arrs[1].shape[0], arrs[1].shape[2]
    nray, nlam = th.size, lm.size
    want = _system_shapes(atom, ncol, nd, nray, nlam)
    for k, a in zip(_SYSTEM_INPUTS, arrs):
        if a.shape != want[k]:
            raise ValueError("%s must be %s, got %s" % (k, want[k], a.shape))
    if wr.size != nray or lw.size != nlam:
        raise ValueError("ray_weight and lambda_weight must match theta and lambda")
    n = nd * atom.nmap
    M = np.zeros((ncol, n, n))
    beta = np.zeros((ncol, 3, nd, n))
    resp = np.zeros((ncol, 3, n, nd)) if solve else None
    rc = L.snapi_nlte_response_system(C.addressof(atom), ncol, nd, nray, nlam, int(tau_grid), float(boundary),
                                      abi.ptr(th), abi.ptr(wr), abi.ptr(lm), abi.ptr(lw),
                                      *[abi.ptr(a) for a in arrs], abi.ptr(M), abi.ptr(beta),
                                      abi.ptr(resp) if solve else None)
    check("snapi_nlte_response_system", rc)
    if one:
        return (M[0], beta[0], resp[0]) if solve else (M[0], beta[0])
    return (M, beta, resp) if solve else (M, beta)


def response_system_dev(atom, theta, wray, lam_t, lamw_t, tensors, M_t, beta_t, resp_t=None, tau_grid=1,
                        boundary=-1.0, stream=None):
    """snapi_nlte_response_system_dev: theta, wray host arrays [nray];
    lam_t, lamw_t and the tensors dict (sigma, depth, pops, damp, coll, vr,
    opem with their column axis) contiguous float64 device tensors; M_t,
    beta_t (and resp_t) device outputs."""
    import torch
    L = lib()
    _declare(L)
    th, wr = _c(theta).ravel(), _c(wray).ravel()
    ncol, _, nd = tensors["depth"].shape
    nray, nlam = th.size, lam_t.numel()
    want = _system_shapes(atom, ncol, nd, nray, nlam)
    n = nd * atom.nmap
    want.update(M=(ncol, n, n), beta=(ncol, 3, nd, n))
    outs = {"M": M_t, "beta": beta_t}
    if resp_t is not None:
        want["resp"] = (ncol, 3, n, nd)
        outs["resp"] = resp_t
    for k, t in list(tensors.items()) + list(outs.items()):
        if t.dtype != torch.float64 or tuple(t.shape) != want[k] or not t.is_contiguous() or not t.is_cuda:
            raise ValueError("response_system_dev: %s must be a contiguous float64 device tensor of shape %s"
                             % (k, want[k]))
    sp = None if stream is None else C.c_void_p(stream.cuda_stream)
    rc = L.snapi_nlte_response_system_dev(C.addressof(atom), ncol, nd, nray, nlam, int(tau_grid), float(boundary),
                                          abi.ptr(th), abi.ptr(wr), lam_t.data_ptr(), lamw_t.data_ptr(),
                                          *[tensors[k].data_ptr() for k in _SYSTEM_INPUTS], M_t.data_ptr(),
                                          beta_t.data_ptr(), None if resp_t is None else resp_t.data_ptr(), sp)
    check("snapi_nlte_response_system_dev", rc)
