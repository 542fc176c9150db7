"""ALI for one NLTE atom (snapi_ali_accumulate, snapi_ali_newpops,
snapi_ali_solve): atmosphere::nltepops (atmos/atmos_pop.cc:67-240) over a
batch of plane-parallel columns.

Arrays (double, C order):
  colstate [ncol][A_NF][nd]  T, vt, ne, referent opacity, RT grid, total population
  vproj    [ncol][nray][nd]  velocity along each ray
  gamma    [ncol][ntrans][nd] total damping of each line (bound-free rows unused)
  bgop, bgem [ncol][nray or 1][nd][nlam] population-independent opacity/emissivity
  coll     [ncol][nd][nlevel][nlevel] collision rate i -> j
  pops     [ncol][nd][nlevel]
"""
import ctypes as C

import numpy as np

from . import abi
from ._native import lib, check

_dp = C.POINTER(C.c_double)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _declare(L):
    if getattr(L, "_ali_declared", False):
        return
    AT, GR = C.POINTER(abi.SnapiAliAtom), C.POINTER(abi.SnapiAliGrid)
    L.snapi_ali_accumulate.argtypes = [C.c_int, AT, GR, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.snapi_ali_newpops.argtypes = [AT, C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_int, _dp, _dp]
    L.snapi_ali_solve.argtypes = [C.c_int, AT, GR, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double,
                                  C.c_int, _dp, C.POINTER(C.c_int), _dp]
    for f in (L.snapi_ali_accumulate, L.snapi_ali_newpops, L.snapi_ali_solve):
        f.restype = C.c_int
    L._ali_declared = True


def _shape(atom, grid, colstate):
    cs = _c(colstate)
    if cs.ndim == 2:
        cs = cs[None]
    ncol, nf, nd = cs.shape
    if nf != abi.A_NF:
        raise ValueError("colstate needs %d fields, got %d" % (abi.A_NF, nf))
    return cs, ncol, nd


def _check_inputs(atom, grid, ncol, nd, vproj, gamma, bgop, bgem):
    nr, nl, nt = grid.nray, grid.nlam, atom.ntrans
    nb = nr if grid.bg_per_ray else 1
    want = [(vproj, ncol * nr * nd), (gamma, ncol * nt * nd), (bgop, ncol * nb * nd * nl), (bgem, ncol * nb * nd * nl)]
    for a, n in want:
        if a.size != n:
            raise ValueError("ALI input of %d values where %d are needed" % (a.size, n))


def accumulate(atom, grid, colstate, vproj, gamma, bgop, bgem, pops, precision=abi.SNAPI_FP64):
    """acc [ncol][nd][ntrans][4]: J lower->upper, J upper->lower, Lambda*, norm."""
    L = lib()
    _declare(L)
    cs, ncol, nd = _shape(atom, grid, colstate)
    vp, ga, bo, be, pp = (_c(x) for x in (vproj, gamma, bgop, bgem, pops))
    _check_inputs(atom, grid, ncol, nd, vp, ga, bo, be)
    if pp.size != ncol * nd * atom.nlevel:
        raise ValueError("pops size")
    acc = np.zeros((ncol, nd, atom.ntrans, 4))
    rc = L.snapi_ali_accumulate(precision, C.byref(atom), C.byref(grid), ncol, nd, abi.ptr(cs), abi.ptr(vp),
                                abi.ptr(ga), abi.ptr(bo), abi.ptr(be), abi.ptr(pp), abi.ptr(acc))
    check("snapi_ali_accumulate", rc)
    return acc


def newpops(atom, colstate, coll, pops, acc, alo=1):
    """(newpops [ncol][nd][nlevel], relchange [ncol])."""
    L = lib()
    _declare(L)
    cs = _c(colstate)
    if cs.ndim == 2:
        cs = cs[None]
    ncol, nf, nd = cs.shape
    co, pp, ac = _c(coll), _c(pops), _c(acc)
    n = atom.nlevel
    if co.size != ncol * nd * n * n or pp.size != ncol * nd * n or ac.size != ncol * nd * atom.ntrans * 4:
        raise ValueError("ALI rate-solve input sizes")
    out = np.zeros((ncol, nd, n))
    rcv = np.zeros(ncol)
    rc = L.snapi_ali_newpops(C.byref(atom), ncol, nd, abi.ptr(cs), abi.ptr(co), abi.ptr(pp), abi.ptr(ac), int(alo),
                             abi.ptr(out), abi.ptr(rcv))
    check("snapi_ali_newpops", rc)
    return out, rcv


def solve(atom, grid, colstate, vproj, gamma, bgop, bgem, coll, pops, tol=1e-2, max_iter=100,
          precision=abi.SNAPI_FP64):
    """(pops, iterations [ncol], relchange [ncol]) after the iteration."""
    L = lib()
    _declare(L)
    cs, ncol, nd = _shape(atom, grid, colstate)
    vp, ga, bo, be, co = (_c(x) for x in (vproj, gamma, bgop, bgem, coll))
    _check_inputs(atom, grid, ncol, nd, vp, ga, bo, be)
    pp = _c(pops).reshape(ncol, nd, atom.nlevel).copy()
    if co.size != ncol * nd * atom.nlevel ** 2:
        raise ValueError("coll size")
    it = np.zeros(ncol, dtype=np.int32)
    rcv = np.zeros(ncol)
    rc = L.snapi_ali_solve(precision, C.byref(atom), C.byref(grid), ncol, nd, abi.ptr(cs), abi.ptr(vp), abi.ptr(ga),
                           abi.ptr(bo), abi.ptr(be), abi.ptr(co), tol, max_iter, abi.ptr(pp),
                           it.ctypes.data_as(C.POINTER(C.c_int)), abi.ptr(rcv))
    check("snapi_ali_solve", rc)
    return pp, it, rcv


def solve_dev(atom, grid, colstate_t, vproj_t, gamma_t, bgop_t, bgem_t, coll_t, pops_t, iters_t, relchange_t,
              tol=1e-4, max_iter=200, precision=abi.SNAPI_MIXED, stream=None):
    """snapi_ali_solve_dev on torch float64 device tensors (int32 for the
    iteration counts); the grid's arrays are host arrays."""
    import torch
    L = lib()
    _declare(L)
    f = L.snapi_ali_solve_dev
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.POINTER(abi.SnapiAliAtom), C.POINTER(abi.SnapiAliGrid), C.c_int, C.c_int] + \
        [C.c_void_p] * 6 + [C.c_double, C.c_int] + [C.c_void_p] * 4
    ncol, nf, nd = colstate_t.shape
    nb = grid.nray if grid.bg_per_ray else 1
    want = [(colstate_t, torch.float64, (ncol, abi.A_NF, nd)), (vproj_t, torch.float64, (ncol, grid.nray, nd)),
            (gamma_t, torch.float64, (ncol, atom.ntrans, nd)), (bgop_t, torch.float64, (ncol, nb, nd, grid.nlam)),
            (bgem_t, torch.float64, (ncol, nb, nd, grid.nlam)),
            (coll_t, torch.float64, (ncol, nd, atom.nlevel, atom.nlevel)),
            (pops_t, torch.float64, (ncol, nd, atom.nlevel)), (iters_t, torch.int32, (ncol,)),
            (relchange_t, torch.float64, (ncol,))]
    for t, d, shp in want:
        if t.dtype != d or tuple(t.shape) != shp or not t.is_contiguous() or not t.is_cuda:
            raise ValueError("solve_dev: expected a contiguous device tensor %s of shape %s, got %s %s"
                             % (d, shp, t.dtype, tuple(t.shape)))
    sp = None if stream is None else C.c_void_p(stream.cuda_stream)
    rc = f(precision, C.byref(atom), C.byref(grid), ncol, nd, colstate_t.data_ptr(), vproj_t.data_ptr(),
           gamma_t.data_ptr(), bgop_t.data_ptr(), bgem_t.data_ptr(), coll_t.data_ptr(), tol, max_iter,
           pops_t.data_ptr(), iters_t.data_ptr(), relchange_t.data_ptr(), sp)
    check("snapi_ali_solve_dev", rc)
