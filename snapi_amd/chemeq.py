"""Chemical equilibrium and LTE populations (snapi_chemeq, snapi_chemeq_dev):
atmosphere::ltepops (atmos/atmos_pop.cc:30-45) -- atmosphere::chemeq
(atmos/atmos_chm.cc:28-363) then atom::lte (atmos/atmol/atom/atom.cc:779-787)
-- at every depth point of a batch of columns, in double.

The species table is an abi.SnapiChemTable (abi.make_chem_table builds one
from per-species arrays).  Arrays (double, C order):
  T, Nt       [ncol][nd]            temperature [K], total particle density [cm^-3]
  ne          [ncol][nd]            electron density
  stage       [ncol][nstage][nd]    every ionisation stage of every element, table order
  molecule    [ncol][nmolecule][nd]
  level       [ncol][nlevel][nd]    LTE populations of the table's levels
  opref       [ncol][nd]            continuum opacity at the table's lambda_ref
                                    (atmosphere::opacity_continuum, atmos/atmos_rad.cc:286-294)
  iterations  [ncol][nd]            Newton steps taken (-1: singular system)

Equation-of-state fields of a column state (snapi_eos_state[_dev]): from T
and NT, the synthesis path's Ne, neutral hydrogen, referent opacity,
background and line level populations, and the response path's state and
thermodynamic derivatives (atmos/atmos_rad_synthesis.cc:173-190).

Hydrostatic equilibrium (snapi_hequilibrium, snapi_hequilibrium_dev):
atmos_pp::enforce_hequilibrium (atmos/plpar/atmos_pp.cc:372-455) on columns
of temperature and log referent optical depth, top first; Nt, ne, opref and
rho [ncol][nd] come out at the final state, iterations counts the HE
iterations at each depth.
"""
import ctypes as C

import numpy as np

from . import abi
from ._native import lib, check

_dp = C.POINTER(C.c_double)


def _declare(L):
    if getattr(L, "_chem_declared", False):
        return
    TB = C.POINTER(abi.SnapiChemTable)
    L.snapi_chemeq.argtypes = [TB, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.POINTER(C.c_int)]
    L.snapi_chemeq.restype = C.c_int
    L.snapi_chemeq_dev.argtypes = [TB, C.c_int, C.c_int] + [C.c_void_p] * 9
    L.snapi_chemeq_dev.restype = C.c_int
    L.snapi_hequilibrium.argtypes = [TB, C.c_int, C.c_int] + [_dp] * 9 + [C.POINTER(C.c_int)]
    L.snapi_hequilibrium.restype = C.c_int
    L.snapi_hequilibrium_dev.argtypes = [TB, C.c_int, C.c_int] + [C.c_void_p] * 11
    L.snapi_hequilibrium_dev.restype = C.c_int
    L.snapi_hequilibrium_tangent.argtypes = [TB, C.c_int, C.c_int, _dp, _dp, _dp, C.c_int, _dp, _dp]
    L.snapi_hequilibrium_tangent.restype = C.c_int
    SM = C.POINTER(abi.SnapiStateMap)
    L.snapi_eos_state.argtypes = [TB, SM, C.c_int, C.c_int, _dp, _dp, _dp]
    L.snapi_eos_state.restype = C.c_int
    L.snapi_eos_state_dev.argtypes = [TB, SM, C.c_int, C.c_int] + [C.c_void_p] * 4
    L.snapi_eos_state_dev.restype = C.c_int
    L._chem_declared = True


def _shape(T, Nt):
    T = np.ascontiguousarray(T, dtype=np.float64)
    Nt = np.ascontiguousarray(Nt, dtype=np.float64)
    if T.ndim == 1:
        T, Nt = T[None], Nt[None]
    if T.shape != Nt.shape or T.ndim != 2:
        raise ValueError("T and Nt must both be [ncol][nd], got %s and %s" % (T.shape, Nt.shape))
    return T, Nt


def chemeq(table, T, Nt):
    """dict(ne, stage, molecule, level, opref, iterations) for host arrays T, Nt;
    opref is None when the table has no referent wavelength."""
    L = lib()
    _declare(L)
    T, Nt = _shape(T, Nt)
    ncol, nd = T.shape
    ne = np.zeros((ncol, nd))
    st = np.zeros((ncol, abi.chem_stage_count(table), nd))
    mo = np.zeros((ncol, table.nmolecule, nd))
    lv = np.zeros((ncol, table.nlevel, nd))
    it = np.zeros((ncol, nd), dtype=np.int32)
    kr = np.zeros((ncol, nd)) if table.lambda_ref > 0 else None
    rc = L.snapi_chemeq(C.byref(table), ncol, nd, abi.ptr(T), abi.ptr(Nt), abi.ptr(ne), abi.ptr(st),
                        abi.ptr(mo) if table.nmolecule else None, abi.ptr(lv) if table.nlevel else None,
                        None if kr is None else abi.ptr(kr), it.ctypes.data_as(C.POINTER(C.c_int)))
    check("snapi_chemeq", rc)
    return dict(ne=ne, stage=st, molecule=mo, level=lv, opref=kr, iterations=it)


def chemeq_dev(table, T_t, Nt_t, ne_t, stage_t=None, molecule_t=None, level_t=None, iterations_t=None,
               stream=None, opref_t=None):
    """snapi_chemeq_dev on torch device tensors (float64; int32 iterations);
    optional outputs may be None."""
    L = lib()
    _declare(L)
    import torch
    if T_t.dim() != 2:
        raise ValueError("T must be [ncol][nd]")
    ncol, nd = T_t.shape
    want = [(T_t, torch.float64, (ncol, nd)), (Nt_t, torch.float64, (ncol, nd)), (ne_t, torch.float64, (ncol, nd)),
            (stage_t, torch.float64, (ncol, abi.chem_stage_count(table), nd)),
            (molecule_t, torch.float64, (ncol, table.nmolecule, nd)),
            (level_t, torch.float64, (ncol, table.nlevel, nd)), (opref_t, torch.float64, (ncol, nd)),
            (iterations_t, torch.int32, (ncol, nd))]
    for t, d, shp in want:
        if t is None:
            continue
        if t.dtype != d or tuple(t.shape) != shp or not t.is_contiguous() or not t.is_cuda:
            raise ValueError("chemeq_dev: expected a contiguous device tensor %s of shape %s, got %s %s"
                             % (d, shp, t.dtype, tuple(t.shape)))
    ptrs = [None if t is None else t.data_ptr() for t, _, _ in want]
    sp = None if stream is None else C.c_void_p(stream.cuda_stream)
    rc = L.snapi_chemeq_dev(C.byref(table), ncol, nd, *ptrs, sp)
    check("snapi_chemeq_dev", rc)


def hequilibrium(table, T, logtau, populations=False):
    """dict(Nt, ne, opref, rho, iterations) for host arrays T, logtau
    [ncol][nd] (top first); with populations also stage, molecule, level."""
    L = lib()
    _declare(L)
    T, logtau = _shape(T, logtau)
    ncol, nd = T.shape
    out = {k: np.zeros((ncol, nd)) for k in ("Nt", "ne", "opref", "rho")}
    out["iterations"] = np.zeros((ncol, nd), dtype=np.int32)
    st = mo = lv = None
    if populations:
        st = out["stage"] = np.zeros((ncol, abi.chem_stage_count(table), nd))
        mo = out["molecule"] = np.zeros((ncol, table.nmolecule, nd))
        lv = out["level"] = np.zeros((ncol, table.nlevel, nd))
    rc = L.snapi_hequilibrium(C.byref(table), ncol, nd, abi.ptr(T), abi.ptr(logtau), abi.ptr(out["Nt"]),
                              abi.ptr(out["ne"]), abi.ptr(out["opref"]), abi.ptr(out["rho"]),
                              None if st is None else abi.ptr(st),
                              abi.ptr(mo) if (mo is not None and table.nmolecule) else None,
                              abi.ptr(lv) if (lv is not None and table.nlevel) else None,
                              out["iterations"].ctypes.data_as(C.POINTER(C.c_int)))
    check("snapi_hequilibrium", rc)
    return out


def hequilibrium_tangent(table, T, logtau, Nt, dT):
    """snapi_hequilibrium_tangent: dNt [ncol][ndir][nd] of columns in
    hydrostatic equilibrium (T, logtau, Nt [ncol][nd]) along the temperature
    perturbations dT [ncol][ndir][nd]."""
    L = lib()
    _declare(L)
    T, logtau = _shape(T, logtau)
    ncol, nd = T.shape
    Nt = np.ascontiguousarray(Nt, dtype=np.float64).reshape(ncol, nd)
    dT = np.ascontiguousarray(dT, dtype=np.float64).reshape(ncol, -1, nd)
    dN = np.zeros_like(dT)
    rc = L.snapi_hequilibrium_tangent(C.byref(table), ncol, nd, abi.ptr(T), abi.ptr(logtau), abi.ptr(Nt),
                                      dT.shape[1], abi.ptr(dT), abi.ptr(dN))
    check("snapi_hequilibrium_tangent", rc)
    return dN


def hequilibrium_dev(table, T_t, logtau_t, Nt_t, ne_t=None, opref_t=None, rho_t=None, stage_t=None,
                     molecule_t=None, level_t=None, iterations_t=None, stream=None):
    """snapi_hequilibrium_dev on torch device tensors (float64; int32
    iterations); optional outputs may be None."""
    L = lib()
    _declare(L)
    import torch
    if T_t.dim() != 2:
        raise ValueError("T must be [ncol][nd]")
    ncol, nd = T_t.shape
    pt = (ncol, nd)
    want = [(T_t, torch.float64, pt), (logtau_t, torch.float64, pt), (Nt_t, torch.float64, pt),
            (ne_t, torch.float64, pt), (opref_t, torch.float64, pt), (rho_t, torch.float64, pt),
            (stage_t, torch.float64, (ncol, abi.chem_stage_count(table), nd)),
            (molecule_t, torch.float64, (ncol, table.nmolecule, nd)),
            (level_t, torch.float64, (ncol, table.nlevel, nd)), (iterations_t, torch.int32, pt)]
    for t, d, shp in want:
        if t is None:
            continue
        if t.dtype != d or tuple(t.shape) != shp or not t.is_contiguous() or not t.is_cuda:
            raise ValueError("hequilibrium_dev: expected a contiguous device tensor %s of shape %s, got %s %s"
                             % (d, shp, t.dtype, tuple(t.shape)))
    ptrs = [None if t is None else t.data_ptr() for t, _, _ in want]
    sp = None if stream is None else C.c_void_p(stream.cuda_stream)
    rc = L.snapi_hequilibrium_dev(C.byref(table), ncol, nd, *ptrs, sp)
    check("snapi_hequilibrium_dev", rc)


def eos_state(table, smap, state, derivatives=True):
    """state [ncol][nfield][nd] (T, NT set) -> (state with the equation-of-state
    fields filled, the response path's state, dstate [ncol][2][nfield][nd]);
    the last two are None without derivatives."""
    L = lib()
    _declare(L)
    st = np.array(state, dtype=np.float64, order="C", copy=True)
    if st.ndim != 3 or st.shape[1] != abi.nfield(smap.nline):
        raise ValueError("state must be [ncol][%d][nd]" % abi.nfield(smap.nline))
    ncol, nf, nd = st.shape
    resp = np.zeros_like(st) if derivatives else None
    ds = np.zeros((ncol, 2, nf, nd)) if derivatives else None
    rc = L.snapi_eos_state(C.byref(table), C.byref(smap), ncol, nd, abi.ptr(st),
                           None if resp is None else abi.ptr(resp), None if ds is None else abi.ptr(ds))
    check("snapi_eos_state", rc)
    return st, resp, ds


def eos_state_dev(table, smap, state_t, resp_t=None, dstate_t=None, stream=None):
    """snapi_eos_state_dev on float64 device tensors: state_t [ncol][nfield][nd]
    updated in place, resp_t like it, dstate_t [ncol][2][nfield][nd]."""
    L = lib()
    _declare(L)
    ncol, nf, nd = state_t.shape
    for t, shp in ((state_t, (ncol, nf, nd)), (resp_t, (ncol, nf, nd)), (dstate_t, (ncol, 2, nf, nd))):
        if t is None:
            continue
        import torch
        if t.dtype != torch.float64 or tuple(t.shape) != shp or not t.is_contiguous() or not t.is_cuda:
            raise ValueError("eos_state_dev: expected a contiguous float64 device tensor of shape %s" % (shp,))
    if nf != abi.nfield(smap.nline):
        raise ValueError("state must have %d fields" % abi.nfield(smap.nline))
    sp = None if stream is None else C.c_void_p(stream.cuda_stream)
    rc = L.snapi_eos_state_dev(C.byref(table), C.byref(smap), ncol, nd, state_t.data_ptr(),
                               None if resp_t is None else resp_t.data_ptr(),
                               None if dstate_t is None else dstate_t.data_ptr(), sp)
    check("snapi_eos_state_dev", rc)
