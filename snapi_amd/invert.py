"""Batched LM inversion (snapi_invert, snapi_invert_dev) and the node model it
uses.  lm_start defaults to the reference's observable start_lambda (1.0,
cmlib/obs.cc:50)."""
import ctypes as C

import numpy as np

from . import abi
from ._native import lib, check

PARAM = {"T": 1, "vt": 2, "vlos": 3, "B": 4, "theta": 5, "phi": 6}


class SnapiNodeModel(C.Structure):
    _fields_ = [("nnode", C.c_int), ("param", C.c_int * abi.MAX_NODES), ("logtau", C.c_double * abi.MAX_NODES),
                ("reg_type", C.c_int * 7), ("reg_alpha", C.c_double * 7)]


class SnapiInstrument(C.Structure):
    _fields_ = [("broadening", C.c_double), ("npsf", C.c_int), ("psf", C.POINTER(C.c_double))]


REG_NONE, REG_ZERO, REG_TIKHONOV = 0, 1, 2


def node_model(nodes, reg=None):
    """nodes: list of (parameter name, log tau) in the reference's order;
    reg: {parameter name: (type, alpha)} with type REG_ZERO (departure from
    zero) or REG_TIKHONOV (model::*_reg_type / *_reg_alpha)."""
    m = SnapiNodeModel()
    m.nnode = len(nodes)
    for i, (p, lt) in enumerate(nodes):
        m.param[i] = PARAM[p]
        m.logtau[i] = lt
    for p, (t, a) in (reg or {}).items():
        m.reg_type[PARAM[p]] = int(t)
        m.reg_alpha[PARAM[p]] = float(a)
    return m


def instrument(broadening=0.0, psf=None):
    """snapi_instrument of a fit (None when there is nothing to apply); the
    PSF samples are kept alive on the returned structure."""
    if not broadening and psf is None:
        return None
    inst = SnapiInstrument()
    inst.broadening = float(broadening)
    if psf is not None:
        p = np.ascontiguousarray(psf, dtype=np.float64)
        inst.npsf = p.size
        inst.psf = abi.ptr(p)
        inst._keep = p
    return inst


def _inst_ptr(inst):
    return None if inst is None else C.byref(inst)


def invert(table, ray, model, state, lam_vac, obs, nodes0, mask=None, wstokes=(1.0, 1.0, 1.0, 1.0), lm_start=1.0,
           max_iter=30, nslot=0, precision=abi.SNAPI_FP64, jacobian=0):
    """Returns (nodes [npix][nnode], chi2 [npix], iterations [npix], total pixel-iterations)."""
    dt = np.float64 if precision == abi.SNAPI_FP64 else np.float32
    st = np.ascontiguousarray(state, dtype=dt)
    npix, nf, nd = st.shape
    if nf != abi.nfield(table.nline):
        raise ValueError("state has %d fields, line table needs %d" % (nf, abi.nfield(table.nline)))
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


def invert_dev(table, ray, model, state_t, lam_t, obs_t, mask_t, nodes0_t, nodes_t, chi2_t, iters_t,
               wstokes=(1.0, 1.0, 1.0, 1.0), lm_start=1.0, max_iter=30, nslot=0, precision=abi.SNAPI_MIXED,
               jacobian=0, stream=None):
    """snapi_invert_dev on torch device tensors (state [npix][nfield][nd],
    obs [npix][nlambda][4], mask [npix][nlambda] or None, in the precision's
    type; lam float64; nodes0/nodes float64 [npix][nnode]; chi2 float64;
    iterations int32).  Returns the total number of pixel-iterations."""
    import torch
    npix, nf, nd = state_t.shape
    nl = lam_t.shape[0]
    nn = model.nnode
    dt = torch.float64 if precision == abi.SNAPI_FP64 else torch.float32
    want = [(state_t, dt, (npix, abi.nfield(table.nline), nd)), (lam_t, torch.float64, (nl,)),
            (obs_t, dt, (npix, nl, 4)), (nodes0_t, torch.float64, (npix, nn)), (nodes_t, torch.float64, (npix, nn)),
            (chi2_t, torch.float64, (npix,)), (iters_t, torch.int32, (npix,))]
    if mask_t is not None:
        want.append((mask_t, dt, (npix, nl)))
    for t, d, shp in want:
        if t.dtype != d or tuple(t.shape) != shp or not t.is_contiguous() or not t.is_cuda:
            raise ValueError("invert_dev: expected a contiguous device tensor %s of shape %s, got %s %s"
                             % (d, shp, t.dtype, tuple(t.shape)))
    ws = np.ascontiguousarray(wstokes, dtype=np.float64)
    total = C.c_longlong(0)
    L = lib()
    f = L.snapi_invert_dev
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_int, C.POINTER(abi.SnapiLineTable), C.POINTER(abi.SnapiRay),
                  C.POINTER(SnapiNodeModel), C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                  C.c_void_p, C.POINTER(C.c_double), C.c_double, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                  C.c_void_p, C.c_void_p, C.POINTER(C.c_longlong), C.c_void_p]
    sp = None if stream is None else C.c_void_p(stream.cuda_stream)
    rc = f(precision, jacobian, C.byref(table), C.byref(ray), C.byref(model), npix, nd, state_t.data_ptr(), nl,
           lam_t.data_ptr(), obs_t.data_ptr(), None if mask_t is None else mask_t.data_ptr(), abi.ptr(ws), lm_start,
           max_iter, nslot, nodes0_t.data_ptr(), nodes_t.data_ptr(), chi2_t.data_ptr(), iters_t.data_ptr(),
           C.byref(total), sp)
    check("snapi_invert_dev", rc)
    return total.value


def node_jacobian(table, ray, model, state, nodes, lam_vac, precision=abi.SNAPI_FP64, jacobian=2):
    """(stokes [npix][nlambda][4], J [npix][nlambda*4][nnode]) of snapi_node_jacobian."""
    dt = np.float64 if precision == abi.SNAPI_FP64 else np.float32
    st = np.ascontiguousarray(state, dtype=dt)
    if st.ndim == 2:
        st = st[None]
    npix, nf, nd = st.shape
    if nf != abi.nfield(table.nline):
        raise ValueError("state has %d fields, line table needs %d" % (nf, abi.nfield(table.nline)))
    nv = np.ascontiguousarray(nodes, dtype=np.float64).reshape(npix, model.nnode)
    lam = np.ascontiguousarray(lam_vac, dtype=np.float64)
    stokes = np.zeros((npix, len(lam), 4), dtype=dt)
    J = np.zeros((npix, len(lam) * 4, model.nnode), dtype=dt)
    f = lib().snapi_node_jacobian
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_int, C.POINTER(abi.SnapiLineTable), C.POINTER(abi.SnapiRay),
                  C.POINTER(SnapiNodeModel), C.c_int, C.c_int, C.c_void_p, C.POINTER(C.c_double), C.c_int,
                  C.POINTER(C.c_double), C.c_void_p, C.c_void_p]
    rc = f(precision, jacobian, C.byref(table), C.byref(ray), C.byref(model), npix, nd, abi.vptr(st), abi.ptr(nv),
           len(lam), abi.ptr(lam), abi.vptr(stokes), abi.vptr(J))
    check("snapi_node_jacobian", rc)
    return stokes, J


def node_jacobian_eos(table, ray, model, chem, chem_he, smap, state, nodes, lam_vac, precision=abi.SNAPI_FP64,
                      analytic_maps=False, jacobian=abi.SNAPI_JAC_ADJOINT):
    """(stokes, J) of snapi_node_jacobian_eos: temperature nodes allowed, the
    thermodynamic fields of state rebuilt from the nodes (hydrostatic
    equilibrium with chem_he, equation of state with chem and smap); adjoint
    or (jacobian = SNAPI_JAC_TANGENT) forward-mode Jacobian; the Stokes
    vector of the response path.  analytic_maps: node responses by exact
    derivatives and the hydrostatic tangent (SNAPI_JAC_ANALYTIC_MAPS) instead
    of the reference's differences."""
    dt = np.float64 if precision == abi.SNAPI_FP64 else np.float32
    st = np.ascontiguousarray(state, dtype=dt)
    if st.ndim == 2:
        st = st[None]
    npix, nf, nd = st.shape
    if nf != abi.nfield(table.nline):
        raise ValueError("state has %d fields, line table needs %d" % (nf, abi.nfield(table.nline)))
    nv = np.ascontiguousarray(nodes, dtype=np.float64).reshape(npix, model.nnode)
    lam = np.ascontiguousarray(lam_vac, dtype=np.float64)
    stokes = np.zeros((npix, len(lam), 4), dtype=dt)
    J = np.zeros((npix, len(lam) * 4, model.nnode), dtype=dt)
    f = lib().snapi_node_jacobian_eos
    f.restype = C.c_int
    TB = C.POINTER(abi.SnapiChemTable)
    f.argtypes = [C.c_int, C.c_int, C.POINTER(abi.SnapiLineTable), C.POINTER(abi.SnapiRay),
                  C.POINTER(SnapiNodeModel), TB, TB, C.POINTER(abi.SnapiStateMap), C.c_int, C.c_int, C.c_void_p,
                  C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double), C.c_void_p, C.c_void_p]
    jac = (abi.SNAPI_JAC_ANALYTIC_MAPS if analytic_maps else 0) | int(jacobian)
    rc = f(precision, jac, C.byref(table), C.byref(ray), C.byref(model), C.byref(chem), C.byref(chem_he),
           C.byref(smap), npix, nd, abi.vptr(st), abi.ptr(nv), len(lam), abi.ptr(lam), abi.vptr(stokes), abi.vptr(J))
    check("snapi_node_jacobian_eos", rc)
    return stokes, J


def invert_eos(table, ray, model, chem, chem_he, smap, state, lam_vac, obs, nodes0, mask=None,
               wstokes=(1.0, 1.0, 1.0, 1.0), lm_start=1.0, max_iter=30, nslot=0, precision=abi.SNAPI_FP64,
               broadening=0.0, analytic_maps=False, jacobian=abi.SNAPI_JAC_ADJOINT, psf=None):
    """snapi_invert_eos: the inversion with temperature nodes (adjoint
    Jacobian; analytic_maps as node_jacobian_eos); broadening [cm] and the
    sampled spectral psf as snapi_instrument. Returns (nodes, chi2,
    iterations, total pixel-iterations)."""
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
    inst = instrument(broadening, psf)
    f = lib().snapi_invert_eos
    f.restype = C.c_int
    TB = C.POINTER(abi.SnapiChemTable)
    f.argtypes = [C.c_int, C.c_int, C.POINTER(abi.SnapiLineTable), C.POINTER(abi.SnapiRay), C.POINTER(SnapiNodeModel),
                  TB, TB, C.POINTER(abi.SnapiStateMap), C.c_int, C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_double),
                  C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.POINTER(SnapiInstrument), C.c_double, C.c_int,
                  C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_longlong)]
    jac = (abi.SNAPI_JAC_ANALYTIC_MAPS if analytic_maps else 0) | int(jacobian)
    rc = f(precision, jac, C.byref(table), C.byref(ray), C.byref(model), C.byref(chem), C.byref(chem_he), C.byref(smap),
           npix, nd, abi.vptr(st), len(lam), abi.ptr(lam), abi.vptr(ob), None if mk is None else abi.vptr(mk),
           abi.ptr(ws), _inst_ptr(inst), lm_start, max_iter, nslot, abi.ptr(nodes), abi.ptr(chi2),
           iters.ctypes.data_as(C.POINTER(C.c_int)), C.byref(total))
    check("snapi_invert_eos", rc)
    return nodes, chi2, iters, total.value


def build_from_nodes(model, chem, chem_he, smap, state, nodes):
    """snapi_build_from_nodes: state [npix][nfield][nd] (double, GRID set)
    with the node fields, T, NT and the equation-of-state fields filled."""
    st = np.array(state, dtype=np.float64, order="C", copy=True)
    if st.ndim == 2:
        st = st[None]
    npix, nf, nd = st.shape
    nv = np.ascontiguousarray(nodes, dtype=np.float64).reshape(npix, model.nnode)
    f = lib().snapi_build_from_nodes
    f.restype = C.c_int
    TB = C.POINTER(abi.SnapiChemTable)
    f.argtypes = [C.POINTER(SnapiNodeModel), TB, TB, C.POINTER(abi.SnapiStateMap), C.c_int, C.c_int,
                  C.POINTER(C.c_double), C.POINTER(C.c_double)]
    check("snapi_build_from_nodes", f(C.byref(model), C.byref(chem), C.byref(chem_he), C.byref(smap), npix, nd,
                                      abi.ptr(st), abi.ptr(nv)))
    return st


def build_from_nodes_dev(model, chem, chem_he, smap, state_t, nodes_t, stream=None):
    """snapi_build_from_nodes_dev on float64 device tensors (state in place)."""
    npix, nf, nd = state_t.shape
    f = lib().snapi_build_from_nodes_dev
    f.restype = C.c_int
    TB = C.POINTER(abi.SnapiChemTable)
    f.argtypes = [C.POINTER(SnapiNodeModel), TB, TB, C.POINTER(abi.SnapiStateMap), C.c_int, C.c_int, C.c_void_p,
                  C.c_void_p, C.c_void_p]
    sp = None if stream is None else C.c_void_p(stream.cuda_stream)
    check("snapi_build_from_nodes_dev", f(C.byref(model), C.byref(chem), C.byref(chem_he), C.byref(smap), npix, nd,
                                          state_t.data_ptr(), nodes_t.data_ptr(), sp))


def invert_eos_dev(table, ray, model, chem, chem_he, smap, state_t, lam_t, obs_t, mask_t, nodes0_t, nodes_t, chi2_t,
                   iters_t, wstokes=(1.0, 1.0, 1.0, 1.0), lm_start=1.0, max_iter=30, nslot=0,
                   precision=abi.SNAPI_MIXED, stream=None, broadening=0.0, analytic_maps=False,
                   jacobian=abi.SNAPI_JAC_ADJOINT, psf=None):
    """snapi_invert_eos_dev on torch device tensors; returns the total
    pixel-iterations."""
    npix, nf, nd = state_t.shape
    nl = lam_t.shape[0]
    ws = np.ascontiguousarray(wstokes, dtype=np.float64)
    total = C.c_longlong(0)
    inst = instrument(broadening, psf)
    f = lib().snapi_invert_eos_dev
    f.restype = C.c_int
    TB = C.POINTER(abi.SnapiChemTable)
    f.argtypes = [C.c_int, C.c_int, C.POINTER(abi.SnapiLineTable), C.POINTER(abi.SnapiRay), C.POINTER(SnapiNodeModel),
                  TB, TB, C.POINTER(abi.SnapiStateMap), C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                  C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.POINTER(SnapiInstrument), C.c_double, C.c_int,
                  C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_longlong), C.c_void_p]
    sp = None if stream is None else C.c_void_p(stream.cuda_stream)
    jac = (abi.SNAPI_JAC_ANALYTIC_MAPS if analytic_maps else 0) | int(jacobian)
    rc = f(precision, jac, C.byref(table), C.byref(ray), C.byref(model), C.byref(chem), C.byref(chem_he), C.byref(smap),
           npix, nd, state_t.data_ptr(), nl, lam_t.data_ptr(), obs_t.data_ptr(),
           None if mask_t is None else mask_t.data_ptr(), abi.ptr(ws), _inst_ptr(inst), lm_start, max_iter, nslot,
           nodes0_t.data_ptr(), nodes_t.data_ptr(), chi2_t.data_ptr(), iters_t.data_ptr(), C.byref(total), sp)
    check("snapi_invert_eos_dev", rc)
    return total.value
