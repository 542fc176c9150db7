"""Line tables from level data (snapi_line_add, snapi_zeeman_pattern,
snapi_vdw_damping, snapi_radiative_damping): the atom class's line setup
(atmos/atmol/atom/atom.cc:233-253, 926-935, 987-1308, 1620-1773) on the host,
through the C ABI.  Energies in erg, wavelengths in cm."""
import ctypes as C

import numpy as np

from . import abi
from ._native import lib, check


class SnapiLevel(C.Structure):
    _fields_ = [("energy", C.c_double), ("g", C.c_double), ("j2", C.c_int), ("l", C.c_int),
                ("g_lande", C.c_double)]


def _declare(L):
    if getattr(L, "_lines_declared", False):
        return
    dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
    LV = C.POINTER(SnapiLevel)
    L.snapi_zeeman_pattern.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, ip, ip,
                                       C.POINTER(abi.SnapiComponent), C.c_int]
    L.snapi_vdw_damping.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, dp, dp]
    L.snapi_radiative_damping.argtypes = [C.c_int, dp, C.c_int, C.c_int, dp]
    L.snapi_line_add.argtypes = [C.POINTER(abi.SnapiLineTable), C.c_double, C.c_int, C.c_double, LV, LV, C.c_double,
                                 C.c_double]
    for f in (L.snapi_zeeman_pattern, L.snapi_vdw_damping, L.snapi_radiative_damping, L.snapi_line_add):
        f.restype = C.c_int
    L._lines_declared = True


def level(energy, g, j2, l, g_lande):
    return SnapiLevel(energy=energy, g=g, j2=int(j2), l=int(l), g_lande=g_lande)


def zeeman_pattern(j2_lower, j2_upper, g_lower, g_upper):
    """(split, [(shift, strength) per pi / sigma_b / sigma_r])."""
    L = lib()
    _declare(L)
    comp = (abi.SnapiComponent * abi.MAX_COMPONENTS)()
    split = C.c_int(0)
    nc = (C.c_int * 3)()
    check("snapi_zeeman_pattern", L.snapi_zeeman_pattern(j2_lower, j2_upper, g_lower, g_upper, C.byref(split), nc,
                                                         comp, abi.MAX_COMPONENTS))
    out, k = [], 0
    for g in range(3):
        out.append([(comp[k + i].shift, comp[k + i].strength) for i in range(nc[g])])
        k += nc[g]
    return split.value, out


def vdw_damping(stage, e_ion, e_lower, e_upper, l_lower, l_upper):
    """(sigma, alpha) of the van der Waals broadening."""
    L = lib()
    _declare(L)
    s, a = C.c_double(0), C.c_double(0)
    check("snapi_vdw_damping", L.snapi_vdw_damping(stage, e_ion, e_lower, e_upper, l_lower, l_upper, C.byref(s),
                                                   C.byref(a)))
    return s.value, a.value


def radiative_damping(A, lower, upper):
    L = lib()
    _declare(L)
    A = np.ascontiguousarray(A, dtype=np.float64)
    g = C.c_double(0)
    check("snapi_radiative_damping", L.snapi_radiative_damping(A.shape[0], abi.ptr(A), lower, upper, C.byref(g)))
    return g.value


def empty_table():
    t = abi.SnapiLineTable()
    t.nline = 0
    t.ncomp_total = 0
    return t


def add_line(table, mass, stage, e_ion, lower, upper, aul, grad):
    """Appends one line (levels as SnapiLevel) to table in place."""
    L = lib()
    _declare(L)
    check("snapi_line_add", L.snapi_line_add(C.byref(table), mass, stage, e_ion, C.byref(lower), C.byref(upper), aul,
                                             grad))
    return table
