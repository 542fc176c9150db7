"""ctypes mirror of include/snapi_gpu.h (structs, enums and sizes).

The structs here are the plain-C boundary of the library; the field layout
must match the header exactly.
"""
import ctypes as C

import numpy as np

SNAPI_OK = 0
SNAPI_ERR_ARG = -1
SNAPI_ERR_DEVICE = -2
SNAPI_ERR_UNSUPPORTED = -3
SNAPI_ERR_NONFINITE = -4

SNAPI_FP64 = 0
SNAPI_MIXED = 1

F_T, F_VT, F_VLOS, F_BMAG, F_BINC, F_BAZI, F_NE, F_NHI, F_OPREF, F_GRID, F_BGOP, F_BGEM, F_NT = range(13)
NF_BASE = 13

MAX_LINES = 16
MAX_COMPONENTS = 96
MAX_ND = 256
MAX_NODES = 16

STATUS_TEXT = {
    SNAPI_OK: "ok",
    SNAPI_ERR_ARG: "bad argument",
    SNAPI_ERR_DEVICE: "HIP runtime error",
    SNAPI_ERR_UNSUPPORTED: "input outside what the kernels implement",
    SNAPI_ERR_NONFINITE: "non-finite result",
}


class SnapiLine(C.Structure):
    _fields_ = [("lam0", C.c_double), ("aul", C.c_double), ("blu", C.c_double), ("bul", C.c_double),
                ("grad", C.c_double), ("sigma", C.c_double), ("alpha", C.c_double), ("mass", C.c_double),
                ("split", C.c_int), ("ncomp", C.c_int * 3), ("comp0", C.c_int)]


class SnapiComponent(C.Structure):
    _fields_ = [("shift", C.c_double), ("strength", C.c_double)]


class SnapiLineTable(C.Structure):
    _fields_ = [("nline", C.c_int), ("ncomp_total", C.c_int), ("line", SnapiLine * MAX_LINES),
                ("comp", SnapiComponent * MAX_COMPONENTS)]


class SnapiRay(C.Structure):
    _fields_ = [("theta", C.c_double), ("lam_mid", C.c_double), ("tau_grid", C.c_int), ("hminus", C.c_int),
                ("hminus_mid", C.c_int)]


# ALI (snapi_ali_*): per-depth fields, transitions, atom, grid
A_T, A_VT, A_NE, A_OPREF, A_GRID, A_NA, A_NF = range(7)
ALI_MAX_LEVELS = 16
ALI_MAX_TRANS = 64
TR_LINE, TR_BF = 0, 1


class SnapiAliTransition(C.Structure):
    _fields_ = [("kind", C.c_int), ("lower", C.c_int), ("upper", C.c_int), ("lam0", C.c_double),
                ("aul", C.c_double), ("blu", C.c_double), ("bul", C.c_double), ("energy", C.c_double),
                ("gratio", C.c_double), ("de", C.c_double)]


class SnapiAliAtom(C.Structure):
    _fields_ = [("nlevel", C.c_int), ("ntrans", C.c_int), ("mass", C.c_double),
                ("tr", SnapiAliTransition * ALI_MAX_TRANS)]


class SnapiAliGrid(C.Structure):
    _fields_ = [("nlam", C.c_int), ("lam", C.POINTER(C.c_double)), ("lamw", C.POINTER(C.c_double)),
                ("fudge", C.POINTER(C.c_double)), ("sigma", C.POINTER(C.c_double)), ("nray", C.c_int),
                ("theta", C.POINTER(C.c_double)), ("wray", C.POINTER(C.c_double)), ("bg_per_ray", C.c_int),
                ("boundary", C.c_double), ("tau_grid", C.c_int)]


def make_ali_atom(nlevel, mass, trans):
    """trans: list of dicts (kind, lower, upper and, for lines, lam0, aul, blu,
    bul, energy; for bound-free, gratio, de) in the atom's transition order."""
    a = SnapiAliAtom()
    a.nlevel = nlevel
    a.ntrans = len(trans)
    a.mass = mass
    for i, t in enumerate(trans):
        for k, v in t.items():
            setattr(a.tr[i], k, v)
    return a


def make_ali_grid(lam, lamw, fudge, sigma, theta, wray, bg_per_ray, boundary=-1.0, tau_grid=1):
    """The arrays are kept alive on the returned structure (attribute _keep)."""
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (lam, lamw, fudge, sigma, theta, wray)]
    g = SnapiAliGrid()
    g.nlam = len(arrs[0])
    g.lam, g.lamw, g.fudge, g.sigma = (ptr(x) for x in arrs[:4])
    g.nray = len(arrs[4])
    g.theta, g.wray = ptr(arrs[4]), ptr(arrs[5])
    g.bg_per_ray = int(bg_per_ray)
    g.boundary = boundary
    g.tau_grid = int(tau_grid)
    g._keep = arrs
    return g


def nfield(nline):
    return NF_BASE + 2 * nline


def make_line_table(lines):
    """lines: list of dicts with keys lam0, aul, blu, bul, grad, sigma, alpha,
    mass, split, and (for split lines) comps = [pi, sigma_b, sigma_r] lists of
    (shift, strength)."""
    t = SnapiLineTable()
    if len(lines) > MAX_LINES:
        raise ValueError("at most %d lines" % MAX_LINES)
    t.nline = len(lines)
    c = 0
    for i, d in enumerate(lines):
        L = t.line[i]
        for k in ("lam0", "aul", "blu", "bul", "grad", "sigma", "alpha", "mass"):
            setattr(L, k, float(d[k]))
        L.split = int(d["split"])
        L.comp0 = c
        groups = d.get("comps", [[], [], []]) if L.split else [[], [], []]
        for g in range(3):
            L.ncomp[g] = len(groups[g])
            for sh, st in groups[g]:
                if c >= MAX_COMPONENTS:
                    raise ValueError("at most %d Zeeman components" % MAX_COMPONENTS)
                t.comp[c].shift = float(sh)
                t.comp[c].strength = float(st)
                c += 1
    t.ncomp_total = c
    return t


def ptr(a, ctype=C.c_double):
    """Pointer to a C-contiguous numpy array (None for None)."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(C.POINTER(ctype))


def vptr(a):
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return C.c_void_p(a.ctypes.data)
