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
SNAPI_ERR_CAPACITY = -5

SNAPI_FP64 = 0
SNAPI_MIXED = 1

# jacobian argument (include/snapi_gpu.h enum snapi_jacobian)
SNAPI_JAC_ADJOINT, SNAPI_JAC_DIFFERENCES, SNAPI_JAC_TANGENT = 0, 1, 2
SNAPI_JAC_ANALYTIC_MAPS = 16

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
    SNAPI_ERR_CAPACITY: "problem larger than the fixed work area",
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


class SnapiNlteAtom(C.Structure):
    _fields_ = [("nmap", C.c_int), ("ntr", C.c_int), ("nstage", C.c_int), ("damping", C.c_int),
                ("mass", C.c_double), ("level_stage", C.POINTER(C.c_int)), ("level_index", C.POINTER(C.c_int)),
                ("level_energy", C.POINTER(C.c_double)), ("level_g", C.POINTER(C.c_double)),
                ("stage_ip", C.POINTER(C.c_double)), ("tr_type", C.POINTER(C.c_int)),
                ("tr_lower", C.POINTER(C.c_int)), ("tr_upper", C.POINTER(C.c_int)), ("tr_A", C.POINTER(C.c_double)),
                ("tr_Blu", C.POINTER(C.c_double)), ("tr_Bul", C.POINTER(C.c_double)),
                ("tr_damp_rad", C.POINTER(C.c_double)), ("tr_lambda_crit", C.POINTER(C.c_double))]


SNAPI_NLTE_NFIELD = 13
NLTE_FIELDS = ("T", "vt", "ne", "Nt", "opref", "grid", "dne_dT", "dne_dNt", "dopref_dT", "dopref_dNt", "dopref_dvt",
               "dNact_dT", "dNact_dNt")


def make_nlte_atom(level_stage, level_energy, level_g, stage_ip, tr_type, tr_lower, tr_upper, tr_A, tr_Blu, tr_Bul,
                   tr_damp_rad, tr_lambda_crit, mass, damping=1):
    """snapi_nlte_atom from per-level and per-transition arrays (map order;
    level_index is derived from level_stage).  The arrays are kept alive on
    the returned structure (attribute _keep)."""
    st = np.ascontiguousarray(level_stage, dtype=np.int32)
    idx = np.zeros_like(st)
    for i in range(1, st.size):
        idx[i] = idx[i - 1] + 1 if st[i] == st[i - 1] else 0
    ints = [st, idx] + [np.ascontiguousarray(x, dtype=np.int32) for x in (tr_type, tr_lower, tr_upper)]
    dbls = [np.ascontiguousarray(x, dtype=np.float64) for x in (level_energy, level_g, stage_ip, tr_A, tr_Blu, tr_Bul,
                                                                 tr_damp_rad, tr_lambda_crit)]
    a = SnapiNlteAtom()
    a.nmap, a.ntr, a.nstage, a.damping, a.mass = st.size, ints[2].size, dbls[2].size, int(damping), float(mass)
    ip = lambda x: x.ctypes.data_as(C.POINTER(C.c_int))  # noqa: E731
    a.level_stage, a.level_index, a.tr_type, a.tr_lower, a.tr_upper = (ip(x) for x in ints)
    (a.level_energy, a.level_g, a.stage_ip, a.tr_A, a.tr_Blu, a.tr_Bul, a.tr_damp_rad,
     a.tr_lambda_crit) = (ptr(x) for x in dbls)
    a._keep = ints + dbls
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


# Chemical equilibrium (snapi_chemeq*): species table
CHEM_MAX_ELEMENTS = 14
CHEM_MAX_STAGES = 10
CHEM_MAX_TERMS = 6
CHEM_MAX_COEFS = 8
CHEM_MAX_MOLECULES = 32
CHEM_MAX_COMPONENTS = 4
CHEM_MAX_LEVELS = 256
CHEM_MAX_BF = 64
PF_NONE, PF_CONST, PF_TRAV, PF_IRWIN = range(4)
EQC_NONE, EQC_CONST, EQC_POLY, EQC_HMIN = range(4)


class SnapiPartition(C.Structure):
    _fields_ = [("type", C.c_int), ("nterm", C.c_int), ("value", C.c_double), ("charge", C.c_double),
                ("g0", C.c_double), ("g2", C.c_double * CHEM_MAX_TERMS), ("xion", C.c_double * CHEM_MAX_TERMS),
                ("xl", C.c_double * CHEM_MAX_TERMS), ("ncoef", C.c_int * CHEM_MAX_TERMS),
                ("gamma", (C.c_double * CHEM_MAX_COEFS) * CHEM_MAX_TERMS),
                ("alpha", (C.c_double * CHEM_MAX_COEFS) * CHEM_MAX_TERMS),
                ("irwin", C.c_double * CHEM_MAX_COEFS)]


class SnapiChemElement(C.Structure):
    _fields_ = [("nstage", C.c_int), ("abundance", C.c_double), ("eion", C.c_double * CHEM_MAX_STAGES),
                ("pf", SnapiPartition * CHEM_MAX_STAGES)]


class SnapiChemMolecule(C.Structure):
    _fields_ = [("ncomp", C.c_int), ("element", C.c_int * CHEM_MAX_COMPONENTS),
                ("stage", C.c_int * CHEM_MAX_COMPONENTS), ("eqc", C.c_int), ("ncoef", C.c_int),
                ("coef", C.c_double * CHEM_MAX_COEFS)]


class SnapiChemLevel(C.Structure):
    _fields_ = [("element", C.c_int), ("stage", C.c_int), ("g", C.c_double), ("energy", C.c_double)]


class SnapiChemBf(C.Structure):
    _fields_ = [("element", C.c_int), ("stage", C.c_int), ("g", C.c_double), ("energy", C.c_double),
                ("g_up", C.c_double), ("energy_up", C.c_double), ("sigma", C.c_double)]


class SnapiChemTable(C.Structure):
    _fields_ = [("nelement", C.c_int), ("nmolecule", C.c_int), ("nlevel", C.c_int), ("nbf", C.c_int),
                ("hminus", C.c_int), ("lambda_ref", C.c_double),
                ("element", SnapiChemElement * CHEM_MAX_ELEMENTS),
                ("molecule", SnapiChemMolecule * CHEM_MAX_MOLECULES),
                ("level", SnapiChemLevel * CHEM_MAX_LEVELS),
                ("bf", SnapiChemBf * CHEM_MAX_BF)]


def make_chem_table(d, levels=True, referent=0):
    """Species table from arrays shaped as tests/golden/chemeq.npz holds them:
    per element Z, abund, eion, pf_* and term_*/irwin parameters; per molecule
    mol_eqc/mol_ncoef/mol_coef/mol_ncomp/mol_elem/mol_stage; levels
    (element 0-based, stage, level, g) with level_ee [erg]; the bound-free
    transitions bf_int/bf_real with their cross-sections at the referent
    wavelength lref[referent] (0: airtovac(5000 A), as compute_op_referent
    uses it; 1: 5000 A, as the hydrostatic integration does); H- included when
    a molecule has the HMIN form."""
    t = SnapiChemTable()
    nel = len(d["Z"])
    t.nelement = nel
    for e in range(nel):
        el = t.element[e]
        el.nstage = int(d["Z"][e]) + 1
        el.abundance = float(d["abund"][e])
        for z in range(el.nstage - 1):
            el.eion[z] = float(d["eion"][e, z])
        for z in range(el.nstage):
            pf = el.pf[z]
            pf.type = int(d["pf_type"][e, z])
            pf.nterm = int(d["pf_nterm"][e, z])
            pf.value = float(d["pf_value"][e, z])
            pf.charge = float(d["pf_charge"][e, z])
            pf.g0 = float(d["pf_g0"][e, z])
            for i in range(CHEM_MAX_TERMS if pf.type == PF_TRAV else 0):
                pf.g2[i] = float(d["term_g2"][e, z, i])
                pf.xion[i] = float(d["term_xion"][e, z, i])
                pf.xl[i] = float(d["term_xl"][e, z, i])
                pf.ncoef[i] = int(d["term_ncoef"][e, z, i])
                for c in range(CHEM_MAX_COEFS):
                    pf.gamma[i][c] = float(d["term_gamma"][e, z, i, c])
                    pf.alpha[i][c] = float(d["term_alpha"][e, z, i, c])
            if pf.type == PF_IRWIN:
                for c in range(pf.nterm):
                    pf.irwin[c] = float(d["irwin"][e, z, c])
    nmol = len(d["mol_eqc"])
    t.nmolecule = nmol
    for m in range(nmol):
        mo = t.molecule[m]
        mo.ncomp = int(d["mol_ncomp"][m])
        mo.eqc = int(d["mol_eqc"][m])
        mo.ncoef = int(d["mol_ncoef"][m])
        for c in range(mo.ncoef):
            mo.coef[c] = float(d["mol_coef"][m, c])
        for i in range(mo.ncomp):
            mo.element[i] = int(d["mol_elem"][m, i])
            mo.stage[i] = int(d["mol_stage"][m, i])
    t.lambda_ref = float(d["lref"][referent]) if "lref" in d else 5000e-8
    t.hminus = int(any(int(e) == EQC_HMIN for e in d["mol_eqc"]))
    if "bf_int" in d:
        bi, br = d["bf_int"], d["bf_real"]
        t.nbf = len(bi)
        for j in range(len(bi)):
            b = t.bf[j]
            b.element, b.stage = int(bi[j, 0]) + 1, int(bi[j, 1])
            b.g, b.energy, b.g_up = float(bi[j, 3]), float(br[j, 0]), float(bi[j, 4])
            b.energy_up, b.sigma = float(br[j, 1]), float(br[j, 2 + referent])
    if levels:
        lv = d["levels"]
        t.nlevel = len(lv)
        for j in range(len(lv)):
            t.level[j].element = int(lv[j, 0]) + 1
            t.level[j].stage = int(lv[j, 1])
            t.level[j].g = float(lv[j, 3])
            t.level[j].energy = float(d["level_ee"][j])
    return t


class SnapiStateMap(C.Structure):
    _fields_ = [("nline", C.c_int), ("line_level", (C.c_int * 2) * MAX_LINES), ("nhi_element", C.c_int),
                ("lambda_mid", C.c_double), ("lambda_bg", C.c_double), ("sigma_bg", C.c_double * CHEM_MAX_BF)]


def make_state_map(table, d, lines, bfsig, lam_mid, nhi_element=1):
    """Where chemical equilibrium fills a column state: each line's lower and
    upper level as the table's level index, found from (species, stage,
    level) of lines [nline][>=4] in the table arrays d (levels: species
    0-based, stage, level, g); the bound-free background from bfsig rows
    (species, stage, level, edge, wavelength, cross-section) as the reference
    records them at the window's last wavelength, matched to the table's bf
    transitions (those with no table entry must have zero cross-section)."""
    m = SnapiStateMap()
    lv = np.asarray(d["levels"])

    def level_index(s, z, l):
        hit = np.nonzero((lv[:, 0] == s) & (lv[:, 1] == z) & (lv[:, 2] == l))[0]
        if len(hit) != 1:
            raise ValueError("level (%d, %d, %d) not in the table" % (s, z, l))
        return int(hit[0])

    m.nline = len(lines)
    for i, row in enumerate(lines):
        s, z, lo, up = (int(x) for x in row[:4])
        m.line_level[i][0] = level_index(s, z, lo)
        m.line_level[i][1] = level_index(s, z, up)
    m.nhi_element = nhi_element
    m.lambda_mid = lam_mid
    lam_bg = [r[4] for r in bfsig]
    m.lambda_bg = float(lam_bg[0]) if len(lam_bg) else 0.0
    if len(lam_bg) and not np.all(np.asarray(lam_bg) == lam_bg[0]):
        raise ValueError("one background wavelength expected")
    bi = np.asarray(d.get("bf_int", np.zeros((0, 5))))
    for r in bfsig:
        s, z, l, sig = int(r[0]), int(r[1]), int(r[2]), float(r[5])
        hit = [j for j in range(len(bi)) if bi[j, 0] == s and bi[j, 1] == z and bi[j, 2] == l]
        if not hit:
            if sig != 0.0:
                raise ValueError("bound-free transition (%d, %d, %d) is missing from the table" % (s, z, l))
            continue
        m.sigma_bg[hit[0]] = sig
    return m


def chem_stage_count(table):
    return sum(table.element[e].nstage for e in range(table.nelement))


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
