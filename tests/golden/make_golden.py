"""Writes the golden vectors under tests/golden/ by running the reference SNAPI
code (built from /root/reference by oracle/ref.mk into oracle/_ref/).

Run here only (the reference is not present where the GPU tests run):

    make -f oracle/ref.mk -j8 && python tests/golden/make_golden.py [fixture ...]

Inputs are written into a scratch directory in the reference's own config
formats (atmos{} block, atom{} files, SPINOR 1-D atmosphere): the atom files
for the background species are read from the reference's cfg/ directory at
generation time, the two-line Fe I atom is generated here from the line
parameters of the 6301.5/6302.5 pair (see `fe_atom_cfg`).  Outputs are small
.npz files of inputs and expected outputs; no reference source is stored.
"""
import math
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("SNAPI_REFERENCE", "/root/reference")
DRIVER = os.path.join(REPO, "oracle", "_ref", "ref_driver")

# CGS constants as the reference's cmlib/const.h states them.
C_LIGHT = 2.99792458e10
H_PLANCK = 6.6260687652e-27
EV = 1.602189e-12
M_E = 9.1093818872e-28
E_CHARGE = 4.8032042510e-10


def airtovac(lam_cm):
    """Air -> vacuum wavelength, the refractive-index formula of
    cmlib/mathtools.cc:1325-1331 (lambda in cm)."""
    s = 1e4 / (lam_cm * 1e8)
    n = 1.0 + 0.00008336624212083 + 0.02408926869968 / (130.1065924522 - s * s) \
        + 0.0001599740894897 / (38.92568793293 - s * s)
    return lam_cm * n


# The two Fe I-like lines: air wavelength [A], lower-level energy [eV],
# log gf, (2J, g, Lande g) of lower and upper level.
LINES = [
    dict(lam_air=6301.5, ee_low=3.65, loggf=-0.72, low=(4, 5, 1.83), up=(4, 5, 1.5)),
    dict(lam_air=6302.5, ee_low=3.650001, loggf=-1.24, low=(2, 3, 2.5), up=(0, 1, 0.0)),
]


def fe_atom_cfg():
    """A four-level Fe I atom in the reference's atom{} format (read by
    atmos/atmol/atom/atomcfg.cc): two lower levels near 3.65 eV and two upper
    levels placed so the line centres fall at the given air wavelengths.
    A_ul follows from log gf = log(g_l f_lu) with
    f_lu = (g_u/g_l) m_e c lambda^2 A_ul / (8 pi^2 e^2) (atom.cc:245-247).
    The second lower level sits 1e-6 eV above the first so that no pair of
    levels is degenerate.  Pairs without a line carry A = 1, the placeholder
    the reference's own atom files use: atom::rtsetup (atom.cc:1516-1541)
    needs a nonzero A between consecutive levels, and the bound-bound opacity
    keeps only A > 1 (atom_rad_synthesis.cc:403).  Partition functions and bound-free data are taken
    from the reference's cfg/iron_6300.cfg ion blocks."""
    src = open(os.path.join(REF, "cfg", "iron_6300.cfg")).read()
    # keep the ion-0 partition block and the ion 1/2 blocks of the template
    i0 = src.index("partition{")
    i1 = src.index("level{")
    part0 = src[i0:i1]
    k1 = src.index("CHARGE=1")
    ion1 = src[src.rindex("ion{", 0, k1) :].rstrip()
    assert ion1.endswith("}")
    ion1 = ion1[:-1]  # drop the closing brace of atom{}
    levels = []
    ups = []
    for ln in LINES:
        lam_vac = airtovac(ln["lam_air"] * 1e-8)
        de = H_PLANCK * C_LIGHT / lam_vac / EV
        ups.append(ln["ee_low"] + de)
    lo = [LINES[0], LINES[1]]
    A = []
    for ln, eu in zip(LINES, ups):
        lam = H_PLANCK * C_LIGHT / ((eu - ln["ee_low"]) * EV)
        gu = ln["up"][1]
        gf = 10.0 ** ln["loggf"]
        A.append(gf * 8.0 * math.pi ** 2 * E_CHARGE ** 2 / (M_E * C_LIGHT * lam * lam * gu))
    bf = "      bf{\n        TYPE=PHY\n        LAMBDA= 100\n        VALUE=%s\n      }\n"

    def level(n, ee, j2, g, gl, lqn, a_list, bfv):
        s = "    level{\n      N=%d\n      EE=%.9f\n      G=%d\n      J=%d\n      L=%d\n      GL=%.4f\n" % (
            n, ee, g, j2, lqn, gl)
        if a_list is not None:
            s += "      A=" + ", ".join("%.6e" % a for a in a_list) + "\n"
        return s + bf % bfv + "    }\n"

    levels.append(level(1, lo[0]["ee_low"], lo[0]["low"][0], lo[0]["low"][1], lo[0]["low"][2], 1, None, "0.092E-18"))
    levels.append(level(2, lo[1]["ee_low"], lo[1]["low"][0], lo[1]["low"][1], lo[1]["low"][2], 1, [1.0], "0.092E-18"))
    # upper levels in ascending energy: the 6302.5 upper level lies lower
    levels.append(level(3, ups[1], LINES[1]["up"][0], LINES[1]["up"][1], LINES[1]["up"][2], 0, [1.0, A[1]],
                        "0.092E-20"))
    levels.append(level(4, ups[0], LINES[0]["up"][0], LINES[0]["up"][1], LINES[0]["up"][2], 0, [A[0], 1.0, 1.0],
                        "0.092E-20"))
    assert ups[1] < ups[0]
    txt = ("atom{\n  NAME=Iron\n  ID=Fe\n  Z=2\n  MASS=55.854\n  ABUND=3.63e-05\n  NLTE=0\n"
           "  ion{\n    CHARGE=0\n    EION=7.896\n    " + part0 + "".join(levels) + "  }\n  " + ion1.strip() + "\n}\n")
    return txt


BACKGROUND = ["hydrogen_3lvl.cfg", "helium.cfg", "carbon.cfg", "nitrogen.cfg", "oxygen.cfg", "calcium.cfg",
              "silicon.cfg", "aluminium.cfg"]


def cfg0_atmosphere(nd=61):
    """The 61-point column: log tau500 uniform on [-6, 1], T(tau) =
    4400 + 2000 max(0, -log tau - 3.5) + 3000 tau^0.5, B = 1000 G at 45/30 deg,
    vmicro 1 km/s, vLOS 0.  Pg comes from the reference's hydrostatic
    integration (SNAPI_REF_HE=1); the Pg/Pe columns here are starting values."""
    lt = np.linspace(-6.0, 1.0, nd)
    tau = 10.0 ** lt
    T = 4400.0 + 2000.0 * np.maximum(0.0, -lt - 3.5) + 3000.0 * np.sqrt(tau)
    rows = []
    for i in range(nd):
        pg = 1e4 * (tau[i] + 1e-3) ** 0.7
        rows.append((lt[i], 1e7 * (1.0 - lt[i]), T[i], pg, pg * 1e-4, 0.0, 0.0, 1000.0, 1e5, 0.0,
                     math.radians(45.0), math.radians(30.0)))
    return rows


def write_spinor(path, rows):
    with open(path, "w") as f:
        f.write("%d cfg0\n" % len(rows))
        for r in rows:
            f.write(" ".join("%.9e" % v for v in r) + "\n")


def cfg0_lambda():
    lam = []
    for c in (6301.5, 6302.5):
        lam.extend(np.round(c + np.linspace(-0.6, 0.6, 201), 6))
    return np.array(lam)


def setup_dir(d, rows, lam, he=True):
    for a in BACKGROUND:
        shutil.copy(os.path.join(REF, "cfg", a), os.path.join(d, a))
    shutil.copy(os.path.join(REF, "cfg", "mol.cfg"), os.path.join(d, "mol.cfg"))
    open(os.path.join(d, "iron_2line.cfg"), "w").write(fe_atom_cfg())
    with open(os.path.join(d, "atoms.cfg"), "w") as f:
        for a in BACKGROUND + ["iron_2line.cfg"]:
            f.write("ATOM=%s\n" % a)
    with open(os.path.join(d, "atmos.cfg"), "w") as f:
        f.write("ID=cfg0\nTYPE=SPINOR\nFILE=atmos.dat\nGEOMETRY=PLANEPARALLEL\nRTSOLVER=Bezier\n"
                "ATOM=atoms.cfg\nMOL=mol.cfg\nELECTRONS=0\nTAU_GRID=1\nLEVEL_POPS=0\n"
                # dimensions given so the constructor allocates the arrays that
                # read_spinor's resize() then frees (atmos/atmos.cc:61-73)
                "NX=1\nNY=1\nNZ=%d\n" % len(rows))
    write_spinor(os.path.join(d, "atmos.dat"), rows)
    np.savetxt(os.path.join(d, "lambda.txt"), lam, fmt="%.6f")


def run(d, args, he=True):
    env = dict(os.environ)
    env["SNAPI_REF_HE"] = "1" if he else "0"
    out = subprocess.run([DRIVER] + args, cwd=d, env=env, capture_output=True, text=True, check=True)
    return out.stdout.splitlines()


def parse_synth(lines):
    rec = {"depth": [], "lambda": [], "stokes": [], "stokes_steps": [], "opem": [], "spop": [], "line": [],
           "zc": [], "lpop": [], "level": [], "species": [], "sdepth": [], "bfop": [], "bfsig": []}
    for ln in lines:
        t = ln.split()
        if not t:
            continue
        k = t[0]
        if k in rec:
            rec[k].append(t[1:])
    return rec


def build_cfg0(d):
    rows = cfg0_atmosphere()
    lam = cfg0_lambda()
    setup_dir(d, rows, lam)
    rec = parse_synth(run(d, ["synth", "atmos.cfg", "lambda.txt", repr(math.pi), "0.0", "20"]))
    save_synth(rec, rows, "cfg0_synth.npz")
    print("cfg0_synth: nd=%d nlambda=%d" % (len(rows), len(lam)))
    return rows


def save_synth(rec, rows, name, extra=None):
    dep = np.array([[float(x) for x in r[1:]] for r in rec["depth"]])
    # columns: T Ne Nt Vt Vr |B| incl_deg azi_deg op_ref tau_ref rt_grid N_HI height
    lamv = np.array([[float(r[1]), float(r[2])] for r in rec["lambda"]])
    stokes = np.array([[float(x) for x in r[1:]] for r in rec["stokes"]])
    steps = np.array([[float(x) for x in r[1:]] for r in rec["stokes_steps"]])
    opem = np.array([[float(x) for x in r] for r in rec["opem"]])
    spop = np.array([[float(x) for x in r] for r in rec["spop"]])
    sdepth = np.array([[float(x) for x in r] for r in rec["sdepth"]])
    bfop = np.array([[float(x) for x in r] for r in rec["bfop"]])
    # species, stage, level, edge wavelength, last (vacuum) wavelength, cross-section there
    bfsig = np.array([[float(x) for x in r] for r in rec["bfsig"]]).reshape(-1, 6)
    lines = []
    for r in rec["line"]:
        s, z, lo, up = (int(x) for x in r[:4])
        kv = dict(zip(r[4::2], r[5::2]))
        lines.append([s, z, lo, up, int(kv["tr"]), float(kv["lam"]), float(kv["A"]), float(kv["Blu"]),
                      float(kv["Bul"]), float(kv["grad"]), float(kv["sigma"]), float(kv["alpha"]),
                      int(kv["split"])])
    zc = [[int(r[0]), int(r[1]), "pbr".index(r[2]), float(r[3]), float(r[4])] for r in rec["zc"]]
    lpop = np.array([[float(x) for x in r] for r in rec["lpop"]])
    species = [r for r in rec["species"]]
    masses = {int(r[0]): float(r[r.index("mass") + 1]) for r in species if r[1] == "atom"}
    levels = [[int(r[0]), int(r[1]), int(r[2]), float(r[3]), int(r[5]), int(r[7]), int(r[9]), float(r[11])]
              for r in rec["level"]]
    np.savez_compressed(
        os.path.join(HERE, name),
        depth=dep, lam=lamv, stokes=stokes, stokes_steps=steps, opem=opem, spop=spop, sdepth=sdepth, bfop=bfop,
        bfsig=bfsig,
        lines=np.array([l[:4] + [l[4]] + [0] * 0 for l in lines], dtype=np.int64),
        line_data=np.array([l[5:12] for l in lines]), line_split=np.array([l[12] for l in lines]),
        zc=np.array(zc, dtype=np.float64), lpop=lpop,
        masses=np.array([[k, v] for k, v in sorted(masses.items())]),
        levels=np.array(levels, dtype=np.float64),
        atmos_rows=np.array(rows, dtype=np.float32), **(extra or {}))


def cfg1_lambda():
    """301 points, 8 mA apart, covering both lines: 6300.8 .. 6303.2 A."""
    return np.round(6300.8 + 0.008 * np.arange(301), 6)


def build_cfg1_template(d):
    """The 71-depth column the cube workload is drawn around (same profile as
    cfg0 on 71 points), its state and Stokes on the 301-point grid."""
    rows = cfg0_atmosphere(71)
    lam = cfg1_lambda()
    setup_dir(d, rows, lam)
    rec = parse_synth(run(d, ["synth", "atmos.cfg", "lambda.txt", repr(math.pi), "0.0", "1000"]))
    save_synth(rec, rows, "cfg1_template.npz")
    print("cfg1_template: nd=%d nlambda=%d" % (len(rows), len(lam)))


def build_resp(d):
    """Per-depth responses of the 61-point column on 31 wavelengths around
    6301.5 A: the state, thermodynamic derivatives and bound-free background
    as obs_stokes_responses sees them, and its Stokes and 7 x nd responses."""
    rows = cfg0_atmosphere()
    lam = np.round(6301.5 + np.linspace(-0.3, 0.3, 31), 6)
    setup_dir(d, rows, lam)
    out = run(d, ["resp", "atmos.cfg", "lambda.txt", repr(math.pi), "0.0"])
    rec = parse_synth(out)
    thermo, dpop, rs = [], [], []
    for ln in out:
        t = ln.split()
        if t and t[0] == "thermo":
            thermo.append([float(x) for x in t[1:]])
        elif t and t[0] == "dpop":
            dpop.append([float(x) for x in t[1:]])
        elif t and t[0] == "resp":
            rs.append([float(x) for x in t[1:]])
    rs = np.array(rs)
    nd = len(rows)
    nl = len(lam)
    resp = np.zeros((7, nd, nl, 4))
    for r in rs:
        resp[int(r[0]) - 1, int(r[1]) - 1, int(r[2]) - 1, :] = r[3:]
    save_synth(rec, rows, "cfg0_resp.npz", extra=dict(thermo=np.array(thermo), dpop=np.array(dpop), resp=resp))
    print("cfg0_resp: nlambda=%d" % nl)


def build_resp_incl(d):
    """Per-depth responses on an inclined ray (el = 2.4 rad, az = 0.6 rad) of
    the 61-point column with a field and velocity varying with depth (|B|
    300-1500 G, inclination 20-150 deg, azimuth -60..80 deg, vz 1 km/s):
    the reference's obs_stokes_responses, whose angle and velocity responses
    transform_responses (atmos/atmos_rad.cc:53-101) takes from the ray's
    frame to the atmosphere's; with the atmosphere-frame field per depth."""
    rows = []
    for i, r in enumerate(cfg0_atmosphere()):
        f = i / 60.0
        r = list(r)
        r[7] = 300.0 + 1200.0 * f
        r[9] = 1e5
        r[10] = math.radians(20.0 + 130.0 * f)
        r[11] = math.radians(-60.0 + 140.0 * f)
        rows.append(tuple(r))
    lam = np.round(6301.5 + np.linspace(-0.3, 0.3, 31), 6)
    setup_dir(d, rows, lam)
    el, az = 2.4, 0.6
    out = run(d, ["resp", "atmos.cfg", "lambda.txt", repr(el), repr(az)])
    rec = parse_synth(out)
    thermo, dpop, rs, bv = [], [], [], []
    for ln in out:
        t = ln.split()
        if t and t[0] == "thermo":
            thermo.append([float(x) for x in t[1:]])
        elif t and t[0] == "dpop":
            dpop.append([float(x) for x in t[1:]])
        elif t and t[0] == "resp":
            rs.append([float(x) for x in t[1:]])
        elif t and t[0] == "bvec":
            bv.append([float(x) for x in t[2:]])
    rs = np.array(rs)
    nd, nl = len(rows), len(lam)
    resp = np.zeros((7, nd, nl, 4))
    for r in rs:
        resp[int(r[0]) - 1, int(r[1]) - 1, int(r[2]) - 1, :] = r[3:]
    save_synth(rec, rows, "cfg0_resp_incl.npz",
               extra=dict(thermo=np.array(thermo), dpop=np.array(dpop), resp=resp, bvec=np.array(bv).T.copy(),
                          el=np.array(el), az=np.array(az)))
    print("cfg0_resp_incl: nlambda=%d" % nl)


def build_formal_scalar(d):
    """Scalar formal solution + Lambda* of the reference on its own scalar
    opacity (atmosphere::opacity/emissivity) for a few wavelengths, outgoing
    and ingoing rays."""
    rows = cfg0_atmosphere()
    lam = np.array([6301.2, 6301.5, 6302.5, 6303.0])
    setup_dir(d, rows, lam)
    rec = {}
    for tag, th in (("out", repr(math.pi)), ("in", "0.0")):
        opem, sl, grid = [], [], []
        for ln in run(d, ["formal_scalar", "atmos.cfg", "lambda.txt", th, "0.0"]):
            t = ln.split()
            if t[0] == "opem1":
                opem.append([float(x) for x in t[1:]])
            elif t[0] == "sl":
                sl.append([float(x) for x in t[1:]])
            elif t[0] == "grid":
                grid.append(float(t[2]))
        rec[tag + "_opem"] = np.array(opem)
        rec[tag + "_sl"] = np.array(sl)
        rec["grid"] = np.array(grid)
    np.savez_compressed(os.path.join(HERE, "formal_scalar.npz"), **rec)
    print("formal_scalar: %d wavelengths" % len(lam))


def build_lmstep(d):
    recs = {}
    for nrow, npar, lm, seed in ((240, 14, 1e-2, 1), (120, 14, 10.0, 2), (60, 7, 1e3, 3)):
        out = run(d, ["lmstep", str(nrow), str(npar), repr(lm), str(seed)], he=False)
        J, r, JTJ, rhs, corr = [], [], [], [], None
        for ln in out:
            t = ln.split()
            if t[0] == "J":
                J.append([float(x) for x in t[2:2 + npar]])
                r.append(float(t[-1]))
            elif t[0] == "JTJ":
                JTJ.append([float(x) for x in t[2:2 + npar]])
                rhs.append(float(t[-1]))
            elif t[0] == "corr":
                corr = [float(x) for x in t[1:]]
        key = "s%d" % seed
        recs[key + "_J"] = np.array(J)
        recs[key + "_r"] = np.array(r)
        recs[key + "_JTJ"] = np.array(JTJ)
        recs[key + "_rhs"] = np.array(rhs)
        recs[key + "_corr"] = np.array(corr)
        recs[key + "_lm"] = np.array(lm)
    np.savez_compressed(os.path.join(HERE, "lmstep.npz"), **recs)
    print("lmstep: %d cases" % 3)


def build_conv(d):
    """Gaussian spectral convolution of the reference (convolve_spectra_with_gauss)
    on seeded spectra: even and uneven grids, widths of 10..60 mA."""
    recs = {}
    for key, args in (("c0", ["201", "0.025", "31", "0"]), ("c1", ["151", "0.04", "32", "1"]),
                      ("c2", ["301", "0.01", "33", "0"]), ("c3", ["97", "0.06", "34", "1"])):
        out = run(d, ["conv"] + args, he=False)
        rin = np.array([[float(x) for x in ln.split()[1:]] for ln in out if ln.startswith("in ")])
        rout = np.array([[float(x) for x in ln.split()[1:]] for ln in out if ln.startswith("out ")])
        recs[key + "_lam"] = rin[:, 0]
        recs[key + "_in"] = rin[:, 1:]
        recs[key + "_out"] = rout[:, 1:]
        recs[key + "_width"] = np.array(float(args[1]) * 1e-8)
    np.savez_compressed(os.path.join(HERE, "conv.npz"), **recs)
    print("conv: 4 cases")


def build_ana(d):
    """ANA f0 files written by the reference's ana_fzwrite in write_file's
    3-D float and 4-D double layouts (the nodes and spectra cubes), and the
    values in the reference's index order."""
    out = run(d, ["ana", os.path.join(d, "ref")], he=False)
    a3 = np.zeros((3, 4, 5))
    a4 = np.zeros((2, 3, 4, 6))
    for ln in out:
        t = ln.split()
        if t[0] == "a3":
            a3[int(t[1]) - 1, int(t[2]) - 1, int(t[3]) - 1] = float(t[4])
        elif t[0] == "a4":
            a4[int(t[1]) - 1, int(t[2]) - 1, int(t[3]) - 1, int(t[4]) - 1] = float(t[5])
    for n in ("3d", "4d"):
        shutil.copyfile(os.path.join(d, "ref_%s.f0" % n), os.path.join(HERE, "ana_ref_%s.f0" % n))
    np.savez_compressed(os.path.join(HERE, "ana.npz"), a3=a3, a4=a4)
    print("ana: 2 files")


def build_interp(d):
    """Node interpolation: non-temperature (is_temp 0) for 1..5 nodes, keys
    n<k>_*; temperature (is_temp 1: flat top, bottom slope from the FAL-C
    table at the deepest node, clamped to 3400..20000 K) for 1..7 nodes with
    the deepest node below, inside and above the table's range, keys t<k>_*."""
    recs = {}
    cases = [("n%d" % nn, [str(nn), "0", str(seed)]) for nn, seed in ((1, 11), (2, 12), (3, 13), (4, 14), (5, 15))]
    cases += [("t1", ["1", "1", "21"]), ("t2", ["2", "1", "22", "-1.0"]), ("t3", ["3", "1", "23", "0.1"]),
              ("t4", ["4", "1", "24", "0.55"]), ("t5", ["5", "1", "25", "1.0"]), ("t6", ["6", "1", "26", "1.5"]),
              ("t7", ["7", "1", "27", "0.9"])]
    for key, args in cases:
        nn = key
        out = run(d, ["interp"] + args, he=False)
        q = []
        for ln in out:
            t = ln.split()
            if t[0] == "nodes":
                v = [float(x) for x in t[1:]]
                recs["%s_tau" % nn] = np.array(v[0::2])
                recs["%s_val" % nn] = np.array(v[1::2])
            elif t[0] == "q":
                q.append([float(t[1]), float(t[2])])
        q = np.array(q)
        recs["%s_grid" % nn] = q[:, 0]
        recs["%s_q" % nn] = q[:, 1]
    np.savez_compressed(os.path.join(HERE, "interp.npz"), **recs)
    print("interp: %d node sets" % len(cases))


NLTE_ATOM = "calcium_ir_rh.cfg"


def build_nlte(d):
    """A Ca II 5-level + continuum NLTE atom (the reference's calcium_ir_rh.cfg)
    in the cfg0 column with the background set: inputs of one ALI iteration,
    its sums and new populations, the iteration history and the converged
    populations of the reference's nltepops (atmos/atmos_pop.cc:67-240)."""
    bg = [a for a in BACKGROUND if a != "calcium.cfg"]
    for a in bg + [NLTE_ATOM]:
        shutil.copy(os.path.join(REF, "cfg", a), os.path.join(d, a))
    shutil.copy(os.path.join(REF, "cfg", "mol.cfg"), os.path.join(d, "mol.cfg"))
    with open(os.path.join(d, "atoms.cfg"), "w") as f:
        for a in bg + [NLTE_ATOM]:
            f.write("ATOM=%s\n" % a)
    rows = cfg0_atmosphere()
    with open(os.path.join(d, "atmos.cfg"), "w") as f:
        f.write("ID=cfg0\nTYPE=SPINOR\nFILE=atmos.dat\nGEOMETRY=PLANEPARALLEL\nRTSOLVER=Bezier\n"
                "ATOM=atoms.cfg\nMOL=mol.cfg\nELECTRONS=0\nTAU_GRID=1\nLEVEL_POPS=0\n"
                "NX=1\nNY=1\nNZ=%d\n" % len(rows))
    write_spinor(os.path.join(d, "atmos.dat"), rows)
    out = run(d, ["nlte", "atmos.cfg", "100"])
    r = {}
    trans, sig, lam, rays, depth, pop0, gam, coll, vproj, bg, acc, newpop, iters, final, ref = \
        [], [], [], [], [], [], [], [], [], [], [], [], [], [], []
    for ln in out:
        t = ln.split()
        k = t[0]
        if k in ("nd", "ntp", "nlam", "tau_grid", "ref_rv"):
            r[k] = int(t[1])
        elif k == "boundary":
            r[k] = float(t[1])
        elif k == "atom":
            r["nmap"], r["ntr"], r["mass"] = int(t[3]), int(t[5]), float(t[9])
        elif k == "ray":
            rays.append([float(t[2]), float(t[3])])
        elif k == "lam":
            lam.append([float(x) for x in t[2:5]])
        elif k == "trans":
            if t[2] == "line":
                trans.append([0, int(t[3]), int(t[4])] + [float(x) for x in t[5:10]] + [0.0, 0.0])
            else:
                trans.append([1, int(t[3]), int(t[4]), 0, 0, 0, 0, 0, float(t[5]), float(t[6])])
        elif k == "sigma":
            sig.append([int(t[1]), int(t[2]), float(t[3])])
        elif k == "depth":
            depth.append([float(x) for x in t[2:8]])
        elif k == "pop0":
            pop0.append([int(t[1]), int(t[2]), float(t[3])])
        elif k == "gamma":
            gam.append([int(t[1]), int(t[2]), float(t[3])])
        elif k == "coll":
            coll.append([int(t[1]), int(t[2]), int(t[3]), float(t[4])])
        elif k == "vproj":
            vproj.append([int(t[1]), int(t[2]), float(t[3])])
        elif k == "bg":
            bg.append([int(t[1]), int(t[2]), int(t[3]), float(t[4]), float(t[5])])
        elif k == "acc":
            acc.append([int(t[1]), int(t[2])] + [float(x) for x in t[3:7]])
        elif k == "newpop":
            newpop.append([int(t[1]), int(t[2]), float(t[3])])
        elif k == "iter":
            iters.append([int(t[1]), float(t[2]), int(t[3])])
        elif k == "final":
            final.append([int(t[1]), int(t[2]), float(t[3])])
        elif k == "ref_final":
            ref.append([int(t[1]), int(t[2]), float(t[3])])
    nd, nl, nt, nm, ntp = r["nd"], r["nlam"], r["ntr"], r["nmap"], r["ntp"]
    z0 = 1  # the reference's depth index starts at x3l = 1

    def grid3(rows, shape, idx):
        a = np.zeros(shape)
        for row in rows:
            a[idx(row)] = row[-1]
        return a
    rec = dict(nd=nd, nlam=nl, ntr=nt, nmap=nm, ntp=ntp, mass=r["mass"], boundary=r["boundary"],
               tau_grid=r["tau_grid"], ref_rv=r["ref_rv"], rays=np.array(rays), lam=np.array(lam),
               trans=np.array(trans), depth=np.array(depth))
    rec["sigma"] = grid3(sig, (nt, nl), lambda w: (w[0] - 1, w[1]))
    rec["pop0"] = grid3(pop0, (nd, nm), lambda w: (w[0] - z0, w[1]))
    rec["gamma"] = grid3(gam, (nt, nd), lambda w: (w[0] - 1, w[1] - z0))
    rec["coll"] = grid3(coll, (nd, nm, nm), lambda w: (w[0] - z0, w[1], w[2]))
    rec["vproj"] = grid3(vproj, (ntp, nd), lambda w: (w[0] - 1, w[1] - z0))
    b = np.array(bg)
    bgop = np.zeros((ntp, nd, nl))
    bgem = np.zeros((ntp, nd, nl))
    bgop[b[:, 0].astype(int) - 1, b[:, 2].astype(int) - z0, b[:, 1].astype(int)] = b[:, 3]
    bgem[b[:, 0].astype(int) - 1, b[:, 2].astype(int) - z0, b[:, 1].astype(int)] = b[:, 4]
    rec["bgop"], rec["bgem"] = bgop, bgem
    a = np.zeros((nd, nt, 4))
    for w in acc:
        a[w[0] - z0, w[1] - 1] = w[2:6]
    rec["acc"] = a
    rec["newpop"] = grid3(newpop, (nd, nm), lambda w: (w[0] - z0, w[1]))
    rec["iters"] = np.array(iters)
    rec["final"] = grid3(final, (nd, nm), lambda w: (w[0] - z0, w[1]))
    rec["ref_final"] = grid3(ref, (nd, nm), lambda w: (w[0] - z0, w[1]))
    np.savez_compressed(os.path.join(HERE, "nlte.npz"), **rec)
    print("nlte: %d depths, %d wavelengths, %d rays, %d iterations" % (nd, nl, ntp, len(iters)))


CHEM_MAXS, CHEM_MAXT, CHEM_MAXC = 10, 6, 8


def chemeq_grid_atmosphere():
    """A column whose depth points sweep temperature and gas pressure
    independently: 12 temperatures 3000-40000 K times 6 pressures
    1e-1-1e6 dyn cm^-2 (SPINOR rows; Pe = 1e-4 Pg is only a starting value)."""
    rows = []
    temps = [3000.0, 3500.0, 4000.0, 4500.0, 5000.0, 5800.0, 6500.0, 8000.0, 10000.0, 15000.0, 25000.0, 40000.0]
    pres = [1e-1, 1e1, 1e3, 1e4, 1e5, 1e6]
    i = 0
    for T in temps:
        for pg in pres:
            lt = -7.0 + 0.1 * i
            rows.append((lt, 1e7 * (1.0 - lt), T, pg, pg * 1e-4, 0.0, 0.0, 0.0, 1e5, 0.0, 0.0, 0.0))
            i += 1
    return rows


def parse_chemeq(out):
    """Parses the driver's `chemeq` records into dense arrays (element index
    1-based in the records, 0-based in the arrays)."""
    S, T_, C = CHEM_MAXS, CHEM_MAXT, CHEM_MAXC
    recs = [ln.split() for ln in out if ln.strip()]
    nel = next(int(t[1]) for t in recs if t[0] == "nelement")
    nmol = sum(1 for t in recs if t[0] == "mol")
    pts = [t for t in recs if t[0] == "point"]
    npt = len(pts)
    pidx = {int(t[1]): i for i, t in enumerate(pts)}
    r = dict(
        Z=np.zeros(nel, np.int32), abund=np.zeros(nel), eion=np.full((nel, S), np.nan),
        pf_type=np.zeros((nel, S), np.int32), pf_value=np.zeros((nel, S)), pf_charge=np.zeros((nel, S)),
        pf_g0=np.zeros((nel, S)), pf_nterm=np.zeros((nel, S), np.int32),
        term_g2=np.zeros((nel, S, T_)), term_xion=np.zeros((nel, S, T_)), term_xl=np.zeros((nel, S, T_)),
        term_ncoef=np.zeros((nel, S, T_), np.int32), term_gamma=np.zeros((nel, S, T_, C)),
        term_alpha=np.zeros((nel, S, T_, C)), irwin=np.zeros((nel, S, C)),
        mol_eqc=np.zeros(nmol, np.int32), mol_ncoef=np.zeros(nmol, np.int32), mol_coef=np.zeros((nmol, C)),
        mol_ncomp=np.zeros(nmol, np.int32), mol_elem=np.zeros((nmol, 4), np.int32),
        mol_stage=np.zeros((nmol, 4), np.int32),
        T=np.zeros(npt), Nt=np.zeros(npt), ne_in=np.zeros(npt),
        U=np.zeros((npt, nel, S)), dU=np.zeros((npt, nel, S)), fk=np.zeros((npt, nel, S)),
        df=np.zeros((npt, nel, S)), K=np.zeros((npt, nmol)), ne=np.zeros(npt), Nel=np.zeros((npt, nel)),
        N=np.zeros((npt, nel, S)), hmin=np.zeros(npt), opref=np.zeros(npt))
    lvl, pops, bfs, he = [], {}, [], []
    for t in recs:
        k = t[0]
        if k == "element":
            e = int(t[1]) - 1
            r["Z"][e], r["abund"][e] = int(t[2]), float(t[3])
        elif k == "eion":
            r["eion"][int(t[1]) - 1, int(t[2])] = float(t[3])
        elif k == "pf":
            e, z, ty = int(t[1]) - 1, int(t[2]), int(t[3])
            r["pf_type"][e, z] = ty
            if ty == 1:
                r["pf_value"][e, z] = float(t[4])
            elif ty == 3:
                n = int(t[4])
                r["pf_nterm"][e, z] = n
                r["irwin"][e, z, :n] = [float(x) for x in t[5:5 + n]]
            elif ty == 2:
                r["pf_charge"][e, z], r["pf_g0"][e, z], r["pf_nterm"][e, z] = float(t[4]), float(t[5]), int(t[6])
        elif k == "term":
            e, z, i = int(t[1]) - 1, int(t[2]), int(t[3])
            r["term_g2"][e, z, i], r["term_xion"][e, z, i], r["term_xl"][e, z, i] = map(float, t[4:7])
            m = int(t[8])
            r["term_ncoef"][e, z, i] = m
            v = [float(x) for x in t[9:9 + 2 * m]]
            r["term_gamma"][e, z, i, :m] = v[0::2]
            r["term_alpha"][e, z, i, :m] = v[1::2]
        elif k == "lvl":
            lvl.append([int(t[1]) - 1, int(t[2]), int(t[3]), int(t[4]), float(t[5])])
        elif k == "mol":
            m = int(t[1]) - 1
            r["mol_ncomp"][m], r["mol_eqc"][m], r["mol_ncoef"][m] = int(t[2]), int(t[3]), int(t[4])
            r["mol_coef"][m, :int(t[4])] = [float(x) for x in t[5:5 + int(t[4])]]
        elif k == "comp":
            m, i = int(t[1]) - 1, int(t[2])
            r["mol_elem"][m, i], r["mol_stage"][m, i] = int(t[3]), int(t[4])
        elif k == "point":
            i = pidx[int(t[1])]
            r["T"][i], r["Nt"][i], r["ne_in"][i] = float(t[2]), float(t[3]), float(t[4])
        elif k == "U":
            i, e, z = pidx[int(t[1])], int(t[2]) - 1, int(t[3])
            r["U"][i, e, z], r["dU"][i, e, z] = float(t[4]), float(t[5])
        elif k == "frc":
            i, e, z = pidx[int(t[1])], int(t[2]) - 1, int(t[3])
            r["fk"][i, e, z], r["df"][i, e, z] = float(t[4]), float(t[5])
        elif k == "K":
            r["K"][pidx[int(t[1])], int(t[2]) - 1] = float(t[3])
        elif k == "ne":
            r["ne"][pidx[int(t[1])]] = float(t[2])
        elif k == "Nt":
            r["Nel"][pidx[int(t[1])], int(t[2]) - 1] = float(t[3])
        elif k == "N":
            r["N"][pidx[int(t[1])], int(t[2]) - 1, int(t[3])] = float(t[4])
        elif k == "n":
            pops[(pidx[int(t[1])], int(t[2]) - 1, int(t[3]), int(t[4]))] = float(t[5])
        elif k == "hmin":
            r["hmin"][pidx[int(t[1])]] = float(t[2])
        elif k == "opref":
            r["opref"][pidx[int(t[1])]] = float(t[2])
        elif k == "bf":
            bfs.append([int(t[1]) - 1, int(t[2]), int(t[3]), int(t[4]), float(t[5]), int(t[6]), float(t[7]),
                        float(t[8]), float(t[9])])
        elif k == "lref":
            r["lref"] = np.array([float(t[1]), float(t[2])])
        elif k == "he":
            he.append([float(x) for x in t[2:8]])
    # bound-free entries: element (0-based), stage, level, g, energy, g and energy of the next
    # stage's ground level, cross-section at airtovac(5000 A) and at 5000 A
    r["bf_int"] = np.array([[b[0], b[1], b[2], b[3], b[5]] for b in bfs], np.int32).reshape(-1, 5)
    r["bf_real"] = np.array([[b[4], b[6], b[7], b[8]] for b in bfs]).reshape(-1, 4)
    r["he"] = np.array(he).reshape(-1, 6)  # T, log tau, Nt, Ne, op_ref (5000 A), rho after the integration
    r["levels"] = np.array([row[:4] for row in lvl], np.int32)
    r["level_ee"] = np.array([row[4] for row in lvl])
    n = np.zeros((npt, len(lvl)))
    for j, row in enumerate(lvl):
        for i in range(npt):
            n[i, j] = pops[(i, row[0], row[1], row[2])]
    r["n"] = n
    return r


def build_chemeq(d):
    """Chemical equilibrium and LTE populations (atmos/atmos_chm.cc:28-363,
    atom.cc:763-787) for the cfg0 species set (eight background atoms, the
    two-line Fe atom and the reference's 13 molecules) on two columns: the
    temperature/pressure sweep of `chemeq_grid_atmosphere` and the cfg0
    column after the reference's hydrostatic integration."""
    recs = []
    for rows, he in ((chemeq_grid_atmosphere(), False), (cfg0_atmosphere(), True)):
        setup_dir(d, rows, cfg0_lambda())
        recs.append(parse_chemeq(run(d, ["chemeq", "atmos.cfg"], he=he)))
    a, b = recs
    pointwise = ("T", "Nt", "ne_in", "U", "dU", "fk", "df", "K", "ne", "Nel", "N", "hmin", "n", "opref")
    out = {k: (np.concatenate([a[k], b[k]]) if k in pointwise else a[k]) for k in a}
    out["he"] = b["he"]
    for k in a:
        if k not in pointwise and k != "he":
            assert np.array_equal(np.nan_to_num(a[k]), np.nan_to_num(b[k])), k
    out["ncolumn_points"] = np.array([len(a["T"]), len(b["T"])], np.int32)
    np.savez_compressed(os.path.join(HERE, "chemeq.npz"), **out)
    print("chemeq: %d elements, %d molecules, %d levels, %d points" %
          (len(a["Z"]), len(a["mol_eqc"]), len(a["levels"]), len(out["T"])))


def nlte_dir(d):
    """Config directory of the NLTE column: the background set without Ca,
    the reference's Ca II 5-level + continuum atom, the cfg0 atmosphere."""
    bg = [a for a in BACKGROUND if a != "calcium.cfg"]
    for a in bg + [NLTE_ATOM]:
        shutil.copy(os.path.join(REF, "cfg", a), os.path.join(d, a))
    shutil.copy(os.path.join(REF, "cfg", "mol.cfg"), os.path.join(d, "mol.cfg"))
    with open(os.path.join(d, "atoms.cfg"), "w") as f:
        for a in bg + [NLTE_ATOM]:
            f.write("ATOM=%s\n" % a)
    rows = cfg0_atmosphere()
    with open(os.path.join(d, "atmos.cfg"), "w") as f:
        f.write("ID=cfg0\nTYPE=SPINOR\nFILE=atmos.dat\nGEOMETRY=PLANEPARALLEL\nRTSOLVER=Bezier\n"
                "ATOM=atoms.cfg\nMOL=mol.cfg\nELECTRONS=0\nTAU_GRID=1\nLEVEL_POPS=0\n"
                "NX=1\nNY=1\nNZ=%d\n" % len(rows))
    write_spinor(os.path.join(d, "atmos.dat"), rows)


def build_nlteresp(d):
    """NLTE population responses of the Ca II atom in the cfg0 column
    (atmosphere::compute_nlte_population_responses after nltepops): the
    assembled (nd*nmap)^2 system, its right-hand sides for T, density and
    vmicro at every depth and the level responses; and formal_with_responses
    at every 74th wavelength of the NLTE grid for both rays."""
    nlte_dir(d)
    out = run(d, ["nlteresp", "atmos.cfg", "74"])
    head = {}
    Ms, betas, resps, grid, probs, opem, R = [], [], [], [], [], [], []
    for ln in out:
        t = ln.split()
        k = t[0]
        if k in ("nd", "nmap", "tau_grid"):
            head[k] = int(t[1])
        elif k == "boundary":
            head[k] = float(t[1])
        elif k == "M":
            Ms.append((int(t[1]), int(t[2]), float(t[3])))
        elif k == "beta":
            betas.append((int(t[1]), int(t[2]), float(t[3]), float(t[4]), float(t[5])))
        elif k == "resp":
            resps.append((int(t[1]), int(t[2]), int(t[3]), float(t[4])))
        elif k == "grid":
            grid.append(float(t[2]))
        elif k == "prob":
            probs.append((int(t[2]), int(t[3]), float(t[4]), float(t[5])))
        elif k == "opem":
            opem.append((int(t[1]), int(t[2]), float(t[3]), float(t[4]), float(t[5])))
        elif k == "R":
            R.append((int(t[1]), int(t[2]), int(t[3]), float(t[4]), float(t[5])))
    nd, nm = head["nd"], head["nmap"]
    n = nd * nm
    M = np.zeros((n, n))
    for r, c, v in Ms:
        M[r, c] = v
    beta = np.zeros((3, nd, n))
    for kk, r, a, b, c in betas:
        beta[:, kk, r] = (a, b, c)
    resp = np.zeros((3, n, nd))
    for p_, r, kk, v in resps:
        resp[p_ - 1, r, kk] = v
    npr = len(probs)
    oe = np.zeros((npr, 3, nd))
    for pr, z, a, b, c in opem:
        oe[pr, :, z] = (a, b, c)
    RR = np.zeros((npr, 2, nd, nd))
    for pr, z, zp, a, b in R:
        RR[pr, :, z, zp] = (a, b)
    rec = dict(nd=nd, nmap=nm, tau_grid=head["tau_grid"], boundary=head["boundary"], M=M, beta=beta, resp=resp,
               grid=np.array(grid), prob=np.array(probs), op=oe[:, 0], em=oe[:, 1], I=oe[:, 2], Rop=RR[:, 0],
               Rem=RR[:, 1])
    np.savez_compressed(os.path.join(HERE, "nlteresp.npz"), **rec)
    print("nlteresp: system %d x %d, %d formal problems" % (n, n, npr))


# The 14-node model of the inversion workload (7 T, 1 vt, 2 vLOS, 2 B, 1
# inclination, 1 azimuth; codes as the reference's model orders them) and two
# node sets: a field inclined towards the observer and one pointing away with
# a negative azimuth, so that the acos/atan folds of the field projection are
# both exercised.
FIT_NODES = [(1, -5.0), (1, -4.0), (1, -3.0), (1, -2.0), (1, -1.0), (1, 0.0), (1, 0.5), (2, -2.0), (3, -4.0),
             (3, 0.0), (4, -4.0), (4, 0.0), (5, -2.0), (6, -2.0)]
NODE_SETS = [
    [4600.0, 4450.0, 4500.0, 4900.0, 5300.0, 6300.0, 7600.0, 1.1e5, 0.8e5, -0.5e5, 900.0, 1300.0, 0.7, 0.4],
    [4700.0, 4400.0, 4550.0, 4800.0, 5400.0, 6200.0, 7400.0, 0.6e5, -1.2e5, 0.3e5, 2200.0, 1500.0, 2.6, -0.9],
]


def fit_lambda():
    """Every third point of the 301-point grid: 101 points over both lines."""
    return cfg1_lambda()[::3]


def write_model(path, values, tau_min=-6.0, tau_max=1.0, nodes=None, reg=()):
    """The driver's model file; reg: (code, type, alpha) entries."""
    nodes = FIT_NODES if nodes is None else nodes
    with open(path, "w") as f:
        f.write("%.17g %.17g\n%d\n" % (tau_min, tau_max, len(nodes)))
        for (code, lt), v in zip(nodes, values):
            f.write("%d %.17g %.17g\n" % (code, lt, v))
        f.write("%d\n" % len(reg))
        for code, typ, alpha in reg:
            f.write("%d %d %.17g\n" % (code, typ, alpha))


def run_nodes(d, values, tag):
    write_model(os.path.join(d, "model_%s.txt" % tag), values)
    out = run(d, ["nodes", "atmos.cfg", "lambda.txt", repr(math.pi), "0.0", "model_%s.txt" % tag], he=False)
    rec = {k: [] for k in ("bfn", "proj", "fwd", "stokes", "jac", "nmap", "model", "post")}
    for ln in out:
        t = ln.split()
        if t and t[0] in rec:
            rec[t[0]].append([float(x) for x in t[1:]])
    nl = len(rec["fwd"])
    npar = len(FIT_NODES)
    nd = len(rec["bfn"])
    jac = np.zeros((npar, nl, 4))
    for r in rec["jac"]:
        jac[int(r[0]) - 1, int(r[1]) - 1] = r[2:]
    nmap = np.zeros((npar, 7, nd))
    for r in rec["nmap"]:
        nmap[int(r[0]) - 1, int(r[1]) - 1, int(r[2])] = r[3]
    return dict(bfn=np.array(rec["bfn"])[:, 1:], proj=np.array(rec["proj"])[:, 1:], fwd=np.array(rec["fwd"])[:, 1:],
                stokes=np.array(rec["stokes"])[:, 1:], jac=jac, nmap=nmap, model_after=np.array(rec["model"][0][1:]),
                post=np.array(rec["post"])[:, 1:])


def build_nodes(d):
    """obs_stokes_responses_to_nodes with build_from_nodes for the 14-node
    model on the 71-depth grid (log tau -6..1) at two node sets: the built
    state (tau, T, Nt, Ne, op_ref, rho, vt, vz, Bx, By, Bz), the ray
    projection (vLOS, |B|, inclination and azimuth in degrees), the
    synthesis-path Stokes (forward_evaluate), the response-path Stokes, the
    Jacobian [node][lambda][4] and the node responses [node][7][depth]."""
    rows = cfg0_atmosphere(71)
    lam = fit_lambda()
    setup_dir(d, rows, lam)
    recs = dict(codes=np.array([c for c, _ in FIT_NODES]), logtau=np.array([t for _, t in FIT_NODES]), lam_air=lam)
    for i, vals in enumerate(NODE_SETS):
        r = run_nodes(d, vals, "n%d" % i)
        recs["values%d" % i] = np.array(vals)
        for k, v in r.items():
            recs["%s%d" % (k, i)] = v
    np.savez_compressed(os.path.join(HERE, "nodes.npz"), **recs)
    print("nodes: %d node sets, %d wavelengths" % (len(NODE_SETS), len(lam)))


def _fit_cases():
    """Three pixels: observations are the reference's synthesis-path spectrum
    at node set 0 or 1 plus Gaussian noise of 1e-3 of the first wavelength's
    intensity; starts perturbed by 120 K rms in temperature and +-20 % in the
    other nodes; pixel 1 with 30 % of the wavelengths masked; pixel 2 starts
    with B = 0.5 G and vt = 50 cm/s, below the fit's bracket."""
    rng = np.random.default_rng(2024)
    cases = []
    for k, base in enumerate((0, 1, 0)):
        truth = np.array(NODE_SETS[base])
        start = truth.copy()
        start[:7] += rng.normal(0.0, 120.0, 7)
        start[7:] *= rng.uniform(0.8, 1.2, 7)
        if k == 2:
            start[7] = 50.0
            start[10] = 0.5
        cases.append(dict(base=base, truth=truth, start=start, seed=100 + k, masked=(k == 1)))
    return cases


# the model's regularisation in the regularised fit: temperature Tikhonov,
# vLOS Tikhonov, |B| departure from zero, alphas of the order of the square
# roots of the scaled J^T J diagonal of these parameters
REG_FIT = [(1, 2, 3e15), (3, 2, 1e14), (4, 1, 2e14)]


def build_fitreg(d):
    """The fit of build_fit with the model regularised (REG_FIT), pixels 0
    and 1, 1..6 iterations."""
    build_fit(d, kmax=6, reg=REG_FIT, name="fitreg", pixels=(0, 1))


def build_fit(d, kmax=8, reg=(), name="fit", pixels=(0, 1, 2)):
    """atmosphere::stokes_lm_fit on three pixels of the 14-node model, run with
    1..kmax iterations (the nodes and synthesis-path chi^2 after each, i.e.
    the trajectory), weights (1, 5, 5, 3), start lambda 1; reg the model's
    regularisation entries."""
    rows = cfg0_atmosphere(71)
    lam = fit_lambda()
    setup_dir(d, rows, lam)
    ws = np.array([1.0, 5.0, 5.0, 3.0])
    recs = dict(codes=np.array([c for c, _ in FIT_NODES]), logtau=np.array([t for _, t in FIT_NODES]), lam_air=lam,
                ws=ws, kmax=np.array(kmax), lm_start=np.array(1.0), reg=np.array(reg, dtype=np.float64).reshape(-1, 3))
    procs = []
    for k, cs in enumerate(_fit_cases()):
        if k not in pixels:
            continue
        fwd = run_nodes(d, cs["truth"], "t%d" % k)["fwd"]
        rng = np.random.default_rng(cs["seed"])
        obs = fwd + rng.normal(0.0, 1e-3 * fwd[0, 0], fwd.shape)
        mask = np.ones(len(lam))
        if cs["masked"]:
            mask[rng.permutation(len(lam))[: int(0.3 * len(lam))]] = 0.0
        path = os.path.join(d, "fit_%d.txt" % k)
        write_model(path, cs["start"], reg=reg)
        with open(path, "a") as f:
            f.write("%.17g %.17g %.17g %.17g %.17g %d\n" % (ws[0], ws[1], ws[2], ws[3], 1.0, kmax))
            f.write(" ".join("%.17g" % m for m in mask) + "\n")
            for row in obs:
                f.write(" ".join("%.17g" % v for v in row) + "\n")
        recs.update({"truth%d" % k: cs["truth"], "start%d" % k: cs["start"], "obs%d" % k: obs, "mask%d" % k: mask})
        env = dict(os.environ)
        env["SNAPI_REF_HE"] = "0"
        procs.append(subprocess.Popen([DRIVER, "fit", "atmos.cfg", "lambda.txt", repr(math.pi), "0.0", path], cwd=d,
                                      env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True))
    for k, pr in zip(pixels, procs):
        out, err = pr.communicate()
        if pr.returncode:
            raise RuntimeError("fit driver failed on pixel %d" % k)
        # stderr of the longest run: one "test chi-squared" line per iteration
        # (the current chi^2 / (4 nlambda 1e-6), atmos_fit.cc:157-158)
        runs, cur = {}, None
        for ln in err.splitlines():
            if ln.startswith("fitrun "):
                cur = int(ln.split()[1])
                runs[cur] = []
            elif "test chi-squared" in ln and cur is not None:
                runs[cur].append(float(ln.split("=")[1].split(",")[0]) * 4.0 * len(lam) * 1e-6)
        recs["iters%d" % k] = np.array([len(runs.get(i, [])) for i in range(1, kmax + 1)])
        recs["chicur%d" % k] = np.array(runs.get(kmax, []))
        nodes, chi, S = [], [], []
        for ln in out.splitlines():
            t = ln.split()
            if not t:
                continue
            if t[0] == "fitk":
                nodes.append([float(x) for x in t[2:]])
            elif t[0] == "fitchi":
                chi.append(float(t[2]))
            elif t[0] == "fitS":
                S.append([float(x) for x in t[2:]])
        recs["nodes%d" % k] = np.array(nodes)
        recs["chi%d" % k] = np.array(chi)
        recs["S%d" % k] = np.array(S)
        print("fit pixel %d: chi2 %s" % (k, np.array2string(np.array(chi), precision=3)))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **recs)
    print("%s: %d pixels, %d iterations" % (name, len(procs), kmax))


def build_psf(d):
    """convolve_spectra_with_psf of the reference on seeded spectra and
    PSFs (the driver hands the samples over so that the reference's window,
    read from index 0 of its array, holds them in order): odd sample
    counts, whose window is the defined centred one from the (npsf-1)/2-th
    wavelength on."""
    recs = {}
    for key, args in (("p0", ["101", "7", "41"]), ("p1", ["60", "13", "42"]), ("p2", ["31", "1", "43"]),
                      ("p3", ["40", "31", "44"])):
        out = run(d, ["psf"] + args, he=False)
        recs[key + "_psf"] = np.array([float(x) for x in out[0].split()[1:]])
        recs[key + "_in"] = np.array([[float(x) for x in ln.split()[1:]] for ln in out if ln.startswith("in ")])
        recs[key + "_out"] = np.array([[float(x) for x in ln.split()[1:]] for ln in out if ln.startswith("out ")])
    np.savez_compressed(os.path.join(HERE, "psf.npz"), **recs)
    print("psf: 4 cases")


REG_CASES = [
    # (nodes, [(code, type, alpha)])
    (FIT_NODES, [(1, 2, 0.5), (3, 1, 2.0)]),            # T Tikhonov (7 nodes), vLOS zero (2)
    (FIT_NODES, [(4, 2, 1.3), (2, 1, 0.7)]),            # |B| Tikhonov (2), vt zero (1)
    (FIT_NODES, [(5, 2, 0.9), (1, 1, 0.1)]),            # one inclination node: couples to the azimuth node
    (FIT_NODES, [(6, 1, 3.0), (4, 1, 0.2)]),            # the last parameter's run ends the list: nothing
    ([(1, -4.0), (1, -2.0), (1, 0.0), (4, -3.0), (4, -1.0)], [(4, 2, 1.1), (1, 2, 0.4)]),  # last run, Tikhonov
]


def build_reg(d):
    """atmosphere::regularize_hessian of the reference on a zero J^T J: the
    matrix it adds, for node models with Tikhonov and departure-from-zero
    regularisation, including the runs that end the node list."""
    rows = cfg0_atmosphere(71)
    setup_dir(d, rows, fit_lambda())
    recs = {}
    for k, (nodes, reg) in enumerate(REG_CASES):
        path = os.path.join(d, "reg_%d.txt" % k)
        write_model(path, [5000.0 if c == 1 else 1.0 for c, _ in nodes], nodes=nodes, reg=reg)
        out = run(d, ["reg", "atmos.cfg", path], he=False)
        G = np.array([[float(x) for x in ln.split()[2:-2]] for ln in out if ln.startswith("G ")])
        rhs = np.array([float(ln.split()[-1]) for ln in out if ln.startswith("G ")])
        recs["codes%d" % k] = np.array([c for c, _ in nodes])
        recs["reg%d" % k] = np.array(reg, dtype=np.float64)
        recs["G%d" % k] = G
        recs["rhs%d" % k] = rhs
    np.savez_compressed(os.path.join(HERE, "reg.npz"), **recs)
    print("reg: %d models" % len(REG_CASES))


def build_nlterin(d):
    """The inputs of the NLTE population-response system of the Ca II atom
    in the cfg0 column (nltepops, then what compute_nlte_population_responses
    feeds add_response_contributions_new and its rate part): state and
    derivatives per depth, populations, line damping, collision rates and
    their derivatives, per ray and wavelength the projected velocity, total
    opacity and emissivity and the LTE background's perturbations; the
    profiles and profile norms at every 40th wavelength; and the system
    (M, beta) the reference assembles from them."""
    nlte_dir(d)
    out = run(d, ["nlterin", "atmos.cfg"])
    head, rays, lams, levels, trans, sigma = {}, [], [], [], [], {}
    depth, pops, damp, coll, vr, norm, pp, prof, Ms, betas = [], [], [], [], [], [], [], [], [], []
    for ln in out:
        t = ln.split()
        k = t[0]
        if k in ("nd", "ntp", "nlam", "tau_grid"):
            head[k] = int(t[1])
        elif k == "boundary":
            head[k] = float(t[1])
        elif k == "atom":
            head.update(nmap=int(t[2]), ntr=int(t[4]), Z=int(t[6]), mass=float(t[8]), saha=float(t[10]),
                        dT=float(t[12]), dN=float(t[14]), damp=int(t[16]))
        elif k == "ray":
            rays.append([float(t[2]), float(t[3])])
        elif k == "lam":
            lams.append([float(t[2]), float(t[3])])
        elif k == "level":
            levels.append([float(x) for x in t[1:]])
        elif k == "trans":
            if t[2] == "line":
                trans.append([int(t[1]), 0, int(t[3]), int(t[4]), int(t[5]), int(t[6]), float(t[7]), float(t[8]),
                              float(t[9]), 0.0])
            else:
                trans.append([int(t[1]), 1, int(t[3]), int(t[4]), int(t[5]), int(t[6]), 0.0, 0.0, 0.0, float(t[7])])
        elif k == "sigma":
            sigma.setdefault(int(t[1]), []).append(float(t[3]))
        elif k == "depth":
            depth.append([float(x) for x in t[2:]])
        elif k == "pop":
            pops.append(float(t[3]))
        elif k == "damp":
            damp.append([int(t[1]), int(t[2]), float(t[3]), float(t[4]), float(t[5])])
        elif k == "coll":
            coll.append([float(t[4]), float(t[5]), float(t[6])])
        elif k == "vr":
            vr.append(float(t[3]))
        elif k == "norm":
            norm.append([float(x) for x in t[4:]])
        elif k == "pp":
            pp.append([float(x) for x in t[4:]])
        elif k == "prof":
            prof.append(float(t[5]))
        elif k == "M":
            Ms.append((int(t[1]), int(t[2]), float(t[3])))
        elif k == "beta":
            betas.append((int(t[1]), int(t[2]), float(t[3]), float(t[4]), float(t[5])))
    nd, ntp, nlam, nm, ntr = head["nd"], head["ntp"], head["nlam"], head["nmap"], head["ntr"]
    n = nd * nm
    M = np.zeros((n, n))
    for r, c, v in Ms:
        M[r, c] = v
    beta = np.zeros((3, nd, n))
    for kk, r, a, b, c in betas:
        beta[:, kk, r] = (a, b, c)
    dm = np.zeros((3, ntr, nd))
    for tr, z, a, b, c in damp:
        dm[:, tr - 1, z] = (a, b, c)
    sg = np.zeros((ntr, nlam))
    for tr, v in sigma.items():
        sg[tr - 1] = v
    nsample = len(range(0, nlam, 40))
    recs = dict(head=np.array([nd, ntp, nlam, nm, ntr, head["Z"], head["tau_grid"], head["damp"]]),
                consts=np.array([head["mass"], head["saha"], head["dT"], head["dN"], head["boundary"]]),
                rays=np.array(rays), lam=np.array(lams), levels=np.array(levels), trans=np.array(trans), sigma=sg,
                depth=np.array(depth), pops=np.array(pops).reshape(nd, nm), damp=dm,
                coll=np.array(coll).reshape(nd, nm, nm, 3), vr=np.array(vr).reshape(ntp, nd),
                norm=np.array(norm).reshape(ntp, nd, ntr, 4), pp=np.array(pp).reshape(ntp, nlam, nd, 8),
                prof=np.array(prof).reshape(ntp, nsample, nd, ntr), M=M, beta=beta)
    np.savez_compressed(os.path.join(HERE, "nlterin.npz"), **recs)
    print("nlterin: %d rays x %d wavelengths, %d depths, %d levels" % (ntp, nlam, nd, nm))


def build_tau(d):
    """The reference's own test atmosphere test/inverted_atmos.f0 (the
    falc_1000 column on its height grid as the reference wrote it back): its
    log tau column is compute_tau_referent of its height and referent-opacity
    columns (tau = 1e-8 at the top).  Stored: height, opacity, log tau."""
    sys.path.insert(0, REPO)
    from snapi_amd import cube
    a = np.asarray(cube.ana_read(os.path.join(REF, "test", "inverted_atmos.f0"))).reshape(12, -1)
    np.savez_compressed(os.path.join(HERE, "tau_scale.npz"), logtau=a[0], height=a[1], opref=a[6], T=a[2])
    print("tau_scale: %d depths" % a.shape[1])


BUILDERS = ["cfg0", "cfg1_template", "resp", "resp_incl", "formal_scalar", "lmstep", "interp", "conv", "ana", "nlte",
            "chemeq", "nlteresp", "nodes", "fit", "tau", "psf", "reg", "fitreg", "nlterin"]


def main():
    if not os.path.exists(DRIVER):
        sys.exit("build the reference first: make -f oracle/ref.mk -j8")
    want = sys.argv[1:] or BUILDERS
    d = tempfile.mkdtemp(prefix="snapi_golden_")
    try:
        for name in want:
            globals()["build_" + name](d)
    finally:
        shutil.rmtree(d)


if __name__ == "__main__":
    main()
