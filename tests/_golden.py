"""Turns the reference dumps under tests/golden/ into the library's inputs
(line table, column state, wavelength grid).  Test infrastructure."""
import os

import numpy as np

from snapi_amd import abi
from snapi_amd.physics import planck, THOMSON

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(HERE, name))


def cfg0():
    """Returns dict(table, ray, state[1][nf][nd], lam_vac, stokes_ref, ...)."""
    d = load("cfg0_synth.npz")
    dep = d["depth"]
    nd = dep.shape[0]
    T, Ne, Nt, Vt, Vr, Bm, Bi, Ba, opref, tauref, grid, nhi, hgt = dep.T
    lam_vac = d["lam"][:, 1]
    lam_mid = 0.5 * (lam_vac[0] + lam_vac[-1])
    masses = dict((int(k), v) for k, v in d["masses"])
    lines = []
    lpop = d["lpop"]
    zc = d["zc"]
    pops = []
    for (s, z, lo, up, tr), ld, sp in zip(d["lines"], d["line_data"], d["line_split"]):
        lam0, A, Blu, Bul, grad, sigma, alpha = ld
        comps = [[], [], []]
        for row in zc:
            if int(row[0]) == s and int(row[1]) == tr:
                comps[int(row[2])].append((row[3], row[4]))
        lines.append(dict(lam0=lam0, aul=A, blu=Blu, bul=Bul, grad=grad, sigma=sigma, alpha=alpha,
                          mass=masses[int(s)], split=int(sp), comps=comps, species=int(s), tr=int(tr)))
        m = (lpop[:, 0] == s) & (lpop[:, 1] == tr)
        rows = lpop[m]
        rows = rows[np.argsort(rows[:, 2])]
        pops.append((rows[:, 3], rows[:, 4], rows[:, 5], rows[:, 6]))
    table = abi.make_line_table(lines)
    nf = abi.nfield(len(lines))
    st = np.zeros((1, nf, nd))
    st[0, abi.F_T] = T
    st[0, abi.F_VT] = Vt
    st[0, abi.F_VLOS] = Vr
    st[0, abi.F_BMAG] = Bm
    st[0, abi.F_BINC] = Bi
    st[0, abi.F_BAZI] = Ba
    st[0, abi.F_NE] = Ne
    st[0, abi.F_NHI] = nhi
    st[0, abi.F_OPREF] = opref
    st[0, abi.F_GRID] = grid
    bf = d["bfop"]
    bfop = np.zeros(nd)
    bfem = np.zeros(nd)
    l1 = bf[:, 1].min()
    for row in bf[bf[:, 1] == l1]:
        bfop[int(row[2]) - 1] += row[3]
        bfem[int(row[2]) - 1] += row[4]
    thop = Ne * THOMSON
    st[0, abi.F_BGOP] = thop + bfop
    st[0, abi.F_BGEM] = thop * planck(lam_mid, T) + bfem
    for i, (nl, nu, _, _) in enumerate(pops):
        st[0, abi.NF_BASE + 2 * i] = nl
        st[0, abi.NF_BASE + 2 * i + 1] = nu
    ray = abi.SnapiRay(theta=np.pi, lam_mid=lam_mid, tau_grid=1, hminus=1)
    return dict(table=table, ray=ray, state=st, lam_vac=lam_vac, lam_air=d["lam"][:, 0], stokes=d["stokes"],
                stokes_steps=d["stokes_steps"], opem=d["opem"], spop=d["spop"], sdepth=d["sdepth"], lines=lines,
                pops=pops, bf=bf, depth=dep)
