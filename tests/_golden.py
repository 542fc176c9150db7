"""Reference column dumps under tests/golden/ as library inputs.  Test
infrastructure."""
import os

import numpy as np

from snapi_amd import workloads

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(HERE, name))


def cfg0():
    g = workloads.cfg0_column()
    d = load("cfg0_synth.npz")
    g["state"] = g["state"][None]
    g.update(stokes_steps=d["stokes_steps"], opem=d["opem"], spop=d["spop"], sdepth=d["sdepth"], zc=d["zc"])
    return g


def nlte():
    """The reference's NLTE column (tests/golden/nlte.npz) as ALI inputs:
    atom, grid and per-column arrays in the snapi_ali_* layouts (one column)."""
    from snapi_amd import abi
    d = load("nlte.npz")
    nd, nl, nt, nm, ntp = (int(d[k]) for k in ("nd", "nlam", "ntr", "nmap", "ntp"))
    trans = []
    for row in d["trans"]:
        kind, lo, up = int(row[0]), int(row[1]), int(row[2])
        t = dict(kind=kind, lower=lo, upper=up)
        if kind == abi.TR_LINE:
            t.update(lam0=row[3], aul=row[4], blu=row[5], bul=row[6], energy=row[7])
        else:
            t.update(gratio=row[8], de=row[9])
        trans.append(t)
    atom = abi.make_ali_atom(nm, float(d["mass"]), trans)
    lam = d["lam"]
    rays = d["rays"]
    grid = abi.make_ali_grid(lam[:, 0], lam[:, 1], lam[:, 2], d["sigma"], rays[:, 0], rays[:, 1], 1,
                             float(d["boundary"]), int(d["tau_grid"]))
    dep = d["depth"]  # T vt ne opref grid Na
    cs = np.zeros((abi.A_NF, nd))
    cs[abi.A_T], cs[abi.A_VT], cs[abi.A_NE] = dep[:, 0], dep[:, 1], dep[:, 2]
    cs[abi.A_OPREF], cs[abi.A_GRID], cs[abi.A_NA] = dep[:, 3], dep[:, 4], dep[:, 5]
    return dict(atom=atom, grid=grid, nd=nd, nlam=nl, ntrans=nt, nlevel=nm, nray=ntp, colstate=cs,
                vproj=d["vproj"], gamma=d["gamma"], bgop=d["bgop"], bgem=d["bgem"], coll=d["coll"],
                pop0=d["pop0"], acc=d["acc"], newpop=d["newpop"], iters=d["iters"], final=d["final"],
                ref_final=d["ref_final"])


def chemeq():
    """The reference's chemical equilibrium on 133 depth points (a T/Pg sweep
    and the cfg0 column): species table arrays, per-point inputs, the
    reference's partition functions, fractions, dissociation constants and
    results.  Adds 'table' (abi.SnapiChemTable) and 'stage_ref'
    [npoint][nstage] (stages of all elements in table order)."""
    from snapi_amd import abi
    d = dict(load("chemeq.npz"))
    d["table"] = abi.make_chem_table(d)
    t = d["table"]
    d["stage_ref"] = np.concatenate([d["N"][:, e, :t.element[e].nstage] for e in range(t.nelement)], axis=1)
    return d
