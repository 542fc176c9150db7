"""The reference's NLTE response case (tests/golden/nlterin.npz, the Ca II
atom of the cfg0 column) in the layouts of snapi_nlte_response_system."""
import os

import numpy as np

from snapi_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))


def load():
    d = dict(np.load(os.path.join(HERE, "golden", "nlterin.npz")))
    nd, nray, nlam, nmap, ntr, Z, tau_grid, damping = (int(x) for x in d["head"])
    mass, saha, dT, dN, boundary = d["consts"]
    lev, tr = d["levels"], d["trans"]
    stage = lev[:, 1].astype(int)
    stage_ip = np.zeros(Z + 1)
    for i in range(nmap):
        stage_ip[stage[i]] = lev[i, 5]
    # transitions: tr, kind (0 line, 1 bf), z, lower, upper (within stage), map of the upper level, A, Blu, Bul, edge
    first = {z: int(np.argmax(stage == z)) for z in set(stage)}
    lower = np.array([first[int(t[2])] + int(t[3]) for t in tr])
    upper = np.array([int(t[5]) for t in tr])
    kind = tr[:, 1].astype(int)
    damp_rad = d["damp"][0, :, 0]
    atom = abi.make_nlte_atom(stage, lev[:, 3], lev[:, 4], stage_ip, kind, lower, upper, tr[:, 6], tr[:, 7],
                              tr[:, 8], damp_rad, tr[:, 9], mass, damping)
    dep = d["depth"]  # [nd][14]: the 13 fields, then the hydrogen population
    case = dict(nd=nd, nray=nray, nlam=nlam, nmap=nmap, ntr=ntr, tau_grid=tau_grid, boundary=float(boundary),
                theta=d["rays"][:, 0].copy(), wray=d["rays"][:, 1].copy(), lam=d["lam"][:, 0].copy(),
                lamw=d["lam"][:, 1].copy(), sigma=d["sigma"].copy(),
                depth=np.ascontiguousarray(dep[:, :abi_nfield()].T),
                pops=np.ascontiguousarray(d["pops"].T),
                damp=np.ascontiguousarray(d["damp"][1:]),
                coll=np.ascontiguousarray(d["coll"].transpose(3, 1, 2, 0)),
                vr=d["vr"].copy(),
                opem=np.ascontiguousarray(d["pp"].transpose(0, 1, 3, 2)),
                M=d["M"], beta=d["beta"], norm=d["norm"], prof=d["prof"], consts=(saha, dT, dN))
    return atom, case


def abi_nfield():
    return len(abi.NLTE_FIELDS)
