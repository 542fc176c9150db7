"""The GPU fit path against the reference itself (tests/golden/nodes.npz and
fit.npz, written by the compiled reference): snapi_build_from_nodes against
atmos_pp::build_from_nodes (plpar/atmos_pp.cc:58-248) and the ray projection
(atmos/atmos_rad.cc:15-51); snapi_node_jacobian_eos against
obs_stokes_responses_to_nodes (atmos/atmos_fit.cc:1205-1352); snapi_invert_eos
against the trajectory of stokes_lm_fit (atmos/atmos_fit.cc:18-227), eight
iterations on three pixels (one masked, one starting below the parameter
limits), and six iterations on two pixels with the model regularised
(tests/golden/fitreg.npz: temperature and vLOS Tikhonov, |B| departure from
zero; regularize_hessian, atmos_fit.cc:1371-1480).

Tolerances: the GPU's Jacobian is the exact derivative through the solver,
the reference's a difference once per node direction; in FP64 each node's
column agrees to 3e-4 of its largest entry, and the trajectory's nodes to
2e-3 of their values.  Mixed precision: Stokes to 1e-4 of Ic, Jacobian
columns to 1e-3."""
import numpy as np
import pytest

import _golden
from snapi_amd import abi, invert, workloads

pytestmark = pytest.mark.gpu

NAMES = {1: "T", 2: "vt", 3: "vlos", 4: "B", 5: "theta", 6: "phi"}


def airtovac(lam_cm):
    s = 1e4 / (lam_cm * 1e8)
    n = 1.0 + 0.00008336624212083 + 0.02408926869968 / (130.1065924522 - s * s) \
        + 0.0001599740894897 / (38.92568793293 - s * s)
    return lam_cm * n


@pytest.fixture(scope="module")
def setup():
    g = _golden.load("nodes.npz")
    f = _golden.load("fit.npz")
    t = workloads.template_cfg1()
    chem, chem_he, smap = workloads.eos_tables("cfg1_template", t["ray"].lam_mid)
    nodes = [(NAMES[int(c)], float(l)) for c, l in zip(g["codes"], g["logtau"])]
    st = t["state"].copy()
    st[abi.F_GRID] = g["bfn0"][:, 0]
    lam = np.array([airtovac(x * 1e-8) for x in g["lam_air"]])
    return dict(g=g, f=f, t=t, chem=chem, chem_he=chem_he, smap=smap, nodes=nodes, model=invert.node_model(nodes),
                state=st, lam=lam)


def test_build_from_nodes_matches_reference(setup):
    s = setup
    g = s["g"]
    vals = np.stack([g["values0"], g["values1"]])
    print("build_from_nodes vs reference: 2 node sets, %d depths" % s["state"].shape[1], flush=True)
    st = invert.build_from_nodes(s["model"], s["chem"], s["chem_he"], s["smap"],
                                 np.repeat(s["state"][None], 2, axis=0), vals)
    for k in range(2):
        bfn, proj = g["bfn%d" % k], g["proj%d" % k]
        errs = dict(T=np.abs(st[k, abi.F_T] / bfn[:, 1] - 1).max(), Nt=np.abs(st[k, abi.F_NT] / bfn[:, 2] - 1).max(),
                    Ne=np.abs(st[k, abi.F_NE] / bfn[:, 3] - 1).max(),
                    vlos=np.abs(st[k, abi.F_VLOS] - proj[:, 0]).max() / np.abs(proj[:, 0]).max(),
                    B=np.abs(st[k, abi.F_BMAG] / proj[:, 1] - 1).max(),
                    inc=np.abs(st[k, abi.F_BINC] - proj[:, 2]).max(), azi=np.abs(st[k, abi.F_BAZI] - proj[:, 3]).max())
        print("  set %d: " % k + " ".join("%s %.1e" % kv for kv in errs.items()), flush=True)
        assert max(errs.values()) < 1e-9


@pytest.mark.parametrize("precision", [abi.SNAPI_FP64, abi.SNAPI_MIXED])
def test_node_jacobian_matches_reference(setup, precision):
    s = setup
    g = s["g"]
    vals = np.stack([g["values0"], g["values1"]])
    st = np.repeat(s["state"][None], 2, axis=0)
    print("node Jacobian vs obs_stokes_responses_to_nodes: 2 node sets, %d wavelengths, precision %d"
          % (len(s["lam"]), precision), flush=True)
    S, J = invert.node_jacobian_eos(s["t"]["table"], s["t"]["ray"], s["model"], s["chem"], s["chem_he"], s["smap"],
                                    st, vals, s["lam"], precision=precision)
    bar_s, bar_j = (1e-10, 3e-4) if precision == abi.SNAPI_FP64 else (1e-4, 1e-3)
    for k in range(2):
        ref = g["stokes%d" % k]
        es = np.abs(S[k].astype(np.float64) - ref).max() / ref[:, 0].max()
        jr = np.transpose(g["jac%d" % k], (1, 2, 0)).reshape(-1, len(s["nodes"]))
        ej = np.abs(J[k].astype(np.float64) - jr).max(axis=0) / np.abs(jr).max(axis=0)
        print("  set %d: Stokes %.2e Ic; J per node max %.2e (node %d)" % (k, es, ej.max(), ej.argmax()), flush=True)
        assert es < bar_s
        assert ej.max() < bar_j


def test_inversion_follows_reference_trajectory(setup):
    """snapi_invert_eos run with 1..8 iterations on the three pixels at once:
    after each, the nodes within 2e-3 of the reference's and the same
    iteration counts (the reference runs to its cap here)."""
    s = setup
    f = s["f"]
    K = int(f["kmax"])
    obs = np.stack([f["obs%d" % p] for p in range(3)])
    mask = np.stack([f["mask%d" % p] for p in range(3)])
    start = np.stack([f["start%d" % p] for p in range(3)])
    st = np.repeat(s["state"][None], 3, axis=0)
    print("inversion vs stokes_lm_fit: 3 pixels, 1..%d iterations, FP64" % K, flush=True)
    worst = 0.0
    for k in range(1, K + 1):
        nodes, chi2, iters, total = invert.invert_eos(s["t"]["table"], s["t"]["ray"], s["model"], s["chem"],
                                                      s["chem_he"], s["smap"], st, s["lam"], obs, start, mask=mask,
                                                      wstokes=f["ws"], lm_start=float(f["lm_start"]), max_iter=k,
                                                      nslot=3, precision=abi.SNAPI_FP64)
        for p in range(3):
            ref = f["nodes%d" % p][k - 1]
            rel = np.abs(nodes[p] - ref) / np.abs(ref)
            worst = max(worst, rel.max())
            assert rel.max() < 2e-3, (k, p, rel.max(), rel.argmax())
            assert iters[p] == f["iters%d" % p][k - 1]
    print("  worst node difference %.2e" % worst, flush=True)


@pytest.mark.parametrize("precision", [abi.SNAPI_FP64, abi.SNAPI_MIXED])
def test_tangent_node_jacobian_with_temperature_nodes(setup, precision):
    """The forward-mode sweep with the temperature and density rows against
    the reference's Jacobian (3e-4 per node in FP64, 1e-3 mixed) and, in
    FP64, against the adjoint sweep (both exact derivatives of the same
    discrete model: 1e-9)."""
    s = setup
    g = s["g"]
    vals = np.stack([g["values0"], g["values1"]])
    st = np.repeat(s["state"][None], 2, axis=0)
    print("tangent node Jacobian with temperature nodes: 2 node sets, %d wavelengths, precision %d"
          % (len(s["lam"]), precision), flush=True)
    S, J = invert.node_jacobian_eos(s["t"]["table"], s["t"]["ray"], s["model"], s["chem"], s["chem_he"], s["smap"],
                                    st, vals, s["lam"], precision=precision, jacobian=abi.SNAPI_JAC_TANGENT)
    if precision == abi.SNAPI_FP64:
        Sa, Ja = invert.node_jacobian_eos(s["t"]["table"], s["t"]["ray"], s["model"], s["chem"], s["chem_he"],
                                          s["smap"], st, vals, s["lam"], precision=precision)
        ea = float((np.abs(J - Ja).max(axis=1) / np.abs(Ja).max(axis=1)).max())
        print("  tangent vs adjoint: max %.2e of each node's largest entry" % ea, flush=True)
        assert ea < 1e-9
        assert np.abs(S - Sa).max() / Sa[..., 0].max() < 1e-12
    bar = 3e-4 if precision == abi.SNAPI_FP64 else 1e-3
    for k in range(2):
        jr = np.transpose(g["jac%d" % k], (1, 2, 0)).reshape(-1, len(s["nodes"]))
        ej = np.abs(J[k].astype(np.float64) - jr).max(axis=0) / np.abs(jr).max(axis=0)
        print("  set %d vs reference: J per node max %.2e (node %d)" % (k, ej.max(), ej.argmax()), flush=True)
        assert ej.max() < bar


def test_tangent_inversion_follows_reference_trajectory(setup):
    """The inversion with temperature nodes on the tangent Jacobian: the same
    trajectory as the reference over eight iterations (2e-3)."""
    s = setup
    f = s["f"]
    K = int(f["kmax"])
    obs = np.stack([f["obs%d" % p] for p in range(3)])
    mask = np.stack([f["mask%d" % p] for p in range(3)])
    start = np.stack([f["start%d" % p] for p in range(3)])
    st = np.repeat(s["state"][None], 3, axis=0)
    print("inversion (tangent Jacobian) vs stokes_lm_fit: 3 pixels, %d iterations, FP64" % K, flush=True)
    nodes, chi2, iters, total = invert.invert_eos(s["t"]["table"], s["t"]["ray"], s["model"], s["chem"],
                                                  s["chem_he"], s["smap"], st, s["lam"], obs, start, mask=mask,
                                                  wstokes=f["ws"], lm_start=float(f["lm_start"]), max_iter=K,
                                                  nslot=3, precision=abi.SNAPI_FP64,
                                                  jacobian=abi.SNAPI_JAC_TANGENT)
    for p in range(3):
        rel = np.abs(nodes[p] - f["nodes%d" % p][K - 1]) / np.abs(f["nodes%d" % p][K - 1])
        print("  pixel %d: iterations %d, worst node %.2e" % (p, iters[p], rel.max()), flush=True)
        assert rel.max() < 2e-3 and iters[p] == K


def test_difference_jacobian_inversion_with_temperature_nodes(setup):
    """The FP64 central-difference Jacobian with temperature nodes (every node
    set rebuilt: temperature, hydrostatic equilibrium, equation of state, the
    synthesis path): eight iterations from the fixture's starts reach the
    reference's fit, nodes within 1 % where the lines form and chi^2 within
    20 % of the reference's final value."""
    s = setup
    f = s["f"]
    K = int(f["kmax"])
    obs = np.stack([f["obs%d" % p] for p in range(2)])
    mask = np.stack([f["mask%d" % p] for p in range(2)])
    start = np.stack([f["start%d" % p] for p in range(2)])
    st = np.repeat(s["state"][None], 2, axis=0)
    print("inversion with the FP64 difference Jacobian: 2 pixels, %d iterations" % K, flush=True)
    nodes, chi2, iters, total = invert.invert_eos(s["t"]["table"], s["t"]["ray"], s["model"], s["chem"],
                                                  s["chem_he"], s["smap"], st, s["lam"], obs, start, mask=mask,
                                                  wstokes=f["ws"], lm_start=float(f["lm_start"]), max_iter=K,
                                                  nslot=2, precision=abi.SNAPI_FP64,
                                                  jacobian=abi.SNAPI_JAC_DIFFERENCES)
    built = invert.build_from_nodes(s["model"], s["chem"], s["chem_he"], s["smap"], st, nodes)
    from snapi_amd import api
    Sf = api.synth_stokes(s["t"]["table"], s["t"]["ray"], built, s["lam"], abi.SNAPI_FP64)
    for p in range(2):
        ref = f["nodes%d" % p][K - 1]
        w = mask[p][:, None] * f["ws"][None, :]
        chif = float((((Sf[p] - obs[p]) * w) ** 2).sum())
        rel = np.abs(nodes[p] - ref) / np.abs(ref)
        print("  pixel %d: chi2 %.4e (reference %.4e); nodes vs reference max %.2e, T(-3..0) %.2e" % (
            p, chif, f["chi%d" % p][-1], rel.max(), rel[2:6].max()), flush=True)
        assert chif < 1.2 * f["chi%d" % p][-1]
        assert rel[2:6].max() < 1e-2



def test_regularised_inversion_follows_reference_trajectory(setup):
    """snapi_invert_eos with the regularisation in the node model, 1..6
    iterations on two pixels: the nodes within 2e-3 of the reference's
    regularised fit after each, the same iteration counts, and a trajectory
    that differs from the unregularised one."""
    s = setup
    f = _golden.load("fitreg.npz")
    reg = {NAMES[int(c)]: (int(t), a) for c, t, a in f["reg"]}
    model = invert.node_model(s["nodes"], reg=reg)
    K = int(f["kmax"])
    pix = (0, 1)
    obs = np.stack([f["obs%d" % p] for p in pix])
    mask = np.stack([f["mask%d" % p] for p in pix])
    start = np.stack([f["start%d" % p] for p in pix])
    st = np.repeat(s["state"][None], 2, axis=0)
    print("regularised inversion vs stokes_lm_fit: 2 pixels, 1..%d iterations, FP64" % K, flush=True)
    worst = 0.0
    for k in range(1, K + 1):
        nodes, chi2, iters, total = invert.invert_eos(s["t"]["table"], s["t"]["ray"], model, s["chem"],
                                                      s["chem_he"], s["smap"], st, s["lam"], obs, start, mask=mask,
                                                      wstokes=f["ws"], lm_start=float(f["lm_start"]), max_iter=k,
                                                      nslot=2, precision=abi.SNAPI_FP64)
        for i, p in enumerate(pix):
            ref = f["nodes%d" % p][k - 1]
            rel = np.abs(nodes[i] - ref) / np.abs(ref)
            worst = max(worst, rel.max())
            assert rel.max() < 2e-3, (k, p, rel.max(), rel.argmax())
            assert iters[i] == f["iters%d" % p][k - 1]
    print("  worst node difference %.2e" % worst, flush=True)
    free = s["f"]
    assert np.max(np.abs(f["nodes0"][K - 1] / free["nodes0"][K - 1] - 1.0)) > 1e-2
