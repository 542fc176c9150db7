"""The oracle's fit path against the reference itself (fixtures written by the
compiled reference through oracle/ref_driver.cc, tests/golden/nodes.npz and
fit.npz): build_from_nodes (plpar/atmos_pp.cc:58-248), the ray projection of
the field (atmos/atmos_rad.cc:15-51), the synthesis-path spectrum
(forward_evaluate, atmos/atmos_fit.cc:717-737), the node responses and node
Jacobian of obs_stokes_responses_to_nodes (atmos/atmos_fit.cc:1205-1352), and
the trajectory of stokes_lm_fit (atmos/atmos_fit.cc:18-227) over eight
iterations on three pixels, one with a wavelength mask and one whose start
lies below the parameter limits.  CPU only."""
import math

import numpy as np
import pytest

import _golden
import _oracle
from snapi_amd import abi, workloads

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
    st[abi.F_GRID] = g["bfn0"][:, 0]  # build_from_nodes' own tau grid
    lam = np.array([airtovac(x * 1e-8) for x in g["lam_air"]])
    return dict(g=g, f=f, t=t, chem=chem, chem_he=chem_he, smap=smap, nodes=nodes, state=st, lam=lam)


@pytest.mark.parametrize("k", [0, 1])
def test_build_from_nodes_and_projection_match_reference(setup, k):
    s = setup
    g = s["g"]
    vals, bfn, proj = g["values%d" % k], g["bfn%d" % k], g["proj%d" % k]
    fs = _oracle.forward_state_eos(s["nodes"], vals, s["state"], s["chem"], s["chem_he"], s["smap"])
    # bfn columns: tau, T, Nt, Ne, op_ref(5000 A), rho, vt, vz, Bx, By, Bz
    assert np.max(np.abs(fs[abi.F_T] / bfn[:, 1] - 1)) < 1e-13
    assert np.max(np.abs(fs[abi.F_NT] / bfn[:, 2] - 1)) < 1e-12
    assert np.max(np.abs(fs[abi.F_NE] / bfn[:, 3] - 1)) < 1e-12
    assert np.max(np.abs(fs[abi.F_VT] - bfn[:, 6])) < 1e-9
    # projection on the mu = 1 ray: vLOS, |B|, inclination, azimuth [deg]
    assert np.max(np.abs(fs[abi.F_VLOS] - proj[:, 0])) < 1e-9
    assert np.max(np.abs(fs[abi.F_BMAG] / proj[:, 1] - 1)) < 1e-13
    assert np.max(np.abs(fs[abi.F_BINC] - proj[:, 2])) < 1e-11
    assert np.max(np.abs(fs[abi.F_BAZI] - proj[:, 3])) < 1e-11
    # the hydrostatic densities of the march (referent opacity at 5000 A)
    Nt, ne, opref, rho, _ = _oracle.hequilibrium(s["chem_he"], fs[abi.F_T], np.log10(-s["state"][abi.F_GRID]))
    assert np.max(np.abs(opref[0] / bfn[:, 4] - 1)) < 1e-12
    assert np.max(np.abs(rho[0] / bfn[:, 5] - 1)) < 1e-12


@pytest.mark.parametrize("k", [0, 1])
def test_forward_spectrum_matches_reference(setup, k):
    s = setup
    g = s["g"]
    fs = _oracle.forward_state_eos(s["nodes"], g["values%d" % k], s["state"], s["chem"], s["chem_he"], s["smap"])
    S, rc = _oracle.synth(s["t"]["table"], _oracle.synthesis_ray(s["t"]["ray"]), fs[None], s["lam"])
    ref = g["fwd%d" % k]
    assert rc == 0
    assert np.max(np.abs(S[0] - ref)) / ref[:, 0].max() < 1e-12


@pytest.mark.parametrize("k", [0, 1])
def test_node_maps_match_reference(setup, k):
    """dq_p(z)/dnode as the model stores them: temperature rows from the
    hydrostatic rebuilds, the others from the interpolation's differences."""
    s = setup
    g = s["g"]
    maps = _oracle.node_maps_eos(s["nodes"], g["values%d" % k], s["state"], s["chem_he"])
    ref = g["nmap%d" % k]
    for n in range(len(s["nodes"])):
        scale = np.maximum(np.abs(ref[n]).max(axis=1), 1e-300)
        err = np.max(np.abs(maps[n] - ref[n]) / scale[:, None])
        assert err < 1e-6, (n, s["nodes"][n], err)


@pytest.mark.parametrize("k", [0, 1])
def test_node_jacobian_matches_reference(setup, k):
    """The response path's Stokes vector to rounding; each node's column of
    the Jacobian within 2e-4 of its largest entry.  The reference differences
    once per node through the solver along the node's combined perturbation,
    the oracle per depth and parameter before contracting with the node
    responses (measured 1e-4 at most, the inclination of node set 1)."""
    s = setup
    g = s["g"]
    S, J = _oracle.node_jacobian_eos(s["t"]["table"], _oracle.response_ray(s["t"]["ray"]), s["nodes"],
                                     g["values%d" % k], s["state"], s["lam"], s["chem"], s["chem_he"], s["smap"])
    ref = g["stokes%d" % k]
    assert np.max(np.abs(S - ref)) / ref[:, 0].max() < 1e-12
    jr = np.transpose(g["jac%d" % k], (1, 2, 0)).reshape(-1, len(s["nodes"]))
    for n in range(len(s["nodes"])):
        err = np.abs(J[:, n] - jr[:, n]).max() / np.abs(jr[:, n]).max()
        assert err < 2e-4, (n, s["nodes"][n], err)


def test_lm_fit_follows_reference_trajectory(setup):
    """Nodes after each of the reference's eight iterations within 5e-4 of
    each value (the Jacobians differ as above), the same number of
    iterations (the reference accepts until about the sixth and then rejects
    every trial, so it runs to its cap: its stopping rule needs accepted
    steps), and the final synthesis-path chi^2 within 1e-3."""
    s = setup
    f = s["f"]
    K = int(f["kmax"])
    for p in range(3):
        v, chi, it, hist = _oracle.lm_fit_eos(s["t"]["table"], s["t"]["ray"], s["nodes"], f["start%d" % p], s["state"],
                                              s["lam"], f["obs%d" % p], s["chem"], s["chem_he"], s["smap"], f["ws"],
                                              max_iter=K, lm_start=float(f["lm_start"]), mask=f["mask%d" % p],
                                              history=True)
        assert it == f["iters%d" % p][-1] == K
        for j in range(K):
            rel = np.abs(hist[j] - f["nodes%d" % p][j]) / np.abs(f["nodes%d" % p][j])
            assert rel.max() < 5e-4, (p, j, rel.max(), rel.argmax())
        fs = _oracle.forward_state_eos(s["nodes"], v, s["state"], s["chem"], s["chem_he"], s["smap"])
        Sf, _ = _oracle.synth(s["t"]["table"], _oracle.synthesis_ray(s["t"]["ray"]), fs[None], s["lam"])
        w = (f["mask%d" % p][:, None] * f["ws"][None, :])
        chif = float((((Sf[0] - f["obs%d" % p]) * w) ** 2).sum())
        assert abs(chif / f["chi%d" % p][-1] - 1.0) < 1e-3, (p, chif, f["chi%d" % p][-1])


def test_regularised_lm_fit_follows_reference_trajectory(setup):
    """The fit with the model regularised (tests/golden/fitreg.npz:
    temperature and vLOS Tikhonov, |B| departure from zero): the oracle's
    nodes after each of the reference's six iterations within 5e-4, the same
    iteration counts, on both pixels."""
    s = setup
    f = _golden.load("fitreg.npz")
    names = {1: "T", 2: "vt", 3: "vlos", 4: "B", 5: "theta", 6: "phi"}
    reg = {names[int(c)]: (int(t), a) for c, t, a in f["reg"]}
    K = int(f["kmax"])
    for p in (0, 1):
        v, chi, it, hist = _oracle.lm_fit_eos(s["t"]["table"], s["t"]["ray"], s["nodes"], f["start%d" % p], s["state"],
                                              s["lam"], f["obs%d" % p], s["chem"], s["chem_he"], s["smap"], f["ws"],
                                              max_iter=K, lm_start=float(f["lm_start"]), mask=f["mask%d" % p],
                                              history=True, reg=reg)
        assert it == f["iters%d" % p][-1]
        for j in range(it):
            rel = np.abs(hist[j] - f["nodes%d" % p][j]) / np.abs(f["nodes%d" % p][j])
            assert rel.max() < 5e-4, (p, j, rel.max(), rel.argmax())


def test_bracket_and_polish_are_the_reference_rules():
    nodes = [("T", 0.0), ("vt", 0.0), ("vlos", 0.0), ("B", 0.0), ("theta", 0.0), ("phi", 0.0)]
    b = _oracle.bracket(nodes, [3000.0, 50.0, -2e6, 0.5, 3.5, 2.0])
    assert list(b[:4]) == [3400.0, 100.0, -15e5, 1.0] and b[4] == 3.5 and b[5] == 2.0
    assert list(_oracle.bracket(nodes, [13000.0, 2e6, 2e6, 2e4, 0.1, 0.1])[:4]) == [12000.0, 15e5, 15e5, 1e4]
    p = _oracle.polish_angles(nodes, [5000.0, 1e5, 0.0, 100.0, 3.5, 2.0])
    assert abs(p[4] - (2 * math.pi - 3.5)) < 1e-14 and abs(p[5] - (2.0 - math.pi)) < 1e-14
