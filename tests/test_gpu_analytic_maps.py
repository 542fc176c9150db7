"""Analytic node responses on the GPU (SNAPI_JAC_ANALYTIC_MAPS): the exact
derivatives of the node interpolation and field projection (dual numbers
through the same code that builds the state) and, for temperature nodes, the
hydrostatic tangent (snapi_hequilibrium_tangent) in place of the
reference's central differences and its two hydrostatic rebuilds per
temperature node (atmos_fit.cc:1225-1305).

Bars: the hydrostatic tangent equals the oracle's (same equations, same
difference steps for the density and opacity derivatives) to 1e-8 of each
direction's largest value; the analytic node Jacobian follows the oracle's
analytic Jacobian to the node-Jacobian bar of the reference-way test (1e-4
of each column's largest value on 99 % of the rows, 3e-3 everywhere) and the
GPU's reference-way Jacobian to 1e-3 (the Jacobian bar; the two node
responses differ by the differences' truncation, 1e-5 in the oracle); the
inversion with analytic maps meets the convergence bar of the reference-way
inversion on the same twelve pixels."""
import numpy as np
import pytest

import _golden
import _oracle
from snapi_amd import abi, api, chemeq, invert, workloads
from test_gpu_invert_eos import NODES, problem  # noqa: F401  (module fixture)

pytestmark = pytest.mark.gpu


def test_hydrostatic_tangent_matches_oracle():
    d = _golden.chemeq()
    tb = abi.make_chem_table({k: v for k, v in d.items() if k != "table"}, referent=1)
    rec = d["he"]
    lt = rec[:, 1]
    T = np.stack([rec[:, 0], rec[:, 0] * 1.04, rec[:, 0] * 0.97])
    LT = np.stack([lt] * 3)
    N = chemeq.hequilibrium(tb, T, LT)["Nt"]
    dT = np.stack([np.stack([np.exp(-((lt + c) / 0.8) ** 2) for c in (5.0, 2.0, 0.0)] + [np.ones_like(lt)])] * 3)
    print("hydrostatic tangent: 3 columns x 4 directions x %d depths" % len(lt), flush=True)
    dN = chemeq.hequilibrium_tangent(tb, T, LT, N, dT)
    for c in range(3):
        ref = _oracle.he_tangent(tb, T[c], lt, N[c], dT[c])
        e = np.abs(dN[c] - ref).max(axis=1) / np.abs(ref).max(axis=1)
        print("  column %d: max rel %.2e" % (c, e.max()), flush=True)
        assert e.max() < 1e-8
    assert np.all(dN[:, 3] < 0)  # uniform heating lowers the density


def test_analytic_temperature_node_jacobian(problem):  # noqa: F811
    p = problem
    t = p["t"]
    model = invert.node_model(NODES)
    sel = [0, 3, 7]
    v = p["truth"][sel]
    st = np.repeat(t["state"][None], len(sel), axis=0)
    lam = p["lam"][::3]
    print("analytic node Jacobian with temperature nodes: %d pixels, %d wavelengths" % (len(sel), len(lam)),
          flush=True)
    s1, J1 = invert.node_jacobian_eos(t["table"], t["ray"], model, p["chem"], p["chem_he"], p["smap"], st, v, lam,
                                      analytic_maps=True)
    s0, J0 = invert.node_jacobian_eos(t["table"], t["ray"], model, p["chem"], p["chem_he"], p["smap"], st, v, lam)
    assert np.array_equal(s0, s1)
    for i in range(len(sel)):
        so, Jo = _oracle.node_jacobian_eos(t["table"], t["ray"], NODES, v[i], t["state"], lam, p["chem"],
                                           p["chem_he"], p["smap"], fd_scale=0.1, analytic=True)
        scale = np.maximum(np.abs(Jo).max(axis=0, keepdims=True), 1e-300)
        e = np.abs(J1[i] - Jo) / scale
        ef = np.abs(J1[i] - J0[i]) / np.maximum(np.abs(J0[i]).max(axis=0, keepdims=True), 1e-300)
        print("  pixel %d: vs oracle analytic max %.2e p99 %.2e; vs reference-way maps %.2e" % (
            sel[i], e.max(), np.quantile(e, 0.99), ef.max()), flush=True)
        assert np.quantile(e, 0.99) < 1e-4 and e.max() < 3e-3
        assert ef.max() < 1e-3


def test_analytic_maps_kinematic_magnetic_jacobian():
    """snapi_node_jacobian (no temperature nodes), tangent sweep, with the
    analytic and the reference-step maps."""
    t = workloads.template_cfg1()
    nodes = [n for n in NODES if n[0] != "T"]
    model = invert.node_model(nodes)
    v = np.array([[0.9e5, -1.1e5, 0.4e5, 1800.0, 700.0, 2.1, 0.6], [0.5e5, 0.8e5, 1.2e5, 300.0, 900.0, 0.7, 1.2]])
    st = np.repeat(t["state"][None], 2, axis=0)
    lam = t["lam_vac"][::4]
    print("analytic maps, kinematic/magnetic node Jacobian: 2 pixels, %d wavelengths" % len(lam), flush=True)
    s0, J0 = invert.node_jacobian(t["table"], t["ray"], model, st, v, lam, jacobian=abi.SNAPI_JAC_TANGENT)
    s1, J1 = invert.node_jacobian(t["table"], t["ray"], model, st, v, lam,
                                  jacobian=abi.SNAPI_JAC_TANGENT | abi.SNAPI_JAC_ANALYTIC_MAPS)
    assert np.array_equal(s0, s1)
    e = np.abs(J1 - J0) / np.maximum(np.abs(J0).max(axis=1, keepdims=True), 1e-300)
    print("  analytic vs reference-step maps: max %.2e" % e.max(), flush=True)
    assert e.max() < 1e-4


@pytest.mark.parametrize("precision", [abi.SNAPI_FP64, abi.SNAPI_MIXED])
def test_inversion_with_analytic_maps(problem, precision):  # noqa: F811
    p = problem
    t = p["t"]
    print("inversion with analytic maps: %d pixels, precision %d" % (len(p["truth"]), precision), flush=True)
    model = invert.node_model(NODES)
    ws = np.array([1.0, 5.0, 5.0, 3.0])
    npix = len(p["truth"])
    st0 = np.stack([_oracle.forward_state_eos(NODES, p["start"][i], t["state"], p["chem"], p["chem_he"], p["smap"])
                    for i in range(npix)])
    s0 = api.synth_stokes(t["table"], t["ray"], st0, p["lam"], abi.SNAPI_FP64)
    chi0 = (((s0 - p["obs"]) * ws) ** 2).sum(axis=(1, 2))
    nodes, chi2, iters, total = invert.invert_eos(t["table"], t["ray"], model, p["chem"], p["chem_he"], p["smap"],
                                                  p["states"], p["lam"], p["obs"], p["start"], wstokes=ws,
                                                  max_iter=30, nslot=5, precision=precision, analytic_maps=True)
    st1 = invert.build_from_nodes(model, p["chem"], p["chem_he"], p["smap"],
                                  np.repeat(t["state"][None], npix, axis=0), nodes)
    s1 = api.synth_stokes(t["table"], t["ray"], st1, p["lam"], abi.SNAPI_FP64)
    chi1 = (((s1 - p["obs"]) * ws) ** 2).sum(axis=(1, 2))
    print("  chi2 start median %.3e, end median %.3e; iterations %s" % (np.median(chi0), np.median(chi1),
                                                                        np.bincount(iters)), flush=True)
    assert total == iters.sum()
    assert np.median(chi1 / chi0) < 0.1 and np.mean(chi1 < chi0) >= 0.75
