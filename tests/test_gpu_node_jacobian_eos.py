"""Node Jacobians with temperature nodes on the GPU (snapi_node_jacobian_eos:
the 14-node model of the cfg2 workload, 7 of them temperature nodes) against
the oracle's reference-algorithm Jacobian: build_from_nodes with the
temperature interpolation and hydrostatic equilibrium, the response path's
state and thermodynamic derivatives, per-depth responses contracted with the
node responses (temperature nodes: dT/dnode and dNt/dnode from rebuilds at
+- 2.5 K).

Tolerance: the GPU sweeps the exact derivative through the solver while the
oracle differences it (10x smaller steps than the reference's), as for the
other node Jacobians: 1e-4 of each (pixel, node) column's largest value on
99 % of the rows, 3e-3 everywhere.  The Stokes vectors (both at the response
path's state) agree to 1e-10 of the continuum."""
import numpy as np
import pytest

import _oracle
from snapi_amd import abi, invert, workloads

pytestmark = pytest.mark.gpu

NODES = [("T", -5.0), ("T", -4.0), ("T", -3.0), ("T", -2.0), ("T", -1.0), ("T", 0.0), ("T", 1.0),
         ("vt", -2.0), ("vlos", -4.0), ("vlos", 0.0), ("B", -4.0), ("B", 0.0), ("theta", -2.0), ("phi", -2.0)]


@pytest.fixture(scope="module")
def pix():
    t = workloads.template_cfg1()
    chem, chem_he, smap = workloads.eos_tables("cfg1_template", t["ray"].lam_mid)
    rng = np.random.default_rng(33)
    npix = 3
    lt = np.log10(-t["state"][abi.F_GRID])
    Tn = np.interp([n[1] for n in NODES[:7]], lt, t["state"][abi.F_T])
    v = np.zeros((npix, len(NODES)))
    v[:, :7] = Tn[None] + rng.normal(0.0, 100.0, (npix, 7))
    v[:, 7] = rng.uniform(0.3e5, 1.8e5, npix)
    v[:, 8] = rng.normal(0, 1.5e5, npix)
    v[:, 9] = v[:, 8] + rng.normal(0, 0.5e5, npix)
    v[:, 10] = rng.uniform(200, 2500, npix)
    v[:, 11] = v[:, 10] * rng.uniform(0.7, 1.3, npix)
    v[:, 12] = np.arccos(rng.uniform(-0.9, 0.9, npix))
    v[:, 13] = rng.uniform(0.1, 1.4, npix)
    st = np.repeat(t["state"][None], npix, axis=0)
    return dict(t=t, v=v, st=st, lam=t["lam_vac"][::6], model=invert.node_model(NODES), chem=chem,
                chem_he=chem_he, smap=smap)


def test_temperature_node_jacobian_against_oracle(pix):
    p = pix
    t = p["t"]
    print("node Jacobian with temperature nodes: %d pixels, %d wavelengths, 14 nodes, adjoint FP64"
          % (len(p["v"]), len(p["lam"])), flush=True)
    s1, J1 = invert.node_jacobian_eos(t["table"], t["ray"], p["model"], p["chem"], p["chem_he"], p["smap"], p["st"],
                                      p["v"], p["lam"])
    for i in range(len(p["v"])):
        so, Jo = _oracle.node_jacobian_eos(t["table"], t["ray"], NODES, p["v"][i], p["st"][i], p["lam"], p["chem"],
                                           p["chem_he"], p["smap"], fd_scale=0.1)
        es = np.max(np.abs(s1[i] - so)) / np.max(so[:, 0])
        scale = np.maximum(np.abs(Jo).max(axis=0, keepdims=True), 1e-300)
        e = np.abs(J1[i] - Jo) / scale
        print("  pixel %d: Stokes %.2e; J max %.2e, p99 %.2e; T nodes max %.2e" %
              (i, es, e.max(), np.quantile(e, 0.99), e[:, :7].max()), flush=True)
        assert es < 1e-10
        assert np.quantile(e, 0.99) < 1e-4 and e.max() < 3e-3, i
    assert np.all(np.abs(J1[:, :, :7]).max(axis=1) > 0)


@pytest.mark.parametrize("jacobian,analytic", [(abi.SNAPI_JAC_TANGENT, True), (abi.SNAPI_JAC_TANGENT, False),
                                               (abi.SNAPI_JAC_ADJOINT, False)])
def test_mixed_temperature_node_jacobian_against_fp64(pix, jacobian, analytic):
    """The mixed-precision Jacobian of the 14-node model (float opacity,
    formal solution and sweep) against the FP64 one at the same nodes: the
    inversion's bar is 1e-3 of each (pixel, node) column's largest value, and
    the Stokes vector within 1e-4 of the continuum."""
    p = pix
    t = p["t"]
    print("14-node Jacobian, mixed vs FP64: jacobian %d, analytic maps %s" % (jacobian, analytic), flush=True)
    args = (t["table"], t["ray"], p["model"], p["chem"], p["chem_he"], p["smap"])
    s64, J64 = invert.node_jacobian_eos(*args, p["st"], p["v"], p["lam"], precision=abi.SNAPI_FP64,
                                        analytic_maps=analytic, jacobian=jacobian)
    s32, J32 = invert.node_jacobian_eos(*args, p["st"].astype(np.float32), p["v"], p["lam"],
                                        precision=abi.SNAPI_MIXED, analytic_maps=analytic, jacobian=jacobian)
    es = np.max(np.abs(s32.astype(np.float64) - s64)) / np.max(s64[..., 0])
    scale = np.maximum(np.abs(J64).max(axis=1, keepdims=True), 1e-300)
    e = np.abs(J32.astype(np.float64) - J64) / scale
    print("  Stokes %.2e Ic; J max %.2e of each column's largest value (T nodes %.2e)"
          % (es, e.max(), e[:, :, :7].max()), flush=True)
    assert es < 1e-4
    assert e.max() < 1e-3


def test_tangent_row_sets_node_order(pix):
    """The tangent Jacobian does not depend on the order of the nodes.  In
    the natural order the two passes of up to 8 nodes launch the kernels
    carrying the (vmicro, T, Nt) rows and the (vLOS, B, inclination,
    azimuth) rows; reordered so that the first pass mixes temperature nodes
    with a field node and the second holds vmicro with the field nodes, they
    launch the kernels carrying all seven rows and the five kinematic rows.
    The columns agree to rounding in FP64 and in mixed precision."""
    p = pix
    t = p["t"]
    order = [0, 10, 1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 12, 13]  # each parameter's nodes stay ascending
    model2 = invert.node_model([NODES[i] for i in order])
    v2 = np.ascontiguousarray(p["v"][:, order])
    for prec, st, tol in ((abi.SNAPI_FP64, p["st"], 1e-12), (abi.SNAPI_MIXED, p["st"].astype(np.float32), 1e-5)):
        print("node order, precision %d" % prec, flush=True)
        s1, J1 = invert.node_jacobian_eos(t["table"], t["ray"], p["model"], p["chem"], p["chem_he"], p["smap"], st,
                                          p["v"], p["lam"], precision=prec, analytic_maps=True,
                                          jacobian=abi.SNAPI_JAC_TANGENT)
        s2, J2 = invert.node_jacobian_eos(t["table"], t["ray"], model2, p["chem"], p["chem_he"], p["smap"], st, v2,
                                          p["lam"], precision=prec, analytic_maps=True,
                                          jacobian=abi.SNAPI_JAC_TANGENT)
        J1o = J1[..., order].astype(np.float64)
        scale = np.maximum(np.abs(J1o).max(axis=1, keepdims=True), 1e-300)
        e = np.abs(J2.astype(np.float64) - J1o) / scale
        es = np.max(np.abs(s2.astype(np.float64) - s1)) / np.max(s1[..., 0])
        print("  Stokes %.2e Ic; J max %.2e of each column's largest value" % (es, e.max()), flush=True)
        assert es < tol
        assert e.max() < tol
