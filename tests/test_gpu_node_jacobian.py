"""Node Jacobians dS/dnode on the GPU (tangent sweep, adjoint sweep, central
differences through the synthesis) against each other and against the
oracle's reference-algorithm Jacobian (per-depth responses contracted with
the node map)."""
import numpy as np
import pytest

import _oracle
from snapi_amd import abi, invert, workloads

pytestmark = pytest.mark.gpu

NODES = [("vt", -2.0), ("vlos", -4.0), ("vlos", 0.0), ("B", -4.0), ("B", 0.0), ("theta", -2.0), ("phi", -2.0)]
TAN, ADJ, FD = 2, 0, 1


@pytest.fixture(scope="module")
def pix():
    t = workloads.template_cfg1()
    rng = np.random.default_rng(21)
    npix = 6
    v = np.zeros((npix, len(NODES)))
    v[:, 0] = rng.uniform(0.3e5, 1.8e5, npix)
    v[:, 1] = rng.normal(0, 1.5e5, npix)
    v[:, 2] = v[:, 1] + rng.normal(0, 0.5e5, npix)
    v[:, 3] = rng.uniform(200, 2500, npix)
    v[:, 4] = v[:, 3] * rng.uniform(0.7, 1.3, npix)
    v[:, 5] = np.arccos(rng.uniform(-0.9, 0.9, npix))
    v[:, 6] = rng.uniform(0.1, 1.4, npix)
    st = np.repeat(t["state"][None], npix, axis=0)
    return dict(t=t, v=v, st=st, lam=t["lam_vac"][::6], model=invert.node_model(NODES))


def _rel(a, b):
    """max |a - b| per (pixel, node), relative to max |b| over rows."""
    scale = np.maximum(np.abs(b).max(axis=1, keepdims=True), 1e-300)
    return np.abs(a - b) / scale


def test_tangent_equals_adjoint_fp64(pix):
    p = pix
    print("node Jacobian: %d pixels, %d wavelengths, tangent vs adjoint, FP64" % (len(p["v"]), len(p["lam"])),
          flush=True)
    s1, J1 = invert.node_jacobian(p["t"]["table"], p["t"]["ray"], p["model"], p["st"], p["v"], p["lam"],
                                  abi.SNAPI_FP64, TAN)
    s2, J2 = invert.node_jacobian(p["t"]["table"], p["t"]["ray"], p["model"], p["st"], p["v"], p["lam"],
                                  abi.SNAPI_FP64, ADJ)
    assert np.max(np.abs(s1 - s2)) / np.max(s2[..., 0]) < 1e-13
    e = _rel(J1, J2)
    print("  max rel diff %.2e" % e.max(), flush=True)
    assert e.max() < 1e-9


def test_tangent_against_oracle_reference_algorithm(pix):
    p = pix
    t = p["t"]
    print("node Jacobian: tangent FP64 vs oracle (responses x node map)", flush=True)
    s1, J1 = invert.node_jacobian(t["table"], t["ray"], p["model"], p["st"], p["v"], p["lam"], abi.SNAPI_FP64, TAN)
    for i in range(len(p["v"])):
        so, Jo = _oracle.node_jacobian(t["table"], t["ray"], NODES, p["v"][i], p["st"][i], p["lam"])
        assert np.max(np.abs(s1[i] - so)) / np.max(so[:, 0]) < 1e-10
        e = _rel(J1[i][None], Jo[None])[0]
        print("  pixel %d: max %.2e, p99 %.2e" % (i, e.max(), np.quantile(e, 0.99)), flush=True)
        assert np.quantile(e, 0.99) < 1e-4 and e.max() < 3e-3, i


def test_tangent_against_differences_and_mixed(pix):
    p = pix
    t = p["t"]
    print("node Jacobian: tangent FP64 vs central differences FP64, and mixed tangent", flush=True)
    _, Jt = invert.node_jacobian(t["table"], t["ray"], p["model"], p["st"], p["v"], p["lam"], abi.SNAPI_FP64, TAN)
    _, Jd = invert.node_jacobian(t["table"], t["ray"], p["model"], p["st"], p["v"], p["lam"], abi.SNAPI_FP64, FD)
    e = _rel(Jt, Jd)
    # the inclination column carries the reference's rho_V term (cos instead
    # of -sin of the inclination) and is not the derivative of the synthesis
    for n, (name, _) in enumerate(NODES):
        print("  %-5s max %.2e" % (name, e[:, :, n].max()), flush=True)
        if name != "theta":
            assert e[:, :, n].max() < 2e-3, name
    _, Jm = invert.node_jacobian(t["table"], t["ray"], p["model"], p["st"].astype(np.float32), p["v"], p["lam"],
                                 abi.SNAPI_MIXED, TAN)
    em = _rel(Jm.astype(np.float64), Jt)
    print("  mixed vs FP64 tangent: max %.2e" % em.max(), flush=True)
    assert em.max() < 1e-3
