"""Batched LM inversion on the GPU: synthetic observations from known node
values (the oracle's synthesis), inverted from perturbed starts with the
adjoint Jacobian (FP64 and mixed precision) and with central-difference
Jacobians (FP64)."""
import numpy as np
import pytest

import _oracle
from snapi_amd import abi, invert, workloads

pytestmark = pytest.mark.gpu

NODES = [("vt", -2.0), ("vlos", -4.0), ("vlos", 0.0), ("B", -4.0), ("B", 0.0), ("theta", -2.0), ("phi", -2.0)]


@pytest.fixture(scope="module")
def problem():
    t = workloads.template_cfg1()
    rng = np.random.default_rng(11)
    npix = 24
    truth = np.zeros((npix, len(NODES)))
    truth[:, 0] = rng.uniform(0.3e5, 1.8e5, npix)
    truth[:, 1] = rng.normal(0, 1.5e5, npix)
    truth[:, 2] = truth[:, 1] + rng.normal(0, 0.5e5, npix)
    truth[:, 3] = rng.uniform(200, 2500, npix)
    truth[:, 4] = truth[:, 3] * rng.uniform(0.7, 1.3, npix)
    truth[:, 5] = np.arccos(rng.uniform(-0.9, 0.9, npix))
    truth[:, 6] = rng.uniform(0.1, 1.4, npix)
    base = t["state"]
    lam = t["lam_vac"][::2]
    states = np.stack([_oracle.nodes_to_state(NODES, truth[p], base) for p in range(npix)])
    obs, rc = _oracle.synth(t["table"], t["ray"], states, lam)
    start = truth * rng.uniform(0.8, 1.2, truth.shape)
    return dict(t=t, truth=truth, states=states, lam=lam, obs=obs, start=start)


@pytest.mark.parametrize("precision,jacobian", [(abi.SNAPI_FP64, 0), (abi.SNAPI_MIXED, 0), (abi.SNAPI_FP64, 1),
                                               (abi.SNAPI_FP64, 2), (abi.SNAPI_MIXED, 2)])
def test_inversion_recovers_truth(problem, precision, jacobian):
    p = problem
    print("inversion: %d pixels, %d nodes, %d wavelengths, precision %d, jacobian %d" % (
        len(p["truth"]), len(NODES), len(p["lam"]), precision, jacobian), flush=True)
    model = invert.node_model(NODES)
    ws = np.array([1.0, 5.0, 5.0, 3.0])
    nodes, chi2, iters, total = invert.invert(p["t"]["table"], p["t"]["ray"], model, p["states"], p["lam"],
                                              p["obs"], p["start"], wstokes=ws, max_iter=30, nslot=10,
                                              precision=precision, jacobian=jacobian)
    ic = p["obs"][..., 0].max()
    # start chi^2 for reference
    st0 = np.stack([_oracle.nodes_to_state(NODES, p["start"][i], p["t"]["state"]) for i in range(len(p["start"]))])
    s0, _ = _oracle.synth(p["t"]["table"], p["t"]["ray"], st0, p["lam"])
    chi0 = (((s0 - p["obs"]) * ws) ** 2).sum(axis=(1, 2))
    print("  chi2 start median %.3e, end median %.3e; iterations %s; total %d" % (
        np.median(chi0), np.median(chi2), np.bincount(iters), total), flush=True)
    assert total == iters.sum()
    assert np.all(iters >= 1) and np.all(iters <= 30)
    assert np.median(chi2 / chi0) < 1e-3
    # recovered magnetic field and velocities of the well-constrained pixels
    good = chi2 / (ic * ic) < 1e-6
    assert good.mean() > 0.6
    rel_b = np.abs(nodes[good, 3] - p["truth"][good, 3]) / p["truth"][good, 3]
    assert np.median(rel_b) < 0.02
    dv = np.abs(nodes[good, 1] - p["truth"][good, 1])
    assert np.median(dv) < 5e3
