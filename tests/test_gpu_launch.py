"""Iteration launches of the batched inversion: on a stream the iterations
after the first are replayed from one captured HIP graph, with the active
count read by the host only every k iterations (k adapting to the
iteration time).  The graph path must give bit for bit the results of the
eager launches (SNAPI_LAUNCH=eager), call after call, with eager work and
synchronisations between the calls; a null stream takes the eager path.
The times of both paths are printed (small batches are launch-bound)."""
import os
import time

import numpy as np
import pytest
import torch

from snapi_amd import abi, api, invert, workloads

pytestmark = pytest.mark.gpu


def _problem(side):
    t = workloads.template_cfg1()
    p = workloads.inversion_problem(t, side, side, seed=21)
    lam = t["lam_vac"]
    obs = api.synth_stokes(t["table"], t["ray"], p["state"], lam, abi.SNAPI_FP64)
    obs = obs + np.random.default_rng(5).normal(0.0, 1e-3 * obs[:, :1, :1], obs.shape)
    return t, p, lam, obs


def _run(t, p, lam, obs, dt, stream, jacobian, max_iter=30):
    dev = "cuda"
    npix = p["state"].shape[0]
    model = invert.node_model(p["nodes"])
    st = torch.from_numpy(p["state"]).to(dev, dt)
    lt = torch.from_numpy(np.ascontiguousarray(lam)).to(dev)
    ot = torch.from_numpy(obs).to(dev, dt)
    mt = torch.from_numpy(p["mask"]).to(dev, dt)
    n0 = torch.from_numpy(p["start"]).to(dev)
    nodes = torch.zeros_like(n0)
    chi2 = torch.zeros(npix, dtype=torch.float64, device=dev)
    its = torch.zeros(npix, dtype=torch.int32, device=dev)
    prec = abi.SNAPI_FP64 if dt == torch.float64 else abi.SNAPI_MIXED
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    total = invert.invert_dev(t["table"], t["ray"], model, st, lt, ot, mt, n0, nodes, chi2, its, wstokes=(1, 5, 5, 3),
                              max_iter=max_iter, precision=prec, jacobian=jacobian, stream=stream)
    if stream is not None:
        stream.synchronize()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3
    return nodes.cpu().numpy(), chi2.cpu().numpy(), its.cpu().numpy(), total, ms


@pytest.mark.parametrize("jacobian", [abi.SNAPI_JAC_TANGENT, abi.SNAPI_JAC_DIFFERENCES])
def test_graph_replay_matches_eager_launches(jacobian):
    t, p, lam, obs = _problem(6)
    s = torch.cuda.Stream()
    dt = torch.float32 if jacobian == abi.SNAPI_JAC_TANGENT else torch.float64
    print("graph vs eager: 36 pixels, jacobian %d, %s" % (jacobian, dt), flush=True)
    os.environ["SNAPI_LAUNCH"] = "eager"
    try:
        ref = _run(t, p, lam, obs, dt, s, jacobian)
    finally:
        del os.environ["SNAPI_LAUNCH"]
    for rep in range(3):
        got = _run(t, p, lam, obs, dt, s, jacobian)
        # eager work between the replays: a synthesis on the same stream
        api.synth_stokes(t["table"], t["ray"], p["state"][:4], lam, abi.SNAPI_FP64)
        for a, b in zip(ref[:3], got[:3]):
            assert np.array_equal(a, b), rep
        assert got[3] == ref[3]
    null = _run(t, p, lam, obs, dt, None, jacobian)
    for a, b in zip(ref[:3], null[:3]):
        assert np.array_equal(a, b)
    print("  %d pixel-iterations; eager %.1f ms, graph %.1f ms, null stream %.1f ms" % (ref[3], ref[4], got[4],
                                                                                       null[4]), flush=True)


def test_small_batch_launch_times():
    """64 and 1,024 pixels, the 7-node tangent fit: eager launches with a
    host read per iteration against the graph replays (times printed; the
    graph path must not be slower by more than its one capture)."""
    s = torch.cuda.Stream()
    for side in (8, 32):
        t, p, lam, obs = _problem(side)
        _run(t, p, lam, obs, torch.float32, s, abi.SNAPI_JAC_TANGENT)  # kernel loading
        os.environ["SNAPI_LAUNCH"] = "eager"
        try:
            e = min(_run(t, p, lam, obs, torch.float32, s, abi.SNAPI_JAC_TANGENT)[4] for _ in range(2))
        finally:
            del os.environ["SNAPI_LAUNCH"]
        g = min(_run(t, p, lam, obs, torch.float32, s, abi.SNAPI_JAC_TANGENT)[4] for _ in range(2))
        print("launch times, %d pixels: eager %.2f ms, graph %.2f ms" % (side * side, e, g), flush=True)
        assert g < 1.5 * e + 20.0
