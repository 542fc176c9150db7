"""The NLTE population-response system on the GPU (snapi_nlte_response_system)
against the reference's own system (tests/golden/nlterin.npz, the Ca II atom
of the cfg0 column: 2 rays x 444 wavelengths, 61 depths, 6 levels) and
against the oracle on perturbed columns.

The GPU sums the 888 (ray, wavelength) contributions per matrix entry in
problem order as the reference does, but groups the terms of one problem
differently (level vectors and per-line sums, fused multiply-adds in the
assembly) and takes its exponentials from the device library.  How far such
regrouping alone moves the result is set by the reference's own central
differences (the Nt derivative of the profile norm differences two Voigt
values a 1e-3 damping step apart): the oracle compiled with fused
multiply-adds moves from itself by 2e-13 (M), 5e-12 (beta T), 7e-10
(beta Nt) and 2e-11 (beta vt) of each row's largest entry, and, through
the solve (cond(M) = 5.9e11), its level responses by 1.3e-7 of each
right-hand side's largest.  Bars, about ten times that: M 1e-11, beta T
and vt 1e-10, beta Nt 1e-8, level responses 1e-6.  (The FP64 tolerance of
the path is 1e-5 relative.)"""
import os

import numpy as np
import pytest
import torch

import _nlte_case
import _oracle
from snapi_amd import nlteresp

pytestmark = pytest.mark.gpu

BAR_M = 1e-11
BAR_BETA = (1e-10, 1e-8, 1e-10)
BAR_RESP = 1e-6


@pytest.fixture(scope="module")
def case():
    return _nlte_case.load()


def _inputs(c, ncol=1, opem=None, pops=None):
    rep = lambda a: np.repeat(a[None], ncol, axis=0)  # noqa: E731
    return dict(sigma=c["sigma"], depth=rep(c["depth"]), pops=rep(c["pops"]) if pops is None else pops,
                damp=rep(c["damp"]), coll=rep(c["coll"]), vr=rep(c["vr"]), opem=rep(c["opem"]) if opem is None else opem)


def _run(atom, c, inp, solve=False):
    return nlteresp.response_system(atom, c["theta"], c["wray"], c["lam"], c["lamw"], inp["sigma"], inp["depth"],
                                    inp["pops"], inp["damp"], inp["coll"], inp["vr"], inp["opem"],
                                    tau_grid=c["tau_grid"], boundary=c["boundary"], solve=solve)


def _rel_rows(a, b):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b).max(axis=-1, keepdims=True), 1e-300))


def test_system_matches_the_reference(case):
    atom, c = case
    print("NLTE response system on the GPU: Ca II, 61 depths, 2 x 444 problems", flush=True)
    M, beta, resp = _run(atom, c, _inputs(c), solve=True)
    eM = _rel_rows(M[0], c["M"])
    eb = [_rel_rows(beta[0, q], c["beta"][q]) for q in range(3)]
    print("  M %.2e  beta T %.2e Nt %.2e vt %.2e" % (eM, *eb), flush=True)
    assert eM < BAR_M and all(e < b for e, b in zip(eb, BAR_BETA))
    ref = np.load(os.path.join(_oracle.REPO, "tests", "golden", "nlteresp.npz"))["resp"]  # [3][n][nd]
    er = max(_rel_rows(resp[0, q].T, ref[q].T) for q in range(3))
    print("  level responses %.2e" % er, flush=True)
    assert er < BAR_RESP


def test_batch_matches_the_oracle(case):
    """Three columns with perturbed opacities, emissivities and populations:
    each column's system equals the oracle's on the same inputs."""
    atom, c = case
    rng = np.random.default_rng(11)
    ncol = 3
    f = np.exp(rng.normal(0.0, 0.1, (ncol,) + c["opem"].shape[:2] + (1, c["nd"])))
    opem = c["opem"][None] * f
    opem[:, :, :, 1] *= np.exp(rng.normal(0.0, 0.05, (ncol,) + c["opem"].shape[:2] + (c["nd"],)))
    pops = c["pops"][None] * np.exp(rng.normal(0.0, 0.05, (ncol,) + c["pops"].shape))
    inp = _inputs(c, ncol, opem=opem, pops=pops)
    print("NLTE response system batch: %d perturbed columns" % ncol, flush=True)
    M, beta = _run(atom, c, inp)
    eM, eb = 0.0, [0.0, 0.0, 0.0]
    for k in range(ncol):
        ck = dict(c, pops=pops[k])
        Mo, bo = _oracle.nlte_response_system(atom, ck, opem=opem[k])
        eM = max(eM, _rel_rows(M[k], Mo))
        eb = [max(eb[q], _rel_rows(beta[k, q], bo[q])) for q in range(3)]
    print("  M %.2e  beta T %.2e Nt %.2e vt %.2e" % (eM, *eb), flush=True)
    assert eM < BAR_M and all(e < b for e, b in zip(eb, BAR_BETA))


def test_device_entry_and_column_passes(case, monkeypatch):
    """The device entry point on torch tensors equals the host entry point bit
    for bit, also when the columns go through the GPU in several passes."""
    atom, c = case
    ncol = 3
    rng = np.random.default_rng(3)
    opem = c["opem"][None] * np.exp(rng.normal(0.0, 0.05, (ncol,) + c["opem"].shape[:2] + (1, c["nd"])))
    inp = _inputs(c, ncol, opem=opem)
    M, beta, resp = _run(atom, c, inp, solve=True)
    dev = "cuda"
    tens = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in inp.items()}
    n = c["nd"] * c["nmap"]
    s = torch.cuda.Stream()
    for chunk in (None, "2"):
        if chunk:
            monkeypatch.setenv("SNAPI_NLTE_COLUMNS_PER_PASS", chunk)
        Mt = torch.zeros((ncol, n, n), dtype=torch.float64, device=dev)
        bt = torch.zeros((ncol, 3, c["nd"], n), dtype=torch.float64, device=dev)
        rt = torch.zeros((ncol, 3, n, c["nd"]), dtype=torch.float64, device=dev)
        print("device entry point, columns per pass %s" % chunk, flush=True)
        nlteresp.response_system_dev(atom, c["theta"], c["wray"], torch.from_numpy(c["lam"]).to(dev),
                                     torch.from_numpy(c["lamw"]).to(dev), tens, Mt, bt, rt,
                                     tau_grid=c["tau_grid"], boundary=c["boundary"], stream=s)
        s.synchronize()
        assert np.array_equal(Mt.cpu().numpy(), M)
        assert np.array_equal(bt.cpu().numpy(), beta)
        assert np.array_equal(rt.cpu().numpy(), resp)
