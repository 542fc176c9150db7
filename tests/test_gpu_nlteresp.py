"""The two numerical kernels of the NLTE population responses on the GPU
against the reference's own values (tests/golden/nlteresp.npz) and against
the oracle on perturbed inputs: formal_with_responses (intensity, Lambda*,
dI/dop, dI/dem for every pair of depths) and the Crout solve of the
response system.  Both run in double with the reference's operation order
(no fused multiply-adds); the formal solution's exponentials come from the
device library, so its results may differ in the last bit: bar 1e-12 of each
matrix's largest entry.  The solve has no transcendental: bar 1e-12 of each
right-hand side's largest entry (measured: equal)."""
import numpy as np
import pytest

import _oracle
from snapi_amd import nlteresp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ref():
    return dict(np.load(_oracle.os.path.join(_oracle.REPO, "tests", "golden", "nlteresp.npz")))


def test_formal_responses_match_reference(ref):
    npr = len(ref["prob"])
    nd = int(ref["nd"])
    print("formal_with_responses on the GPU: %d problems, nd %d" % (npr, nd), flush=True)
    I, L, Rop, Rem = nlteresp.formal_responses(ref["grid"][None], ref["op"][None], ref["em"][None], ref["prob"][:, 2],
                                               float(ref["boundary"]))
    eI = np.max(np.abs(I[0] - ref["I"]) / np.abs(ref["I"]).max(axis=1, keepdims=True))
    eo = max(np.max(np.abs(Rop[0, p] - ref["Rop"][p])) / np.abs(ref["Rop"][p]).max() for p in range(npr))
    ee = max(np.max(np.abs(Rem[0, p] - ref["Rem"][p])) / np.abs(ref["Rem"][p]).max() for p in range(npr))
    print("  I %.2e  dI/dop %.2e  dI/dem %.2e" % (eI, eo, ee))
    assert max(eI, eo, ee) < 1e-12
    # Lambda* is the local weight w1 of each row; the oracle gives it
    for p in (0, npr - 1):
        _, Lo, _, _ = _oracle.formal_responses(ref["grid"], ref["op"][p], ref["em"][p], ref["prob"][p, 2],
                                               float(ref["boundary"]))
        assert np.allclose(L[0, p], Lo, rtol=1e-12, atol=0)


def test_formal_responses_batch_matches_oracle(ref):
    rng = np.random.default_rng(6)
    ncol, npr, nd = 3, len(ref["prob"]), int(ref["nd"])
    grid = np.repeat(ref["grid"][None], ncol, axis=0)
    f = np.exp(rng.normal(0.0, 0.2, (ncol, npr, nd)))
    op = ref["op"][None] * f
    em = ref["em"][None] * f * np.exp(rng.normal(0.0, 0.1, (ncol, npr, nd)))
    theta = ref["prob"][:, 2]
    print("formal_with_responses batch: %d columns x %d problems" % (ncol, npr), flush=True)
    I, L, Rop, Rem = nlteresp.formal_responses(grid, op, em, theta)
    worst = 0.0
    for c in range(ncol):
        for p in range(npr):
            Io, Lo, Ro, Re = _oracle.formal_responses(grid[c], op[c, p], em[c, p], theta[p])
            worst = max(worst, np.max(np.abs(I[c, p] - Io)) / np.abs(Io).max(),
                        np.max(np.abs(Rop[c, p] - Ro)) / np.abs(Ro).max(),
                        np.max(np.abs(Rem[c, p] - Re)) / np.abs(Re).max())
    print("  max %.2e" % worst)
    assert worst < 1e-12


def test_response_solve_matches_reference(ref):
    nd, n = int(ref["nd"]), ref["M"].shape[0]
    B = ref["beta"].reshape(3 * nd, n)
    print("response solve on the GPU: %d x %d, %d right-hand sides" % (n, n, 3 * nd), flush=True)
    X = nlteresp.response_solve(ref["M"], B)
    want = ref["resp"].transpose(0, 2, 1).reshape(3 * nd, n)
    err = np.max(np.abs(X - want) / np.maximum(np.abs(want).max(axis=1, keepdims=True), 1e-300))
    print("  max err %.2e of each row's largest; equal entries %.4f" % (err, np.mean(X == want)))
    assert err < 1e-12


def test_response_solve_batch_matches_oracle(ref):
    rng = np.random.default_rng(2)
    nd, n = int(ref["nd"]), ref["M"].shape[0]
    M = ref["M"][None] * (1.0 + 1e-3 * rng.standard_normal((3, n, n)))
    B = np.repeat(ref["beta"].reshape(3 * nd, n)[None], 3, axis=0) * (1.0 + 1e-3 * rng.standard_normal((3, 3 * nd, n)))
    print("response solve batch: 3 systems", flush=True)
    X = nlteresp.response_solve(M, B)
    err = 0.0
    for s in range(3):
        Xo = _oracle.response_solve(M[s], B[s])
        err = max(err, np.max(np.abs(X[s] - Xo) / np.maximum(np.abs(Xo).max(axis=1, keepdims=True), 1e-300)))
    print("  max err %.2e" % err)
    assert err < 1e-12
