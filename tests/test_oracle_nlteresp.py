"""The oracle's formal solution with intensity responses and its response
system solve against the reference build's own values (tests/golden/
nlteresp.npz: Ca II 5 levels + continuum in the cfg0 column, 12
wavelength/ray problems, the 366 x 366 system with 183 right-hand sides),
plus what the mathematics fixes: a constant source function seen from the
bottom gives I = S and no response to scaling opacity and emissivity
together, and the responses are the derivatives of the formal solution
(central differences).  CPU only."""
import numpy as np
import pytest

import _oracle


@pytest.fixture(scope="module")
def ref():
    return dict(np.load(_oracle.os.path.join(_oracle.REPO, "tests", "golden", "nlteresp.npz")))


def test_formal_responses_match_reference(ref):
    worst = [0.0, 0.0, 0.0]
    for p in range(len(ref["prob"])):
        I, L, Rop, Rem = _oracle.formal_responses(ref["grid"], ref["op"][p], ref["em"][p], ref["prob"][p, 2],
                                                  float(ref["boundary"]))
        worst[0] = max(worst[0], np.max(np.abs(I - ref["I"][p]) / np.abs(ref["I"][p]).max()))
        worst[1] = max(worst[1], np.max(np.abs(Rop - ref["Rop"][p])) / np.abs(ref["Rop"][p]).max())
        worst[2] = max(worst[2], np.max(np.abs(Rem - ref["Rem"][p])) / np.abs(ref["Rem"][p]).max())
    print("formal_with_responses: I %.2e, dI/dop %.2e, dI/dem %.2e of the largest entry" % tuple(worst))
    assert max(worst) < 1e-14


def test_response_solve_matches_reference(ref):
    nd, n = int(ref["nd"]), ref["M"].shape[0]
    B = ref["beta"].reshape(3 * nd, n)
    X = _oracle.response_solve(ref["M"], B)
    want = ref["resp"].transpose(0, 2, 1).reshape(3 * nd, n)
    err = np.max(np.abs(X - want) / np.maximum(np.abs(want).max(axis=1, keepdims=True), 1e-300))
    print("response solve: %d x %d, %d right-hand sides, max err %.2e of each row's largest" % (n, n, 3 * nd, err))
    assert err < 1e-13


def test_constant_source_function_from_below(ref):
    g = ref["grid"]
    nd = len(g)
    op = 10.0 ** np.linspace(-2, 3, nd)
    S = 2.5
    I, L, Rop, Rem = _oracle.formal_responses(g, op, S * op, np.pi - 0.3, -1.0)  # upward ray, I = S at the bottom
    assert np.allclose(I, S, rtol=1e-13)
    scale = (Rop * op[None, :] + Rem * (S * op)[None, :]).sum(axis=1)
    print("constant S: max |I - S| %.2e, response to uniform scaling %.2e" % (np.abs(I - S).max(), np.abs(scale).max()))
    # zero up to the reference's Taylor branch (steps below 1e-2): its
    # e^-dtau, w0 and w1 are truncated at different orders, so the step
    # derivatives cancel only to O(dtau^3) there
    assert np.abs(scale).max() < 1e-8 * S


@pytest.mark.parametrize("p", [1, 7])
def test_responses_are_derivatives(ref, p):
    g, op, em, th = ref["grid"], ref["op"][p].copy(), ref["em"][p].copy(), ref["prob"][p, 2]
    I0, _, Rop, Rem = _oracle.formal_responses(g, op, em, th)
    nd = len(g)
    worst = 0.0
    for j in (3, nd // 2, nd - 4):
        for which, R in (("op", Rop), ("em", Rem)):
            h = 1e-6
            a, b = op.copy(), em.copy()
            (a if which == "op" else b)[j] *= 1 + h
            Ip = _oracle.formal_responses(g, a, b, th)[0]
            a, b = op.copy(), em.copy()
            (a if which == "op" else b)[j] *= 1 - h
            Im = _oracle.formal_responses(g, a, b, th)[0]
            x = (op if which == "op" else em)[j]
            fd = (Ip - Im) / (2 * h * x)
            scale = max(np.abs(R[:, j]).max(), 1e-300)
            worst = max(worst, np.max(np.abs(fd - R[:, j])) / scale)
    print("problem %d: responses vs central differences, max %.2e of the column's largest" % (p, worst))
    assert worst < 1e-5
