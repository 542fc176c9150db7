"""The oracle's assembly of the NLTE population-response system
(oracle/snapi_oracle_nlteasm.c) against the reference build's own values
(tests/golden/nlterin.npz: the Ca II atom of the cfg0 column, 2 rays x 444
wavelengths, 61 depths, 6 levels).

The oracle keeps the reference's operands and loop order, so its profiles,
profile norms and norm derivatives equal the reference's bit for bit, and the
system (M, beta) equals it to the rounding of a few reordered sums (bar:
1e-15 of each row's largest entry).  CPU only."""
import numpy as np
import pytest

import _nlte_case
import _oracle


@pytest.fixture(scope="module")
def case():
    return _nlte_case.load()


def test_profile_norms_equal_the_reference(case):
    atom, c = case
    for ray in range(c["nray"]):
        norm, normd = _oracle.nlte_profile_norm(atom, c, ray)
        ref = c["norm"][ray]  # [nd][ntr][norm, dnorm/dT, /dNt, /dvt]
        assert np.array_equal(norm, ref[:, :, 0])
        for q in range(3):
            assert np.array_equal(normd[q], ref[:, :, 1 + q])
        # bound-free transitions carry no profile
        bf = [t for t in range(c["ntr"]) if atom.tr_type[t] == 1]
        assert not np.any(norm[:, bf])


def test_line_profiles_equal_the_reference(case):
    atom, c = case
    for ray in range(c["nray"]):
        for s, l in enumerate(range(0, c["nlam"], 40)):
            prof = _oracle.nlte_line_profiles(atom, c, ray, c["lam"][l])
            assert np.array_equal(prof, c["prof"][ray, s]), (ray, l)


def test_system_matches_the_reference(case):
    atom, c = case
    M, beta = _oracle.nlte_response_system(atom, c)
    nd, nm = c["nd"], c["nmap"]
    Mr, br = c["M"], c["beta"]
    err = np.abs(M - Mr) / np.abs(Mr).max(axis=1, keepdims=True)
    print("M: max %.2e of the row's largest, %.4f of the entries equal" % (err.max(), np.mean(M == Mr)))
    assert err.max() < 1e-15
    for q in range(3):
        e = np.abs(beta[q] - br[q]) / np.maximum(np.abs(br[q]).max(axis=1, keepdims=True), 1e-300)
        print("beta[%d]: max %.2e, %.4f equal" % (q, e.max(), np.mean(beta[q] == br[q])))
        assert e.max() < 1e-15
    # structure the reference imposes: the last level's row of each depth is
    # the particle conservation row, and the T and Nt right-hand sides carry
    # the derivative of the atom's total population there
    last = np.arange(nd) * nm + nm - 1
    for l in range(nd):
        row = M[last[l]].reshape(nd, nm)
        assert np.all(row[l] == 1.0) and not np.any(np.delete(row, l, axis=0))
    dep = c["depth"]
    for q in range(2):
        rhs = beta[q][:, last]
        assert np.array_equal(np.diag(rhs), dep[11 + q])
        assert not np.any(rhs - np.diag(np.diag(rhs)))


def test_response_solution_matches_the_reference_responses(case):
    """The oracle's system, solved by the oracle's Crout solve, gives the
    reference's level responses (tests/golden/nlteresp.npz)."""
    atom, c = case
    M, beta = _oracle.nlte_response_system(atom, c)
    nd, n = c["nd"], c["nd"] * c["nmap"]
    X = _oracle.response_solve(M, beta.reshape(3 * nd, n))
    ref = np.load(_oracle.os.path.join(_oracle.REPO, "tests", "golden", "nlteresp.npz"))["resp"]
    want = ref.transpose(0, 2, 1).reshape(3 * nd, n)
    err = np.max(np.abs(X - want) / np.maximum(np.abs(want).max(axis=1, keepdims=True), 1e-300))
    print("level responses: max %.2e of each right-hand side's largest" % err)
    assert err < 1e-10
