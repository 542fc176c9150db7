"""Chemical equilibrium and LTE populations on the GPU (snapi_chemeq) against
the reference's own results (tests/golden/chemeq.npz, 133 depth points) and
against the oracle on perturbed columns and on a wide temperature/pressure
sweep; the device-array entry point against the host one; the pure-hydrogen
Saha limit; a table without molecules or levels.  The continuum opacity at
the referent wavelength is checked with the populations.

Tolerance: everything is double precision; the GPU sums the same terms in a
different association and its exp/pow differ from glibc's in the last bit,
so results agree with the oracle to 1e-9 relative wherever both take the same
number of Newton steps (stage and level populations relative to 1e-12 of the
element's total where they are smaller).  Step counts must agree on at least
99 % of the points; where they differ the results differ by at most the last
step's size, bounded by the 1e-4 convergence tolerance."""
import numpy as np
import pytest

import _golden
import _oracle
from snapi_amd import abi, chemeq, workloads
from test_oracle_chemeq import hydrogen_table, saha_ne

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ref():
    return _golden.chemeq()


def rel(a, b, floor):
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def compare(tb, got, want_ne, want_st, want_mo, want_lv, nel_total, tol, mask=None):
    m = np.ones(want_ne.shape, bool) if mask is None else mask
    e_ne = rel(got["ne"], want_ne, 1e-300)[m].max()
    fl = 1e-12 * nel_total
    e_st = rel(got["stage"], want_st, fl[:, None, :])[np.broadcast_to(m[:, None, :], want_st.shape)].max()
    e_mo = 0.0
    if tb.nmolecule and want_mo is not None:
        e_mo = rel(got["molecule"], want_mo, 1e-300)[np.broadcast_to(m[:, None, :], want_mo.shape)].max()
    e_lv = 0.0
    if tb.nlevel and want_lv is not None:
        e_lv = rel(got["level"], want_lv, fl[:, None, :])[np.broadcast_to(m[:, None, :], want_lv.shape)].max()
    print("  ne %.2e  stages %.2e  molecules %.2e  levels %.2e" % (e_ne, e_st, e_mo, e_lv), flush=True)
    assert max(e_ne, e_st, e_mo, e_lv) < tol


def test_chemeq_matches_reference(ref):
    tb = ref["table"]
    n = len(ref["T"])
    print("chemeq vs reference: %d points, %d elements, %d molecules, %d levels" %
          (n, tb.nelement, tb.nmolecule, tb.nlevel), flush=True)
    got = chemeq.chemeq(tb, ref["T"][None], ref["Nt"][None])
    _, _, _, _, it = _oracle.chemeq(tb, ref["T"][None], ref["Nt"][None])
    print("  steps GPU %s" % np.bincount(got["iterations"][0]), flush=True)
    assert np.array_equal(got["iterations"], it)
    nel = ref["Nel"].sum(axis=1)[None]
    compare(tb, got, ref["ne"][None], ref["stage_ref"].T[None], None, ref["n"].T[None], nel, 1e-9)
    e_hm = np.max(rel(got["molecule"][0, 0], ref["hmin"], 1e-300))
    e_kr = np.max(rel(got["opref"][0], ref["opref"], 1e-300))
    print("  H- %.2e  referent opacity %.2e" % (e_hm, e_kr))
    assert e_hm < 1e-9 and e_kr < 1e-9


def oracle_all(tb, T, Nt):
    ne, st, mo, lv, it, kr = _oracle.chemeq(tb, T, Nt, with_opref=True)
    return dict(ne=ne, stage=st, molecule=mo, level=lv, iterations=it, opref=kr)


def check_against_oracle(tb, T, Nt, label):
    print("%s: %d columns x %d depths" % (label, T.shape[0], T.shape[1]), flush=True)
    got = chemeq.chemeq(tb, T, Nt)
    want = oracle_all(tb, T, Nt)
    same = got["iterations"] == want["iterations"]
    print("  steps equal on %.4f of the points; GPU steps %s" % (same.mean(), np.bincount(
        np.maximum(got["iterations"].ravel(), 0))), flush=True)
    assert same.mean() >= 0.99
    assert (got["iterations"] >= 1).all()
    nel = want["stage"].sum(axis=1)
    compare(tb, got, want["ne"], want["stage"], want["molecule"], want["level"], nel, 1e-9, same)
    e_kr = rel(got["opref"], want["opref"], 1e-300)
    print("  referent opacity %.2e" % e_kr[same].max(), flush=True)
    assert e_kr[same].max() < 1e-9
    if not same.all():
        compare(tb, got, want["ne"], want["stage"], want["molecule"], want["level"], nel, 1e-4, ~same)
        assert e_kr[~same].max() < 1e-4
    return got


def test_chemeq_perturbed_columns_match_oracle():
    p = workloads.chemeq_problem(16, 8, seed=11)
    check_against_oracle(p["table"], p["T"], p["Nt"], "chemeq perturbed cfg0 columns")


def test_chemeq_temperature_pressure_sweep_matches_oracle(ref):
    rng = np.random.default_rng(4)
    T = 10.0 ** rng.uniform(np.log10(3000.0), np.log10(30000.0), (4, 300))
    pg = 10.0 ** rng.uniform(-1.0, 6.0, (4, 300))
    Nt = pg / (1.3806503E-16 * T)
    check_against_oracle(ref["table"], T, Nt, "chemeq T/Pg sweep")


def test_chemeq_dev_equals_host(ref):
    import torch
    tb = ref["table"]
    p = workloads.chemeq_problem(8, 4, seed=2)
    host = chemeq.chemeq(tb, p["T"], p["Nt"])
    ncol, nd = p["T"].shape
    dev = torch.device("cuda")
    T_t = torch.from_numpy(p["T"]).to(dev)
    Nt_t = torch.from_numpy(p["Nt"]).to(dev)
    ne_t = torch.empty((ncol, nd), dtype=torch.float64, device=dev)
    st_t = torch.empty((ncol, abi.chem_stage_count(tb), nd), dtype=torch.float64, device=dev)
    mo_t = torch.empty((ncol, tb.nmolecule, nd), dtype=torch.float64, device=dev)
    lv_t = torch.empty((ncol, tb.nlevel, nd), dtype=torch.float64, device=dev)
    it_t = torch.empty((ncol, nd), dtype=torch.int32, device=dev)
    kr_t = torch.empty((ncol, nd), dtype=torch.float64, device=dev)
    print("chemeq_dev vs chemeq: %d x %d" % (ncol, nd), flush=True)
    s = torch.cuda.Stream()
    chemeq.chemeq_dev(tb, T_t, Nt_t, ne_t, st_t, mo_t, lv_t, it_t, stream=s, opref_t=kr_t)
    s.synchronize()
    assert np.array_equal(kr_t.cpu().numpy(), host["opref"])
    assert np.array_equal(ne_t.cpu().numpy(), host["ne"])
    assert np.array_equal(st_t.cpu().numpy(), host["stage"])
    assert np.array_equal(mo_t.cpu().numpy(), host["molecule"])
    assert np.array_equal(lv_t.cpu().numpy(), host["level"])
    assert np.array_equal(it_t.cpu().numpy(), host["iterations"])
    # outputs left out are not written
    ne2 = torch.zeros_like(ne_t)
    chemeq.chemeq_dev(tb, T_t, Nt_t, ne2)
    torch.cuda.synchronize()
    assert np.array_equal(ne2.cpu().numpy(), host["ne"])


def test_pure_hydrogen_is_saha_on_gpu():
    tb = hydrogen_table()
    T = np.array([[4000.0, 5000.0, 6000.0, 8000.0, 10000.0, 20000.0]])
    Nt = np.array([[1e17, 1e16, 1e15, 1e14, 1e13, 1e12]])
    print("chemeq pure hydrogen: 6 points", flush=True)
    got = chemeq.chemeq(tb, T, Nt)
    want = np.array([saha_ne(a, b) for a, b in zip(T[0], Nt[0])])
    err = np.abs(got["ne"][0] / want - 1.0).max()
    print("  vs Saha: max rel %.2e" % err)
    assert err < 1e-8
    assert np.allclose(got["level"][0, 0], got["stage"][0, 0], rtol=1e-14)


def test_table_without_molecules_or_levels(ref):
    import copy
    tb = copy.deepcopy(ref["table"])
    tb.nmolecule = 0
    tb.nlevel = 0
    p = workloads.chemeq_problem(4, 2, seed=8)
    print("chemeq without molecules or levels", flush=True)
    got = chemeq.chemeq(tb, p["T"], p["Nt"])
    want = oracle_all(tb, p["T"], p["Nt"])
    assert np.array_equal(got["iterations"], want["iterations"])
    compare(tb, got, want["ne"], want["stage"], None, None, want["stage"].sum(axis=1), 1e-9)


def test_unsupported_tables_are_refused(ref):
    import copy
    from snapi_amd._native import SnapiError
    T = np.full((1, 3), 5000.0)
    Nt = np.full((1, 3), 1e16)
    bad = copy.deepcopy(ref["table"])
    bad.molecule[0].element[0] = bad.nelement + 1
    with pytest.raises(SnapiError):
        chemeq.chemeq(bad, T, Nt)
    bad = copy.deepcopy(ref["table"])
    bad.element[0].nstage = abi.CHEM_MAX_STAGES + 1
    with pytest.raises(SnapiError):
        chemeq.chemeq(bad, T, Nt)
    # a bound-free transition from the highest stage has no stage above it
    bad = copy.deepcopy(ref["table"])
    bad.bf[0].stage = bad.element[bad.bf[0].element - 1].nstage - 1
    with pytest.raises(SnapiError):
        chemeq.chemeq(bad, T, Nt)
