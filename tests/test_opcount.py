"""The operation count behind the synthesis roofline (oracle/opcount.cc).

The counting harness compiles the oracle's C synthesis with an op-counting
scalar; these checks show that it runs the oracle's arithmetic unchanged
and that its value numbering counts each distinct operation once per
column.  CPU only."""
import json
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "oracle"))

import opcount  # noqa: E402

from tests import _oracle  # noqa: E402


@pytest.fixture(scope="module")
def sample():
    tb, ray, st, lam = opcount.cube512_sample()
    return tb, ray, st, lam


def test_harness_reproduces_the_oracle_bit_for_bit(sample):
    tb, ray, st, lam = sample
    st2, lam2 = st[:2], lam[::13]
    res = opcount.count(tb, ray, st2, lam2)
    ref, rc = _oracle.synth(tb, ray, st2, lam2, nthreads=1)
    assert rc == 0
    for c in range(2):
        assert np.array_equal(res[c][2], ref[c]), np.max(np.abs(res[c][2] - ref[c]))


def test_repeats_are_counted_once(sample):
    """Every wavelength given twice: the raw count doubles, the counted
    operations stay; a column given twice counts the same each time."""
    tb, ray, st, lam = sample
    one = opcount.count(tb, ray, st[:1], lam[::20])
    dup = opcount.count(tb, ray, np.concatenate([st[:1], st[:1]]), np.repeat(lam[::20], 2))
    assert dup[0][0] == one[0][0] == dup[1][0]
    assert dup[0][1] == 2 * one[0][1]


def test_wavelength_independent_work_is_counted_once_per_column(sample):
    """Two wavelength sets counted apart and together: the raw counts add,
    the counted operations of the union fall short of the sum by the work
    both sets share, the per-depth and per-line quantities that do not
    depend on the wavelength (at least ten operations per line and depth)."""
    tb, ray, st, lam = sample
    A, B = lam[:9], lam[150:159]
    a = opcount.count(tb, ray, st[:1], A)[0]
    b = opcount.count(tb, ray, st[:1], B)[0]
    ab = opcount.count(tb, ray, st[:1], np.concatenate([A, B]))[0]
    assert ab[1] == a[1] + b[1]
    shared = a[0] + b[0] - ab[0]
    assert shared >= 10 * tb.nline * st.shape[2]
    assert shared < 0.2 * ab[0]


def test_frozen_table_matches_a_fresh_count(sample):
    tb, ray, st, lam = sample
    table = json.load(open(os.path.join(REPO, "profiles", "synth_flops.json")))
    c = table["cube512"]
    assert (c["columns"], c["depths"], c["wavelengths"]) == (st.shape[0], st.shape[2], lam.size)
    fresh = opcount.count(tb, ray, st[:1], lam)[0][0]
    assert fresh == c["counted_per_column"][0]
    assert np.isclose(c["flops_per_lambda_point"], np.mean(c["counted_per_column"]) / lam.size)


# ---- chemistry (oracle/opcount_chem.cc) ----

@pytest.fixture(scope="module")
def chem_cases():
    from snapi_amd import workloads
    return workloads.chemeq_problem(4, 4), workloads.hequilibrium_problem(4, 4)


def test_chem_harness_reproduces_the_oracle_bit_for_bit(chem_cases):
    pc, ph = chem_cases
    res = opcount.count_chem(pc["table"], pc["T"][:2], pc["Nt"][:2], 0)
    ne = _oracle.chemeq(pc["table"], pc["T"][:2], pc["Nt"][:2], nthreads=1)[0]
    for c in range(2):
        assert np.array_equal(res[c][3], ne[c])
    res = opcount.count_chem(ph["table"], ph["T"][:2], ph["logtau"][:2], 1)
    Nt, _, _, _, it = _oracle.hequilibrium(ph["table"], ph["T"][:2], ph["logtau"][:2], nthreads=1)
    for c in range(2):
        assert np.array_equal(res[c][3], Nt[c])
        assert res[c][2] == int(it[c].sum()) + ph["T"].shape[1]  # the march's iterations plus the final pass


def test_chem_repeats_are_counted_once(chem_cases):
    """A column's points given twice: the raw count doubles, the counted
    operations stay."""
    pc, _ = chem_cases
    one = opcount.count_chem(pc["table"], pc["T"][:1], pc["Nt"][:1], 0)[0]
    two = opcount.count_chem(pc["table"], np.tile(pc["T"][:1], 2), np.tile(pc["Nt"][:1], 2), 0)[0]
    assert two[0] == one[0] and two[1] == 2 * one[1]


def test_frozen_chem_table_matches_a_fresh_count():
    from snapi_amd import workloads
    table = json.load(open(os.path.join(REPO, "profiles", "chem_flops.json")))
    pc = workloads.chemeq_problem(256, 256)
    assert opcount.count_chem(pc["table"], pc["T"][:1], pc["Nt"][:1], 0)[0][0] == \
        table["chemeq"]["counted_per_column"][0]
    ph = workloads.hequilibrium_problem(128, 128)
    assert opcount.count_chem(ph["table"], ph["T"][:1], ph["logtau"][:1], 1)[0][0] == \
        table["hequilibrium"]["counted_per_column"][0]


# ---- ALI accumulation (oracle/opcount_ali.cc) ----

def test_ali_harness_reproduces_the_oracle_bit_for_bit():
    from snapi_amd import workloads
    p = workloads.ali_problem(1, 2, nlam=12, nmu=1)
    res = opcount.count_ali(p, [0, 1])
    for c in range(2):
        q = {k: (v[c] if k in ("colstate", "vproj", "gamma", "bgop", "bgem", "pops") else v) for k, v in p.items()}
        q["nd"], q["ntrans"] = p["colstate"].shape[2], p["atom"].ntrans
        acc = _oracle.ali_accumulate(q, p["pops"][c])
        assert np.array_equal(acc, res[c][2])
    # the wavelength- and ray-independent work (Doppler widths, damping,
    # Planck and Saha factors per depth) is counted once per column
    assert res[0][0] < res[0][1]
