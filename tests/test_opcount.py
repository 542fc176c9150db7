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
