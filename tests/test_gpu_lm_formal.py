"""HIP LM normal-equation step and scalar formal solver against the oracle
and the reference golden vectors."""
import math

import numpy as np
import pytest

import _golden
import _oracle
from snapi_amd import abi, api_extra

pytestmark = pytest.mark.gpu


def test_lm_step_fp64_matches_reference_golden():
    d = _golden.load("lmstep.npz")
    for s in ("s1", "s2", "s3"):
        print("lm golden case", s, flush=True)
        J, r, lm = d[s + "_J"], d[s + "_r"], float(d[s + "_lm"])
        JTJ, rhs, corr = api_extra.lm_step(J[None], r[None], np.array([lm]), abi.SNAPI_FP64)
        assert np.allclose(JTJ[0], d[s + "_JTJ"], rtol=1e-10, atol=1e-10)
        assert np.allclose(rhs[0], d[s + "_rhs"], rtol=1e-10, atol=1e-10)
        ref = d[s + "_corr"]
        assert np.max(np.abs(corr[0] - ref)) / np.max(np.abs(ref)) < 1e-7, s


def test_lm_step_batched_fp64_and_mixed():
    rng = np.random.default_rng(3)
    npix, nrow, npar = 37, 1204, 14
    print("lm batch", npix, nrow, npar, flush=True)
    J = rng.standard_normal((npix, nrow, npar))
    r = rng.standard_normal((npix, nrow))
    lm = 10.0 ** rng.uniform(-3, 3, npix)
    JTJ, rhs, corr = api_extra.lm_step(J, r, lm, abi.SNAPI_FP64)
    JTJ32, rhs32, corr32 = api_extra.lm_step(J.astype(np.float32), r.astype(np.float32), lm, abi.SNAPI_MIXED)
    for p in range(npix):
        G, b, c = _oracle.lm_step(J[p], r[p], lm[p])
        assert np.allclose(JTJ[p], G, rtol=1e-12, atol=1e-9)
        assert np.allclose(rhs[p], b, rtol=1e-12, atol=1e-9)
        assert np.max(np.abs(corr[p] - c)) / np.max(np.abs(c)) < 1e-9
        # float inputs: J^T J accumulated in float on the matrix cores
        Jf = J[p].astype(np.float32).astype(np.float64)
        rf = r[p].astype(np.float32).astype(np.float64)
        G2, b2, c2 = _oracle.lm_step(Jf, rf, lm[p])
        assert np.max(np.abs(JTJ32[p] - G2)) / np.max(np.abs(G2)) < 1e-5
        assert np.max(np.abs(corr32[p] - c2)) / np.max(np.abs(c2)) < 1e-3


def test_lm_step_ragged_rows():
    """Row counts not divisible by the 4 rows an MFMA step consumes."""
    rng = np.random.default_rng(5)
    for nrow in (1, 3, 5, 483):
        J = rng.standard_normal((2, nrow, 7))
        r = rng.standard_normal((2, nrow))
        JTJ, rhs, corr = api_extra.lm_step(J, r, np.array([1.0, 0.1]), abi.SNAPI_FP64)
        for p in range(2):
            G, b, c = _oracle.lm_step(J[p], r[p], [1.0, 0.1][p])
            assert np.allclose(JTJ[p], G, rtol=1e-12, atol=1e-12)
            assert np.allclose(corr[p], c, rtol=1e-8, atol=1e-8)


def test_formal_scalar_matches_reference_golden():
    d = _golden.load("formal_scalar.npz")
    grid = d["grid"]
    for tag, th in (("out", math.pi), ("in", 0.0)):
        opem, sl = d[tag + "_opem"], d[tag + "_sl"]
        ls = np.unique(opem[:, 0]).astype(int)
        op = np.stack([opem[opem[:, 0] == l, 2] for l in ls])
        em = np.stack([opem[opem[:, 0] == l, 3] for l in ls])
        g = np.repeat(grid[None], len(ls), axis=0)
        ct = np.full(len(ls), math.cos(th))
        I, L = api_extra.formal_scalar(g, op, em, ct)
        for i, l in enumerate(ls):
            ref = sl[sl[:, 0] == l]
            assert np.max(np.abs(I[i] - ref[:, 2])) / np.max(np.abs(ref[:, 2])) < 1e-10
            assert np.max(np.abs(L[i] - ref[:, 3])) < 1e-12
