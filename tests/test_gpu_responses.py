"""Adjoint (reverse-sweep) response functions on the GPU against the oracle,
which restates the reference's central difference through the solver, and
against the reference's own obs_stokes_responses.  Tolerance: 1e-4 of the
largest response of each (parameter, Stokes) pair -- the reference's
difference quotient with unit steps differs from the exact derivative by
O(step^2)."""
import math
import os

import numpy as np
import pytest

import _oracle
from snapi_amd import abi, api_extra, workloads

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cfg0_resp.npz")
NAMES = ["T", "Nt", "vt", "vlos", "B", "incl", "azi"]


@pytest.fixture(scope="module")
def g():
    g = workloads.response_column_from_dump(GOLD)
    g["rray"] = abi.SnapiRay(theta=math.pi, lam_mid=g["ray"].lam_mid, tau_grid=1, hminus=1, hminus_mid=1)
    return g


def _cmp(mine, ref, tol, what):
    for p in range(7):
        scale = np.maximum(np.abs(ref[p]).max(axis=(0, 1)), 1e-300)
        err = np.max(np.abs(mine[p] - ref[p]) / scale)
        print("  %s %-5s max err %.2e of max |response|" % (what, NAMES[p], err), flush=True)
        assert err < tol, (what, NAMES[p], err)


def test_responses_match_reference_and_oracle(g):
    print("cfg0 responses: nd=61, 31 wavelengths, 7 parameters", flush=True)
    st, resp = api_extra.stokes_responses(g["table"], g["rray"], g["state"], g["dstate"], g["lam_vac"])
    assert np.max(np.abs(st[0] - g["stokes"])) / g["stokes"][:, 0].max() < 1e-10
    _cmp(resp[0], g["resp"], 1e-4, "vs reference")
    ost, oresp, rc = _oracle.responses(g["table"], g["rray"], g["state"][None], g["dstate"][None], g["lam_vac"])
    _cmp(resp[0], oresp[0], 1e-4, "vs oracle")


def test_adjoint_gradient_equals_responses_contracted(g):
    """grad = sum_{l,s} w_s (S - S_obs) dS/dq, and the dot-product test
    <J v, r> = <v, J^T r> with the response matrix."""
    rng = np.random.default_rng(1)
    st, resp = api_extra.stokes_responses(g["table"], g["rray"], g["state"], g["dstate"], g["lam_vac"])
    obs = st * (1.0 + 1e-3 * rng.standard_normal(st.shape))
    ws = np.array([1.0, 4.0, 4.0, 2.0])
    grad = api_extra.stokes_adjoint(g["table"], g["rray"], g["state"], g["dstate"], g["lam_vac"], obs, ws)
    r = (st - obs) * ws[None, None, :]
    expect = np.einsum("pzls,ls->pz", resp[0], r[0])
    for p in range(7):
        sc = np.abs(expect[p]).max()
        assert np.max(np.abs(grad[0, p] - expect[p])) / sc < 1e-9, NAMES[p]


def test_responses_cube_columns_against_oracle():
    """A few columns of the cube workload (different fields, velocities and
    temperatures), 30 wavelengths across both lines."""
    t = workloads.template_cfg1()
    st = workloads.cube_state(t, 4, 2, seed=5, dtype=np.float64)
    # thermodynamic derivatives of the synthetic state: take the template's
    # relative derivatives (any consistent input exercises the chain rule)
    nf, nd = st.shape[1:]
    ds = np.zeros((st.shape[0], 2, nf, nd))
    rng = np.random.default_rng(2)
    for f in [abi.F_NE, abi.F_NHI, abi.F_OPREF] + list(range(abi.NF_BASE, nf)):
        ds[:, 0, f] = st[:, f] * rng.uniform(-2e-3, 2e-3, (st.shape[0], 1))
        ds[:, 1, f] = st[:, f] / st[:, abi.F_NT] * rng.uniform(0.5, 1.5, (st.shape[0], 1))
    lam = t["lam_vac"][::10]
    ray = abi.SnapiRay(theta=math.pi, lam_mid=t["ray"].lam_mid, tau_grid=1, hminus=1, hminus_mid=1)
    print("cube responses: 8 columns, nd=71, %d wavelengths" % len(lam), flush=True)
    s1, r1 = api_extra.stokes_responses(t["table"], ray, st, ds, lam)
    s2, r2, rc = _oracle.responses(t["table"], ray, st, ds, lam)
    assert np.max(np.abs(s1 - s2)) / s2[..., 0].max() < 1e-10
    for c in range(st.shape[0]):
        _cmp(r1[c], r2[c], 1e-4, "col %d" % c)
