"""The NLTE workload on the GPU at its stated sampling: Ca II 5 levels +
continuum in plane-parallel columns of 81 depths with the chromospheric rise
to 8000 K and seeded +-10 % perturbations, 220 wavelengths, 5 Gauss-Radau mu
x 2 directions, iterated to a relative population change below 1e-4
(snapi_ali_solve against the oracle's nltepops restatement,
atmos/atmos_pop.cc:67-240), then polarized synthesis of Ca II 8542 A with
|B| up to 500 G from the converged populations (snapi_synth_stokes against
the oracle's obs_stokes, atmos/atmos_obs.cc:16-87).

Tolerances: FP64 populations within 1e-6 of the oracle's with equal
iteration counts (the same arithmetic up to contraction and summation
order); mixed precision within 2e-3: the iteration stops once a step changes
the populations by less than 1e-4, and the float solution's rounding moves
where that happens by the tolerance over the contraction rate.  Stokes:
FP64 within 1e-5 of the continuum, mixed within 1e-4."""
import numpy as np
import pytest

import _oracle
from snapi_amd import abi, ali, api, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfg4():
    p = workloads.ali_problem(2, 2, nd=81, nlam=220, nmu=5, seed=9)
    q = dict(p)
    q["nd"] = p["colstate"].shape[2]
    op, oit, orc = _oracle.ali_solve(q, p["pops"], 1e-4, 300, ncol=4)
    return p, op, oit


def _args(p):
    return (p["atom"], p["grid"], p["colstate"], p["vproj"], p["gamma"], p["bgop"], p["bgem"], p["coll"])


def test_cfg4_populations_fp64_match_oracle(cfg4):
    p, op, oit = cfg4
    print("cfg4 ALI FP64: 4 columns x 81 depths, %d wavelengths, %d rays, to 1e-4" % (p["nlam"], p["nray"]),
          flush=True)
    gp, git, grc = ali.solve(*_args(p), p["pops"], tol=1e-4, max_iter=300, precision=abi.SNAPI_FP64)
    err = float(np.max(np.abs(gp / op - 1.0)))
    print("  iterations GPU %s oracle %s; populations max rel %.2e" % (git.tolist(), oit.tolist(), err), flush=True)
    assert np.array_equal(git, oit)
    assert np.all(grc >= 0) and np.all(grc < 1e-4)
    assert err < 1e-6


def test_cfg4_populations_mixed_against_oracle(cfg4):
    p, op, oit = cfg4
    gp, git, grc = ali.solve(*_args(p), p["pops"], tol=1e-4, max_iter=300, precision=abi.SNAPI_MIXED)
    err = float(np.max(np.abs(gp / op - 1.0)))
    print("cfg4 ALI mixed: iterations %s (oracle %s); populations max rel %.2e" % (git.tolist(), oit.tolist(), err),
          flush=True)
    assert np.all(grc >= 0)
    assert err < 2e-3


def test_cfg4_8542_synthesis_from_converged_populations(cfg4):
    """The chain: GPU populations -> GPU Stokes against oracle populations ->
    oracle Stokes, and both syntheses on the same inputs."""
    p, op, _ = cfg4
    gp, _, _ = ali.solve(*_args(p), p["pops"], tol=1e-4, max_iter=300, precision=abi.SNAPI_FP64)
    tb, ray, st_g, lam = workloads.ca8542_synthesis(p, gp)
    _, _, st_o, _ = workloads.ca8542_synthesis(p, op)
    print("cfg4 8542 A synthesis: %d columns, %d wavelengths, |B| up to %.0f G" % (
        st_g.shape[0], len(lam), st_g[:, abi.F_BMAG, 0].max()), flush=True)
    ref, rc = _oracle.synth(tb, ray, st_o, lam)
    assert rc == 0
    ic = float(ref[..., 0].max())
    s64 = api.synth_stokes(tb, ray, st_g, lam, abi.SNAPI_FP64)
    s32 = api.synth_stokes(tb, ray, st_g.astype(np.float32), lam, abi.SNAPI_MIXED)
    same, _ = _oracle.synth(tb, ray, st_g, lam)
    e_chain = float(np.abs(s64 - ref).max() / ic)
    e_same = float(np.abs(s64 - same).max() / ic)
    e_mixed = float(np.abs(s32.astype(np.float64) - same).max() / ic)
    vmax = float(np.abs(ref[..., 3]).max() / ic)
    print("  chain FP64 %.2e Ic, same inputs FP64 %.2e Ic, mixed %.2e Ic; max |V|/Ic %.2e" % (
        e_chain, e_same, e_mixed, vmax), flush=True)
    assert vmax > 1e-4  # the field is seen
    assert e_same < 1e-10 and e_chain < 1e-5 and e_mixed < 1e-4
