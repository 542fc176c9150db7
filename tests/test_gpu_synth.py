"""HIP synthesis against the FP64 oracle and the reference golden vectors.
Tolerances: FP64 path 1e-5 relative to the oracle (and the reference);
mixed-precision path 1e-4 of the continuum intensity Ic."""
import math
import sys

import numpy as np
import pytest

import _oracle
from snapi_amd import abi, api, workloads

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def ic_of(stokes):
    return np.max(stokes[..., 0])


def test_cfg0_fp64_matches_reference_and_oracle(cfg0):
    _say("cfg0 fp64: 1 column, nd=61, 402 wavelengths")
    out = api.synth_stokes(cfg0["table"], cfg0["ray"], cfg0["state"], cfg0["lam_vac"], abi.SNAPI_FP64)
    ref = cfg0["stokes"]
    orc, _ = _oracle.synth(cfg0["table"], cfg0["ray"], cfg0["state"], cfg0["lam_vac"])
    ic = ic_of(ref)
    assert np.max(np.abs(out[0, :, 0] - ref[:, 0]) / ref[:, 0]) < 1e-5
    assert np.max(np.abs(out[0] - ref)) / ic < 1e-5
    assert np.max(np.abs(out - orc)) / ic < 1e-5


def test_cfg0_mixed_within_1e4_ic(cfg0):
    _say("cfg0 mixed: 1 column, nd=61, 402 wavelengths")
    st = cfg0["state"].astype(np.float32)
    out = api.synth_stokes(cfg0["table"], cfg0["ray"], st, cfg0["lam_vac"], abi.SNAPI_MIXED)
    ref = cfg0["stokes"]
    err = np.max(np.abs(out[0].astype(np.float64) - ref)) / ic_of(ref)
    _say("  mixed error vs reference: %.3e Ic" % err)
    assert err < 1e-4


@pytest.fixture(scope="module")
def cube():
    t = workloads.template_cfg1()
    st = workloads.cube_state(t, 16, 16, seed=7, dtype=np.float64)
    return t, st


def test_cube_fp64_and_mixed_against_oracle(cube):
    t, st = cube
    _say("cube 16x16 nd=71 nlambda=301: fp64 + mixed vs oracle")
    orc, rc = _oracle.synth(t["table"], t["ray"], st, t["lam_vac"])
    assert rc == 0
    g64 = api.synth_stokes(t["table"], t["ray"], st, t["lam_vac"], abi.SNAPI_FP64)
    ic = ic_of(orc)
    rel = np.max(np.abs(g64 - orc)) / ic
    _say("  fp64 max error %.3e Ic" % rel)
    assert rel < 1e-5
    g32 = api.synth_stokes(t["table"], t["ray"], st.astype(np.float32), t["lam_vac"], abi.SNAPI_MIXED)
    err = np.abs(g32.astype(np.float64) - orc) / ic
    _say("  mixed max error %.3e Ic, mean %.3e" % (err.max(), err.mean()))
    assert err.max() < 1e-4


def test_wavelength_order_leaves_results_unchanged(cube):
    """The float kernels take a workgroup's wavelengths in order of their
    nearest line and distance from it (up to 2048 wavelengths; beyond that
    in index order).  Each wavelength is independent, so a wavelength's
    Stokes vector is the same bits whatever else the launch holds: the cube
    grid reversed, and the grid followed by 1,800 far points (2,101
    wavelengths, index order)."""
    t, st = cube
    st32 = st[:8].astype(np.float32)
    lam = t["lam_vac"]
    _say("wavelength order: 8 columns, %d / reversed / %d wavelengths, mixed" % (len(lam), len(lam) + 1800))
    base = api.synth_stokes(t["table"], t["ray"], st32, lam, abi.SNAPI_MIXED)
    rev = api.synth_stokes(t["table"], t["ray"], st32, lam[::-1].copy(), abi.SNAPI_MIXED)
    assert np.array_equal(base, rev[:, ::-1])
    far = np.concatenate([lam, lam[-1] + 1e-9 * np.arange(1, 1801)])
    big = api.synth_stokes(t["table"], t["ray"], st32, far, abi.SNAPI_MIXED)
    assert np.array_equal(base, big[:, :len(lam)])


def test_edge_cases_against_oracle(cfg0):
    """Two depth points, zero field, one wavelength, more wavelengths than
    threads in a block, an inclined ray, no H-, and a height grid."""
    base = cfg0["state"][0]
    lam = cfg0["lam_vac"]
    table, ray = cfg0["table"], cfg0["ray"]
    cases = []
    cases.append(("nd=2", base[:, -2:].copy(), lam[:7], ray))
    s = base.copy()
    s[abi.F_BMAG] = 0.0
    cases.append(("B=0", s, lam, ray))
    cases.append(("nlambda=1", base, lam[100:101], ray))
    lam_long = np.concatenate([lam, lam + 2e-9, lam + 4e-9])  # 1206 > threads per block
    cases.append(("nlambda=1206", base, lam_long, ray))
    cases.append(("mu=0.6", base, lam[::3], abi.SnapiRay(theta=math.pi - math.acos(0.6), lam_mid=ray.lam_mid,
                                                           tau_grid=1, hminus=1)))
    cases.append(("no H-", base, lam[::3], abi.SnapiRay(theta=math.pi, lam_mid=ray.lam_mid, tau_grid=1, hminus=0)))
    for name, st, lm, ry in cases:
        _say("edge case %s" % name)
        orc, _ = _oracle.synth(table, ry, st[None], lm)
        g = api.synth_stokes(table, ry, st[None], lm, abi.SNAPI_FP64)
        ic = max(ic_of(orc), 1e-300)
        assert np.max(np.abs(g - orc)) / ic < 1e-5, name
        if name == "B=0":
            assert np.max(np.abs(g[..., 1:])) / ic < 1e-12


def test_bad_arguments_are_rejected(cfg0):
    from snapi_amd._native import SnapiError
    st = cfg0["state"]
    with pytest.raises(ValueError):
        api.synth_stokes(cfg0["table"], cfg0["ray"], st[:, :-2], cfg0["lam_vac"])
    with pytest.raises(SnapiError):
        api.synth_stokes(cfg0["table"], cfg0["ray"], st[:, :, :1], cfg0["lam_vac"])  # nd = 1
