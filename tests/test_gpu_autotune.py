"""The launch-shape autotuner: every shape of the synthesis kernels (three
occupancy bounds per precision) and of the tangent kernels (three workgroup
sizes) gives the same Stokes vectors and node Jacobians bit for bit, and the
tuner times each shape on the first call of a problem shape and keeps one."""
import numpy as np
import pytest

from snapi_amd import abi, api, autotune, invert, workloads

pytestmark = pytest.mark.gpu

NODES = [("vt", -2.0), ("vlos", -4.0), ("vlos", 0.0), ("B", -4.0), ("B", 0.0), ("theta", -2.0), ("phi", -2.0)]


@pytest.fixture(autouse=True)
def _restore():
    old = autotune.mode()
    yield
    autotune.force(autotune.SYNTH, -1)
    autotune.force(autotune.TANGENT, -1)
    autotune.mode(old)


@pytest.mark.parametrize("prec", [abi.SNAPI_FP64, abi.SNAPI_MIXED])
def test_synthesis_shapes_agree_bit_for_bit(prec):
    t = workloads.template_cfg1()
    st = workloads.cube_state(t, 8, 8, seed=4, dtype=np.float64 if prec == abi.SNAPI_FP64 else np.float32)
    lam = t["lam_vac"]
    outs = []
    for c in range(3):
        print("synthesis, precision %d, shape %d: 64 columns x %d wavelengths" % (prec, c, len(lam)), flush=True)
        autotune.force(autotune.SYNTH, c)
        outs.append(api.synth_stokes(t["table"], t["ray"], st, lam, prec))
    for c in (1, 2):
        assert np.array_equal(outs[0], outs[c]), c


def test_tangent_shapes_agree_bit_for_bit():
    t = workloads.template_cfg1()
    rng = np.random.default_rng(3)
    npix = 5
    v = np.column_stack([rng.uniform(0.3e5, 1.8e5, npix), rng.normal(0, 1.5e5, npix), rng.normal(0, 1.5e5, npix),
                         rng.uniform(200, 2500, npix), rng.uniform(200, 2500, npix),
                         np.arccos(rng.uniform(-0.9, 0.9, npix)), rng.uniform(0.1, 1.4, npix)])
    st = np.repeat(t["state"][None], npix, axis=0)
    model = invert.node_model(NODES)
    for prec in (abi.SNAPI_FP64, abi.SNAPI_MIXED):
        sv = st if prec == abi.SNAPI_FP64 else st.astype(np.float32)
        res = []
        for c in range(3):
            print("tangent node Jacobian, precision %d, workgroup shape %d" % (prec, c), flush=True)
            autotune.force(autotune.TANGENT, c)
            res.append(invert.node_jacobian(t["table"], t["ray"], model, sv, v, t["lam_vac"], prec, 2))
        for c in (1, 2):
            assert np.array_equal(res[0][0], res[c][0]) and np.array_equal(res[0][1], res[c][1]), (prec, c)


def test_tuner_times_every_shape_once():
    t = workloads.template_cfg1()
    st = workloads.cube_state(t, 16, 16, seed=5, dtype=np.float32)
    lam = t["lam_vac"][:257]
    autotune.mode(1)
    autotune.clear()
    print("tuned synthesis: 256 columns x %d wavelengths" % len(lam), flush=True)
    a = api.synth_stokes(t["table"], t["ray"], st, lam, abi.SNAPI_MIXED)
    rep = [r for r in autotune.report() if r["kernel"] == autotune.SYNTH and r["nlambda"] == len(lam)]
    print("  report:", rep, flush=True)
    assert len(rep) == 1
    r = rep[0]
    assert r["precision"] == abi.SNAPI_MIXED and r["nd"] == st.shape[2] and 0 <= r["choice"] < 3
    assert len(r["ms"]) == 3 and all(m > 0 for m in r["ms"])
    assert r["ms"][r["choice"]] == min(r["ms"])
    # a second call reuses the choice (no new entry) and gives the same result
    b = api.synth_stokes(t["table"], t["ray"], st, lam, abi.SNAPI_MIXED)
    assert np.array_equal(a, b)
    assert len([x for x in autotune.report() if x["kernel"] == autotune.SYNTH and x["nlambda"] == len(lam)]) == 1


def test_ali_accumulation_shapes_agree_bit_for_bit():
    """The float ALI accumulation of the Ca II atom (10 lines, 5 bound-free)
    in its three occupancy bounds: the same sums bit for bit."""
    from snapi_amd import ali
    p = workloads.ali_problem(4, 4, nlam=64, nmu=2)
    args = [p[k] for k in ("colstate", "vproj", "gamma", "bgop", "bgem", "pops")]
    res = []
    for c in range(3):
        print("ALI accumulation, mixed, occupancy shape %d: 16 columns" % c, flush=True)
        autotune.force(autotune.ALI, c)
        res.append(ali.accumulate(p["atom"], p["grid"], *args, precision=abi.SNAPI_MIXED))
    autotune.force(autotune.ALI, -1)
    for c in (1, 2):
        assert np.array_equal(res[0], res[c]), c
