"""The oracle's restatement of atmosphere::transform_responses
(atmos/atmos_rad.cc:53-101), pinned three ways: its angle coefficients
against central differences of the ray-frame angles of atmosphere::transform
(:23-51) as functions of the atmosphere-frame inclination and azimuth; the
identity on the mu = 1 ray with phi = 0; and, composed with the oracle's
per-depth responses on an inclined ray, against the reference's own
obs_stokes_responses there (tests/golden/cfg0_resp_incl.npz)."""
import os

import numpy as np

import _oracle
from snapi_amd import abi, workloads

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cfg0_resp_incl.npz")


def _field(bm, tb, pb):
    return np.array([bm * np.sin(tb) * np.cos(pb), bm * np.sin(tb) * np.sin(pb), bm * np.cos(tb)])


def test_angle_coefficients_are_derivatives_of_the_ray_angles():
    rng = np.random.default_rng(3)
    n = 40
    bm = rng.uniform(10, 3000, n)
    tb = rng.uniform(0.2, 2.9, n)
    pb = rng.uniform(-1.3, 1.3, n)  # Bx > 0: atan(By/Bx) is the azimuth itself
    for theta, phi in ((2.4, 0.6), (1.9, -1.1), (np.pi, 0.8)):
        bv = _field(bm, tb, pb)
        # unit responses to the ray-frame inclination (slot 5) and azimuth (6)
        r = np.zeros((7, n, 1, 4))
        r[5, :, 0, 0] = 1.0
        r[6, :, 0, 1] = 1.0
        out = _oracle.transform_responses(r, bv, theta, phi)
        h = 1e-6
        i_tp, a_tp = _oracle.ray_angles(_field(bm, tb + h, pb), theta, phi)
        i_tm, a_tm = _oracle.ray_angles(_field(bm, tb - h, pb), theta, phi)
        i_pp, a_pp = _oracle.ray_angles(_field(bm, tb, pb + h), theta, phi)
        i_pm, a_pm = _oracle.ray_angles(_field(bm, tb, pb - h), theta, phi)
        fd = np.array([(i_tp - i_tm), (i_pp - i_pm), (a_tp - a_tm), (a_pp - a_pm)]) / (2 * h)
        ok = np.all(np.abs(np.array([a_tp - a_tm, a_pp - a_pm])) < 1.0, axis=0)  # no atan fold in the step
        got = np.array([out[5, :, 0, 0], out[6, :, 0, 0], out[5, :, 0, 1], out[6, :, 0, 1]])
        assert ok.sum() > n // 2
        assert np.max(np.abs(got - fd)[:, ok]) < 1e-6, (theta, phi)


def test_identity_on_the_vertical_ray():
    rng = np.random.default_rng(4)
    nd = 11
    bv = _field(rng.uniform(10, 3000, nd), rng.uniform(0.2, 2.9, nd), rng.uniform(-1.3, 1.3, nd))
    r = rng.standard_normal((7, nd, 5, 4))
    out = _oracle.transform_responses(r, bv, np.pi, 0.0)
    assert np.max(np.abs(out - r)) < 1e-12
    out = _oracle.transform_responses(r, bv, 2.0, 0.7)
    assert np.array_equal(out[3], r[3] * np.cos(0.7))
    for p in (0, 1, 2, 4):
        assert np.array_equal(out[p], r[p])


def test_oracle_composition_matches_the_reference_on_an_inclined_ray():
    g = workloads.response_column_from_dump(GOLD)
    d = np.load(GOLD)
    el, az = float(d["el"]), float(d["az"])
    ray = abi.SnapiRay(theta=el, lam_mid=g["ray"].lam_mid, tau_grid=1, hminus=1, hminus_mid=1)
    st, resp, rc = _oracle.responses(g["table"], ray, g["state"][None], g["dstate"][None], g["lam_vac"])
    assert rc == 0
    assert np.max(np.abs(st[0] - g["stokes"])) / g["stokes"][:, 0].max() < 1e-10
    mine = _oracle.transform_responses(resp[0], d["bvec"], el, az)
    ref = g["resp"]
    for p in range(7):
        scale = np.maximum(np.abs(ref[p]).max(axis=(0, 1)), 1e-300)
        e = np.max(np.abs(mine[p] - ref[p]) / scale)
        print("  parameter %d: max %.2e of its largest response" % (p + 1, e))
        assert e < 1e-6, (p, e)
