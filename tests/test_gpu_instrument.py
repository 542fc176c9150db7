"""Gaussian spectral convolution on the GPU (snapi_spectral_convolve) against
the reference's convolve_spectra_with_gauss on four seeded spectra
(tests/golden/conv.npz) and the oracle's restatement; on a node Jacobian it
is the Jacobian of the convolved spectrum (the convolution is linear).

Tolerance: both sides resample and sum in double in the same order; the
device exp differs from glibc's in the last bits, which the sums of up to
121 kernel terms carry to about 1e-11 of a component's largest value: the
bar is 1e-10 (mixed precision: 1e-6, the float rounding of the data)."""
import numpy as np
import pytest

import _golden
import _oracle
from snapi_amd import abi, instrument, invert, workloads

pytestmark = pytest.mark.gpu


def test_convolution_matches_reference_and_oracle():
    g = _golden.load("conv.npz")
    for k in range(4):
        c = "c%d" % k
        lam, w = g[c + "_lam"], float(g[c + "_width"])
        print("convolution %s: %d wavelengths, width %.0f mA" % (c, len(lam), w * 1e11), flush=True)
        out = instrument.spectral_convolve(g[c + "_in"][None], lam, w)[0]
        scale = np.max(np.abs(g[c + "_out"]), axis=0)
        e_ref = np.max(np.abs(out - g[c + "_out"]) / scale)
        o = _oracle.convolve_gauss(g[c + "_in"], lam, w)
        e_or = np.max(np.abs(out - o) / scale)
        m = instrument.spectral_convolve(g[c + "_in"][None], lam, w, abi.SNAPI_MIXED)[0]
        e_m = np.max(np.abs(m - g[c + "_out"]) / scale)
        print("  vs reference %.2e, vs oracle %.2e, mixed %.2e" % (e_ref, e_or, e_m), flush=True)
        assert e_ref < 1e-10 and e_or < 1e-10 and e_m < 1e-6


def test_jacobian_components_convolve_independently():
    """convolve_response_with_gauss convolves every (Stokes, node) column of
    the Jacobian on its own: a [npix][nlambda][4 nnode] array convolves as its
    columns do one at a time (the workgroups take up to eight columns at a
    time), in double and in float.  The resampling is not linear
    (Fritsch-Butland slopes), so the convolved Jacobian is the reference's
    approximation, not the Jacobian of the convolved spectrum."""
    t = workloads.template_cfg1()
    nodes = [("vt", -2.0), ("vlos", -4.0), ("B", -2.0), ("theta", -2.0), ("phi", -2.0)]
    model = invert.node_model(nodes)
    v = np.array([[1.0e5, 0.5e5, 800.0, 0.8, 0.4], [1.5e5, -0.5e5, 300.0, 2.0, 1.0]])
    lam = t["lam_vac"]
    w = 3.0e-10
    print("convolved node Jacobian: %d wavelengths, %d nodes, 2 pixels" % (len(lam), len(nodes)), flush=True)
    st = np.repeat(t["state"][None], 2, axis=0)
    S, J = invert.node_jacobian(t["table"], t["ray"], model, st, v, lam, abi.SNAPI_FP64, 2)
    nc = 4 * len(nodes)
    Jv = J.reshape(2, len(lam), nc)
    for prec, tol in ((abi.SNAPI_FP64, 0.0), (abi.SNAPI_MIXED, 0.0)):
        allc = instrument.spectral_convolve(Jv, lam, w, prec)
        one = np.stack([instrument.spectral_convolve(Jv[:, :, c:c + 1], lam, w, prec)[:, :, 0] for c in range(nc)],
                       axis=2)
        print("  precision %d: grouped vs one column at a time max diff %.2e" % (prec, np.max(np.abs(allc - one))),
              flush=True)
        assert np.array_equal(allc, one)
    # the Stokes spectrum convolves as the first four columns would
    sc = instrument.spectral_convolve(S, lam, w)
    assert np.all(np.isfinite(sc)) and sc.shape == S.shape


def test_long_grid_uses_smaller_groups():
    """A narrow Gaussian on a long grid: the fine grid of ~2,000 points leaves
    room for fewer columns per workgroup; results still match the oracle."""
    lam = np.linspace(6.299e-5, 6.305e-5, 400)
    rng = np.random.default_rng(5)
    S = 1.0 - 0.5 * np.exp(-((lam - 6.302e-5) / 2e-9) ** 2)[:, None] * rng.uniform(0.5, 1.0, (1, 4))
    w = 6.0e-11
    print("long grid: 400 wavelengths, width 6 mA", flush=True)
    out = instrument.spectral_convolve(S[None], lam, w)[0]
    o = _oracle.convolve_gauss(S, lam, w)
    e = np.max(np.abs(out - o)) / np.max(np.abs(o))
    print("  vs oracle %.2e" % e, flush=True)
    assert e < 1e-10


def test_psf_convolution_matches_reference_and_oracle():
    """snapi_psf_convolve against the reference's convolve_spectra_with_psf
    where its window is the centred one (tests/golden/psf.npz) and against
    the oracle's defined version everywhere (1e-14 of a component's largest
    value in FP64, 1e-6 mixed)."""
    g = _golden.load("psf.npz")
    for k in range(4):
        p, S, out = g["p%d_psf" % k], g["p%d_in" % k], g["p%d_out" % k]
        h = (p.size - 1) // 2
        print("psf convolution p%d: %d samples, %d wavelengths" % (k, p.size, S.shape[0]), flush=True)
        o = instrument.psf_convolve(S[None], p)[0]
        scale = np.max(np.abs(out), axis=0)
        e_ref = np.max(np.abs(o[h:] - out[h:]) / scale)
        e_or = np.max(np.abs(o - _oracle.psf_convolve(S, p)) / scale)
        m = instrument.psf_convolve(S[None], p, abi.SNAPI_MIXED)[0]
        e_m = np.max(np.abs(m - _oracle.psf_convolve(S, p)) / scale)
        print("  vs reference %.2e, vs oracle %.2e, mixed %.2e" % (e_ref, e_or, e_m), flush=True)
        assert e_ref < 1e-14 and e_or < 1e-14 and e_m < 1e-6


def test_psf_jacobian_columns_and_device_entry_point():
    """A [npix][nlambda][4 nnode] Jacobian convolves column by column (the
    workgroups take up to 16 columns at a time), through the device entry
    point on a stream, in both precisions."""
    import torch
    rng = np.random.default_rng(8)
    J = rng.normal(size=(3, 57, 4 * 14))
    p = rng.uniform(0.1, 1.0, 11)
    ref = np.stack([_oracle.psf_convolve(J[i], p) for i in range(3)])
    for dt, tol in ((torch.float64, 1e-13), (torch.float32, 1e-5)):
        print("psf on a Jacobian, device entry point, %s" % dt, flush=True)
        jt = torch.from_numpy(J).to("cuda", dt)
        pt = torch.from_numpy(p).to("cuda")
        s = torch.cuda.Stream()
        instrument.psf_convolve_dev(jt, pt, s)
        s.synchronize()
        assert np.max(np.abs(jt.cpu().numpy() - ref)) < tol * np.max(np.abs(ref))


def test_scattered_light():
    S = np.random.default_rng(2).normal(size=(4, 23, 4))
    print("scattered light: host and device entry points", flush=True)
    o = instrument.add_scattered_light(S, 0.07, 3.0)
    assert np.allclose(o, _oracle.add_scattered_light(S, 0.07, 3.0), rtol=0, atol=1e-15)
    import torch
    st = torch.from_numpy(S).to("cuda", torch.float32)
    instrument.add_scattered_light_dev(st, 0.07, 3.0)
    torch.cuda.synchronize()
    assert np.allclose(st.cpu().numpy(), _oracle.add_scattered_light(S, 0.07, 3.0), rtol=0, atol=1e-6)
