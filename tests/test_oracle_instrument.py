"""The oracle's PSF convolution, scattered light and fit regularisation
against the compiled reference (tests/golden/psf.npz, reg.npz) and against
what the definitions fix.  CPU only."""
import numpy as np

import _golden
import _oracle

NAMES = {1: "T", 2: "vt", 3: "vlos", 4: "B", 5: "theta", 6: "phi"}


def test_regularisation_matrix_matches_reference():
    """regularize_hessian's additions to a zero J^T J, bit for bit, on five
    node models: Tikhonov over 7 and 2 nodes, departure from zero, a single
    Tikhonov node coupling to the next parameter's node, and runs that end
    the node list (nothing for departure from zero, the first difference
    only for Tikhonov).  The reference leaves the right-hand side alone."""
    g = _golden.load("reg.npz")
    for k in range(5):
        codes, reg, G = g["codes%d" % k], g["reg%d" % k], g["G%d" % k]
        nodes = [(NAMES[int(c)], 0.0) for c in codes]
        rg = {NAMES[int(c)]: (int(t), a) for c, t, a in reg}
        assert np.array_equal(_oracle.reg_matrix(nodes, rg), G), k
        assert not np.any(g["rhs%d" % k])


def test_regularisation_is_alpha_squared_r_transpose_r():
    """Tikhonov over a run of n nodes is alpha^2 times the second-difference
    stencil of a chain (closed form), departure from zero alpha^2 I."""
    nodes = [("T", 0.0)] * 4 + [("B", 0.0)] * 3 + [("phi", 0.0)]
    G = _oracle.reg_matrix(nodes, {"T": (2, 2.0)})
    # rows R_0 = -x0 + x1, R_i = x_i - x_(i-1): R^T R of a 4-chain with the first row doubled
    R = np.zeros((4, 4))
    R[0, 0], R[0, 1] = -1.0, 1.0
    for i in range(1, 4):
        R[i, i], R[i, i - 1] = 1.0, -1.0
    assert np.allclose(G[:4, :4], 4.0 * R.T @ R) and not np.any(G[4:]) and not np.any(G[:, 4:])
    G = _oracle.reg_matrix(nodes, {"B": (1, 3.0)})
    assert np.allclose(G, np.diag([0, 0, 0, 0, 9.0, 9.0, 9.0, 0]))


def test_psf_convolution_matches_reference_interior():
    """convolve_spectra_with_psf of the reference from the (npsf-1)/2-th
    wavelength on, where its window is the centred one (odd sample counts
    1, 7, 13 and 31), bit for bit; before that the reference anchors the
    PSF at the first wavelength and the defined version differs."""
    g = _golden.load("psf.npz")
    for k in range(4):
        p, S, out = g["p%d_psf" % k], g["p%d_in" % k], g["p%d_out" % k]
        h = (p.size - 1) // 2
        o = _oracle.psf_convolve(S, p)
        assert np.array_equal(o[h:], out[h:]), k
        if h:
            assert np.max(np.abs(o[:h] - out[:h])) > 1e-3


def test_psf_convolution_definition():
    """A unit impulse returns the samples reversed and normalised around it;
    a constant stays constant where the whole window is on the grid; an even
    sample count uses all its samples; the convolution is linear."""
    rng = np.random.default_rng(4)
    for n in (1, 4, 9):
        p = rng.uniform(0.1, 1.0, n)
        h = (n - 1) // 2
        S = np.zeros((40, 2))
        S[20, :] = 1.0
        o = _oracle.psf_convolve(S, p)
        ref = np.zeros(40)
        for j in range(n):
            ref[20 + h - j] = p[j] / p.sum()
        assert np.allclose(o[:, 0], ref, rtol=1e-15, atol=1e-17)
        c = _oracle.psf_convolve(np.full((40, 1), 3.0), p)
        assert np.allclose(c[h:40 - (n - 1 - h), 0], 3.0, rtol=1e-14)
        A, B = rng.normal(size=(40, 3)), rng.normal(size=(40, 3))
        assert np.allclose(_oracle.psf_convolve(A + 2 * B, p),
                           _oracle.psf_convolve(A, p) + 2 * _oracle.psf_convolve(B, p), rtol=1e-12, atol=1e-14)


def test_scattered_light_closed_form():
    S = np.random.default_rng(1).normal(size=(3, 11, 4))
    o = _oracle.add_scattered_light(S, 0.05, 2.0)
    assert np.array_equal(o[..., 1:], S[..., 1:])
    assert np.allclose(o[..., 0] - S[..., 0], 0.1, rtol=1e-14)
