"""Analytic node responses in the oracle: the node interpolation's exact
derivatives (complex-step differentiation of atmospheric_interpolation,
cmlib/mathtools.cc:1208-1315, and of the field projection) and the tangent of
the hydrostatic march (the linearised equations of atmos_pp.cc:372-455), in
place of the reference's central differences (atmos_fit.cc:1225-1305).

Pins: the complex-step maps against the reference-step differences of the
same interpolation (their difference is the differences' O(h^2) error) and
the inclination row against its closed form; the hydrostatic tangent against
central differences of the march solved to rounding (orc_hequilibrium_tol),
which a dropped density or opacity derivative, a wrong neighbour term or a
missing kT fails by orders of magnitude, and against the reference's own
+- 2.5 K rebuilds; the whole analytic Jacobian against the reference's way to
the Jacobian bar.  CPU only."""
import numpy as np
import pytest

import _golden
import _oracle
from snapi_amd import abi, workloads

NODES = [("T", -5.0), ("T", -4.0), ("T", -3.0), ("T", -2.0), ("T", -1.0), ("T", 0.0), ("T", 0.5),
         ("vt", -2.0), ("vlos", -4.0), ("vlos", 0.0), ("B", -4.0), ("B", 0.0), ("theta", -2.0), ("phi", -2.0)]


@pytest.fixture(scope="module")
def pix():
    t = workloads.template_cfg1()
    rng = np.random.default_rng(71)
    lt = np.log10(-t["state"][abi.F_GRID])
    v = np.zeros(len(NODES))
    v[:7] = np.interp([n[1] for n in NODES[:7]], lt, t["state"][abi.F_T]) + rng.normal(0.0, 100.0, 7)
    v[7:] = [0.9e5, -1.1e5, 0.4e5, 1800.0, 700.0, 2.1, 0.6]
    return t, v


def _fd_map(nodes, v, state, scale):
    """Central differences of the interpolation with the reference's steps
    times scale (T rows: dT/dnode)."""
    x = np.log10(-state[abi.F_GRID])
    tn = [i for i, (p, _) in enumerate(nodes) if p == "T"]
    out = np.zeros((len(nodes), state.shape[1]))
    fd = _oracle.node_map(*zip(*[(m, vv) for m, vv in zip(nodes, v) if m[0] != "T"]), state) if len(tn) < len(nodes) \
        else None
    k = 0
    for n, (p, _) in enumerate(nodes):
        if p == "T":
            h = 5.0 * scale
            vp, vm = np.array(v, float), np.array(v, float)
            vp[n] += h / 2
            vm[n] -= h / 2
            tpos = np.array([nodes[i][1] for i in tn])
            out[n] = (_oracle._interp_temp(tpos, vp[tn], x) - _oracle._interp_temp(tpos, vm[tn], x)) / h
        else:
            out[n] = fd[k]
            k += 1
    return out


def test_interpolation_maps_match_differences(pix):
    t, v = pix
    st = t["state"]
    an = _oracle.node_map_analytic(NODES, v, st)
    fd = _fd_map(NODES, v, st, 1.0)
    sc = np.maximum(np.abs(an).max(axis=1, keepdims=True), 1e-300)
    err = np.abs(an - fd) / sc
    print("complex-step vs reference-step differences: max %.2e per row %s" % (err.max(), np.round(err.max(1), 9)))
    assert err.max() < 1e-5
    # every node moves its own field somewhere, and the maps are zero
    # outside the interpolation's reach of a single node
    assert np.all(np.abs(an).max(axis=1) > 0)
    # B, inclination and azimuth rows carry the field projection's chain rule
    # (|B| linear in B, inclination [rad] 1:1 in theta away from the clamps)
    b = NODES.index(("theta", -2.0))
    assert np.allclose(an[b], 1.0, atol=1e-12)


def test_hydrostatic_tangent_matches_converged_march():
    d = _golden.chemeq()
    tb = abi.make_chem_table({k: vv for k, vv in d.items() if k != "table"}, referent=1)
    rec = d["he"]
    T, lt = rec[:, 0].copy(), rec[:, 1].copy()
    N = _oracle.hequilibrium_tol(tb, T, lt, 1e-13, 500)
    dT = np.stack([np.exp(-((lt + c) / 0.8) ** 2) for c in (5.0, 3.0, 1.0, 0.0)] + [np.ones_like(lt)])
    dN = _oracle.he_tangent(tb, T, lt, N, dT)
    for k in range(len(dT)):
        eps = 0.5
        fd = (_oracle.hequilibrium_tol(tb, T + eps * dT[k], lt, 1e-13, 500) -
              _oracle.hequilibrium_tol(tb, T - eps * dT[k], lt, 1e-13, 500)) / (2 * eps)
        e = np.abs(dN[k] - fd).max() / np.abs(fd).max()
        # the reference's rebuilds at +- 2.5 K with its 1e-2 stopping rule
        N0 = _oracle.hequilibrium(tb, T, lt)[0][0]
        ref = (_oracle.hequilibrium(tb, T + 2.5 * dT[k], lt)[0][0] -
               _oracle.hequilibrium(tb, T - 2.5 * dT[k], lt)[0][0]) / 5.0
        dN0 = _oracle.he_tangent(tb, T, lt, N0, dT[k])[0]
        er = np.abs(dN0 - ref).max() / np.abs(ref).max()
        print("direction %d: tangent vs converged march %.2e, vs the reference's rebuilds %.2e" % (k, e, er))
        assert e < 1e-4 and er < 1e-4
    # a uniform heating lowers the density everywhere (pressure fixed by the
    # column mass above, P = Nt k T)
    assert np.all(dN[-1] < 0)


def test_analytic_jacobian_matches_reference_way(pix):
    t, v = pix
    chem, chem_he, smap = workloads.eos_tables("cfg1_template", t["ray"].lam_mid)
    lam = t["lam_vac"][::10]
    s0, J0 = _oracle.node_jacobian_eos(t["table"], t["ray"], NODES, v, t["state"], lam, chem, chem_he, smap)
    s1, J1 = _oracle.node_jacobian_eos(t["table"], t["ray"], NODES, v, t["state"], lam, chem, chem_he, smap,
                                       analytic=True)
    assert np.array_equal(s0, s1)
    e = np.abs(J1 - J0) / np.maximum(np.abs(J0).max(axis=0, keepdims=True), 1e-300)
    print("analytic vs reference-way Jacobian: max %.2e, T nodes %.2e" % (e.max(), e[:, :7].max()))
    assert e.max() < 1e-3
