"""Line tables built from level data on the host (snapi_line_add and its
parts) against the reference's own line setup: the lines, van der Waals
parameters and Zeeman components the reference's atoms hold after
zeeman_setup in the cfg0 column (tests/golden/cfg0_synth.npz: Fe I 6301.5 and
6302.5 (s-p, Anstee-O'Mara table), H Ly alpha, Ly beta (Unsold) and H alpha
(s-p)), plus the textbook limits of the Zeeman pattern and the tables' own
grid points.  Host functions only: no device is touched."""
import numpy as np
import pytest

import _golden
from snapi_amd import abi, lines


@pytest.fixture(scope="module")
def dump():
    return _golden.load("cfg0_synth.npz"), _golden.load("chemeq.npz")


def _levels(d, s, z):
    rows = [r for r in d["levels"] if int(r[0]) == s and int(r[1]) == z]
    rows.sort(key=lambda r: int(r[2]))
    out = []
    for r in rows:
        l = int(r[6])
        out.append(lines.level(r[3], r[4], int(r[5]), -1 if l == 255 else l, r[7]))
    return out


def test_line_table_matches_reference(dump):
    d, c = dump
    masses = dict((int(k), v) for k, v in d["masses"])
    t = lines.empty_table()
    for (s, z, lo, up, tr), ld in zip(d["lines"], d["line_data"]):
        lv = _levels(d, int(s), int(z))
        lines.add_line(t, masses[int(s)], int(z), float(c["eion"][int(s), int(z)]), lv[int(lo)], lv[int(up)], ld[1],
                       ld[4])
    assert t.nline == len(d["lines"])
    for i, ld in enumerate(d["line_data"]):
        L = t.line[i]
        got = np.array([L.lam0, L.aul, L.blu, L.bul, L.grad, L.sigma, L.alpha])
        err = np.abs(got / ld - 1.0)
        assert err.max() < 1e-12, (i, err)
        assert L.split == int(d["line_split"][i])
    # Zeeman components in pi, sigma_b, sigma_r order
    for i, (s, z, lo, up, tr) in enumerate(d["lines"]):
        L = t.line[i]
        for g in range(3):
            ref = [(r[3], r[4]) for r in d["zc"] if int(r[0]) == s and int(r[1]) == tr and int(r[2]) == g]
            got = [(t.comp[L.comp0 + sum(L.ncomp[:g]) + k].shift, t.comp[L.comp0 + sum(L.ncomp[:g]) + k].strength)
                   for k in range(L.ncomp[g])]
            assert len(got) == len(ref)
            assert np.allclose(np.array(got).reshape(-1, 2), np.array(ref).reshape(-1, 2), rtol=0, atol=1e-14)


def test_zeeman_pattern_textbook_limits():
    # J = 1 -> 0 (6302.5 A): a normal triplet at 0, -g, +g with unit strengths
    split, (p, b, r) = lines.zeeman_pattern(2, 0, 2.5, 0.0)
    assert split == 1 and p == [(0.0, 1.0)] and b == [(-2.5, 1.0)] and r == [(2.5, 1.0)]
    # every group sums to one; the strength-weighted sigma shift is the
    # effective Lande factor g_eff = (g_u + g_l)/2 + (g_u - g_l)(J_u(J_u+1) - J_l(J_l+1))/4
    for j2l, j2u, gl, gu in ((4, 4, 1.83, 1.5), (2, 4, 1.2, 1.33), (5, 3, 1.2, 1.3333), (1, 3, 2.0, 1.3333)):
        split, groups = lines.zeeman_pattern(j2l, j2u, gl, gu)
        assert split == 1
        for g in groups:
            assert abs(sum(s for _, s in g) - 1.0) < 1e-14
        jl, ju = j2l / 2.0, j2u / 2.0
        geff = 0.5 * (gu + gl) + 0.25 * (gu - gl) * (ju * (ju + 1) - jl * (jl + 1))
        assert abs(sum(sh * s for sh, s in groups[2]) - geff) < 1e-12
        assert abs(sum(sh * s for sh, s in groups[1]) + geff) < 1e-12
        assert abs(sum(sh * s for sh, s in groups[0])) < 1e-12
    # J = 0 -> 0 and unsplit transitions
    assert lines.zeeman_pattern(0, 0, 1.0, 1.0)[0] == 0
    assert lines.zeeman_pattern(0, 4, 1.0, 1.0)[0] == 0


def test_vdw_damping_table_points_and_unsold():
    ev = 1.602189e-12
    ry = 2.18e-11
    # at a grid point of the s-p table (n*_s = 2.0, n*_p = 2.3) the value is
    # the tabulated one (row 10, column 10 of Anstee & O'Mara: 522 a0^2,
    # alpha 0.235 ... read from the table itself)
    from math import gamma, pi
    e_ion = 10.0 * ev
    es = e_ion - ry / 2.0 ** 2
    ep = e_ion - ry / 2.3 ** 2
    s, a = lines.vdw_damping(0, e_ion, es, ep, 0, 1)
    # the table at (2.0, 2.3): sigma and alpha on the grid, with the unit factor
    import re
    src = open(__import__("os").path.join(__import__("os").path.dirname(abi.__file__), "csrc", "abo_tables.h")).read()

    def table(name):
        body = src[src.index("kAbo_sp_%s[" % name):]
        body = body[body.index("{") + 1:body.index("}")]
        return np.array([float(x) for x in re.findall(r"[-+]?\d*\.?\d+", body)]).reshape(21, 18)
    sig0, alp0 = table("sigma")[10, 10], table("alpha")[10, 10]
    assert abs(a - alp0) < 1e-12
    assert abs(s / (sig0 * 5.62e-17 * (4.0 / pi) ** (alp0 / 2.0) * gamma(2.0 - alp0 / 2.0)) - 1.0) < 1e-12
    # ions take Unsold's value: alpha 0.4, sigma from C6
    s, a = lines.vdw_damping(1, 11.87 * ev, 1.7 * ev, 3.15 * ev, 2, 1)
    nl, nu = np.sqrt(4 * ry / (10.17 * ev)), np.sqrt(4 * ry / (8.72 * ev))
    r2 = lambda n: n * n / 2.0 / 4.0 * (5 * n * n + 1.0)
    assert a == 0.4 and abs(s / (17.0 * (1.41 * 6.46e-34 * (r2(nu) - r2(nl))) ** 0.4 / 251.0) - 1.0) < 1e-12


def test_radiative_damping_sums_both_levels():
    A = np.zeros((4, 4))
    A[1, 0], A[2, 0], A[2, 1], A[3, 0], A[3, 1], A[3, 2] = 1.0, 2.0, 3.0, 4.0, 5.0, 6.0
    assert lines.radiative_damping(A, 1, 3) == 4.0 + 5.0 + 6.0 + 1.0
    assert lines.radiative_damping(A, 0, 2) == 5.0
