"""The oracle's equation-of-state pass (chemical equilibrium filling a column
state, and the response path's thermodynamic derivatives) against the
reference build's own records of the cfg0 column:
  * the synthesis path's state after popsetup and compute_op_referent
    (tests/golden/cfg0_synth.npz): Ne, neutral hydrogen, referent opacity,
    Thomson + bound-free background and the five lines' level populations;
  * the response path's state and derivatives (tests/golden/cfg0_resp.npz):
    T raised by delta_T/2, Ne and neutral hydrogen re-solved there, level
    populations and referent opacity kept, the background at that state, and
    d/dT, d/dNt of Ne, neutral hydrogen, the referent opacity
    (compute_op_referent_derivative, ne_lte_derivatives) and of every line
    level (compute_lte_population_responses_analytical).
And, independently of the reference, that the derivatives are what they
claim: a central difference of the oracle's own chemical equilibrium over a
wider step agrees to the truncation error.  CPU only."""
import os

import numpy as np

import _golden
import _oracle
from snapi_amd import abi, workloads


def setup(name):
    d = _golden.chemeq()
    dd = {k: v for k, v in d.items() if k != "table"}
    tb = abi.make_chem_table(dd, referent=0)
    g = _golden.load(name + ".npz")
    col = workloads.column_from_dump(os.path.join(_golden.HERE, name + ".npz"))
    smap = abi.make_state_map(tb, dd, g["lines"], g["bfsig"], col["ray"].lam_mid)
    return tb, smap, g, col


def rel(a, b):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


FIELDS = [("NE", abi.F_NE), ("NHI", abi.F_NHI), ("OPREF", abi.F_OPREF), ("BGOP", abi.F_BGOP),
          ("BGEM", abi.F_BGEM)]


def test_synthesis_state_matches_reference():
    tb, smap, g, col = setup("cfg0_synth")
    ref = col["state"]
    inp = np.zeros_like(ref)
    inp[abi.F_T], inp[abi.F_NT] = ref[abi.F_T], ref[abi.F_NT]
    st, _, _ = _oracle.eos_state(tb, smap, inp[None])
    errs = {n: rel(st[0, f], ref[f]) for n, f in FIELDS}
    errs["pops"] = rel(st[0, abi.NF_BASE:], ref[abi.NF_BASE:])
    print("synthesis state vs reference:", " ".join("%s %.1e" % kv for kv in errs.items()))
    assert max(errs.values()) < 1e-12


def test_response_state_and_derivatives_match_reference():
    tb, smap, g, col = setup("cfg0_resp")
    _, _, _, c0 = setup("cfg0_synth")
    ref = col["state"]
    inp = np.zeros_like(ref)
    inp[abi.F_T], inp[abi.F_NT] = c0["state"][abi.F_T], c0["state"][abi.F_NT]
    st, resp, ds = _oracle.eos_state(tb, smap, inp[None])
    errs = {n: rel(resp[0, f], ref[f]) for n, f in FIELDS + [("T", abi.F_T), ("NT", abi.F_NT)]}
    errs["pops"] = rel(resp[0, abi.NF_BASE:], ref[abi.NF_BASE:])
    print("response state vs reference:", " ".join("%s %.1e" % kv for kv in errs.items()))
    assert max(errs.values()) < 1e-12
    th = g["thermo"]  # z, Nt, then per p (T, Nt): dNe, dNHI, dopref
    derr = {}
    for p in range(2):
        for k, (n, f) in enumerate((("NE", abi.F_NE), ("NHI", abi.F_NHI), ("OPREF", abi.F_OPREF))):
            derr["d%s/d%s" % (n, "TN"[p])] = rel(ds[0, p, f], th[:, 2 + 3 * p + k])
    dp = g["dpop"]  # s z lo up depth, dlow/dT dup/dT dlow/dNt dup/dNt
    worst = 0.0
    for i, (s, z, lo, up, tr) in enumerate(g["lines"]):
        rows = dp[(dp[:, 0] == s) & (dp[:, 1] == z) & (dp[:, 2] == lo) & (dp[:, 3] == up)]
        rows = rows[np.argsort(rows[:, 4])]
        for p in range(2):
            for u in range(2):
                worst = max(worst, rel(ds[0, p, abi.NF_BASE + 2 * i + u], rows[:, 5 + 2 * p + u]))
    derr["dpops"] = worst
    print("derivatives vs reference:", " ".join("%s %.1e" % kv for kv in derr.items()))
    # forward differences over 5 K of solutions converged to 1e-4 relative:
    # the reference's own values carry its convergence noise, equal here
    # because both sides take the same Newton steps
    assert max(derr.values()) < 1e-8


def test_derivatives_are_derivatives():
    """Each derivative against a central difference over a smaller step of
    the oracle's own equilibrium, at the temperature it belongs to: the
    referent opacity's at T (central over 5 K), Ne's and neutral hydrogen's
    in T at T + 2.5 K (the reference's forward difference over 5 K), every Nt
    derivative at T + 2.5 K, where ne_lte_derivatives takes it.  The level
    responses mix both (stage responses and Boltzmann factor at T + 2.5 K).
    The bars are the truncation error of the reference's steps on this
    column; a sign, a missing factor or a wrong level fails them."""
    tb, smap, g, col = setup("cfg0_synth")
    ref = col["state"]
    inp = np.zeros_like(ref)
    inp[abi.F_T], inp[abi.F_NT] = ref[abi.F_T], ref[abi.F_NT]
    _, _, ds = _oracle.eos_state(tb, smap, inp[None])

    def central(dT, dNt, T_shift):
        a, b = inp.copy(), inp.copy()
        a[abi.F_T] += T_shift + dT
        b[abi.F_T] += T_shift - dT
        a[abi.F_NT] += dNt
        b[abi.F_NT] -= dNt
        sa, _, _ = _oracle.eos_state(tb, smap, a[None])
        sb, _, _ = _oracle.eos_state(tb, smap, b[None])
        return (sa[0] - sb[0]) / (2 * (dT if dT != 0.0 else dNt))

    cT0 = central(0.5, 0.0, 0.0)
    cT25 = central(0.5, 0.0, 2.5)
    cN25 = central(0.0, 1e-4 * inp[abi.F_NT], 2.5)
    pops = list(range(abi.NF_BASE, ref.shape[0]))
    checks = [("dOPREF/dT", 0, [abi.F_OPREF], cT0, 1e-5), ("dNe,dNHI/dT", 0, [abi.F_NE, abi.F_NHI], cT25, 1e-5),
              ("dpops/dT", 0, pops, cT25, 1e-3), ("dNe,dNHI,dpops/dNt", 1, [abi.F_NE, abi.F_NHI] + pops, cN25, 1e-6),
              ("dOPREF/dNt", 1, [abi.F_OPREF], central(0.0, 1e-4 * inp[abi.F_NT], 0.0), 1e-6)]
    for name, p, fields, fd, bar in checks:
        e = max(np.max(np.abs(ds[0, p, f] - fd[f]) / np.max(np.abs(fd[f]))) for f in fields)
        print("%s vs central difference: %.2e (bar %.0e)" % (name, e, bar))
        assert e < bar, name
