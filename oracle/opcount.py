"""Algorithmic operation count of one lambda-point of the polarized
synthesis -- MEASUREMENT INFRASTRUCTURE ONLY.

Builds oracle/_ref/opcount (oracle/opcount.cc: the oracle's C synthesis
compiled with an op-counting scalar), runs it on a sample of columns of a
synthesis workload and writes the per-lambda-point count.  The count is the
roofline's algorithmic work: the synthesis kernel is timed, the count is
not re-derived at bench time but read from the table this script writes.

    python oracle/opcount.py            # writes profiles/synth_flops.json
    python oracle/opcount.py chem       # writes profiles/chem_flops.json
    python oracle/opcount.py ali        # writes profiles/ali_flops.json

Workloads (the ones the synthesis roofline is quoted on):
  cube512: the 512 x 512 x 71 cube of bench.py (seed 1, thermodynamics
           solved: hydrostatic equilibrium and the equation of state, float32
           state), 301 wavelengths; bench.py's 64-column sample (numpy
           default_rng(5).choice(262144, 64, replace=False)), built by the
           oracle's restatement of the same chain.
  cfg0:    the reference's one-column case (its depths and wavelength grid).

Chemistry (oracle/opcount_chem.cc, the oracle's chemical equilibrium and
hydrostatic march with the counting scalar):
  chemeq:       bench.py's chemeq workload (65,536 columns of 61 depths),
                its first 64 columns: operations per depth point.
  hequilibrium: bench.py's hydrostatic-equilibrium workload (16,384 columns
                of 61 depths), its first 64 columns: operations per column
                (the march and the final pass) and per chemical
                equilibrium solved.

ALI (oracle/opcount_ali.cc, the oracle's accumulation pass with the counting
scalar):
  ali:          bench.py's ALI workload (64 x 64 columns of 81 depths, 220
                wavelengths, 10 rays) at its starting populations, its first
                8 columns: operations per formal solution (column x ray x
                wavelength, with its opacity over the depths and its share
                of the atom::add sums).
"""
import json
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
BIN = os.path.join(HERE, "_ref", "opcount")


def build():
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    src = os.path.join(HERE, "opcount.cc")
    deps = [src, os.path.join(HERE, "snapi_oracle.c"), os.path.join(HERE, "snapi_oracle.h")]
    if os.path.exists(BIN) and all(os.path.getmtime(BIN) >= os.path.getmtime(d) for d in deps):
        return BIN
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wno-unknown-pragmas", "-o", BIN, src],
                   check=True)
    return BIN


def write_input(path, table, ray, states, lam):
    states = np.ascontiguousarray(states, dtype=np.float64)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    ncol, nf, nd = states.shape
    with open(path, "wb") as f:
        f.write(struct.pack("<9i", table.nline, table.ncomp_total, ncol, nd, nf, lam.size, ray.tau_grid, ray.hminus,
                            ray.hminus_mid))
        f.write(struct.pack("<2d", ray.theta, ray.lam_mid))
        for i in range(table.nline):
            L = table.line[i]
            f.write(struct.pack("<8d", L.lam0, L.aul, L.blu, L.bul, L.grad, L.sigma, L.alpha, L.mass))
            f.write(struct.pack("<5i", L.split, L.ncomp[0], L.ncomp[1], L.ncomp[2], L.comp0))
        for c in range(table.ncomp_total):
            f.write(struct.pack("<2d", table.comp[c].shift, table.comp[c].strength))
        f.write(states.tobytes())
        f.write(lam.tobytes())


def count(table, ray, states, lam):
    """Per column: (counted operations, raw operations, stokes [nlambda][4])."""
    exe = build()
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "in.bin")
        write_input(p, table, ray, states, lam)
        out = subprocess.run([exe, p], capture_output=True, text=True, check=True).stdout.split("\n")
    res = []
    for line in out:
        if not line.strip():
            continue
        w = line.split()
        st = np.array([float.fromhex(x) for x in w[3:]]).reshape(-1, 4)
        res.append((int(w[1]), int(w[2]), st))
    return res


def cube512_sample():
    sys.path.insert(0, REPO)
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from snapi_amd import workloads
    import _oracle
    t = workloads.template_cfg1()
    samp = np.random.default_rng(5).choice(512 * 512, 64, replace=False)
    st = _oracle.cube_state_eos(t, 512, 512, 1, samp).astype(np.float32)
    return t["table"], t["ray"], st.astype(np.float64), t["lam_vac"]


def cfg0_case():
    sys.path.insert(0, REPO)
    from snapi_amd import workloads
    t = workloads.cfg0_column()
    return t["table"], t["ray"], t["state"][None].astype(np.float64), t["lam_vac"]


CHEM_BIN = os.path.join(HERE, "_ref", "opcount_chem")


def build_chem():
    os.makedirs(os.path.dirname(CHEM_BIN), exist_ok=True)
    src = os.path.join(HERE, "opcount_chem.cc")
    deps = [src, os.path.join(HERE, "opcount.cc"), os.path.join(HERE, "snapi_oracle_chem.c"),
            os.path.join(HERE, "snapi_oracle.h")]
    if os.path.exists(CHEM_BIN) and all(os.path.getmtime(CHEM_BIN) >= os.path.getmtime(d) for d in deps):
        return CHEM_BIN
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wno-unknown-pragmas", "-o", CHEM_BIN, src],
                   check=True)
    return CHEM_BIN


def count_chem(table, T, X, mode):
    """Per column: (counted operations, raw operations, equilibria, Ne
    (mode 0) or Nt (mode 1) [nd]).  mode 0: X = Nt, chemical equilibrium of
    every point; mode 1: X = log tau, the hydrostatic march."""
    import ctypes
    exe = build_chem()
    T = np.ascontiguousarray(T, dtype=np.float64)
    X = np.ascontiguousarray(X, dtype=np.float64)
    ncol, nd = T.shape
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "in.bin")
        with open(p, "wb") as f:
            f.write(struct.pack("<3i", mode, ncol, nd))
            f.write(ctypes.string_at(ctypes.addressof(table), ctypes.sizeof(table)))
            f.write(T.tobytes())
            f.write(X.tobytes())
        out = subprocess.run([exe, p], capture_output=True, text=True, check=True).stdout.split("\n")
    res = []
    for line in out:
        if not line.strip():
            continue
        w = line.split()
        res.append((int(w[1]), int(w[2]), int(w[3]), np.array([float.fromhex(x) for x in w[4:]])))
    return res


def chem_main():
    sys.path.insert(0, REPO)
    from snapi_amd import workloads
    table = {}
    p = workloads.chemeq_problem(256, 256)
    res = count_chem(p["table"], p["T"][:64], p["Nt"][:64], 0)
    c = np.array([r[0] for r in res], dtype=np.float64)
    nd = p["T"].shape[1]
    table["chemeq"] = {"flops_per_point": float(c.sum() / (64 * nd)),
                       "raw_per_point": float(sum(r[1] for r in res) / (64 * nd)), "columns": 64, "depths": nd,
                       "counted_per_column": [int(x) for x in c],
                       "workload": "bench.py chemeq leg (chemeq_problem 256 x 256), first 64 columns"}
    print("chemeq", table["chemeq"], flush=True)
    p = workloads.hequilibrium_problem(128, 128)
    res = count_chem(p["table"], p["T"][:64], p["logtau"][:64], 1)
    c = np.array([r[0] for r in res], dtype=np.float64)
    neq = np.array([r[2] for r in res], dtype=np.float64)
    nd = p["T"].shape[1]
    table["hequilibrium"] = {"flops_per_column": float(c.mean()), "flops_per_equilibrium": float(c.sum() / neq.sum()),
                             "equilibria_per_column": float(neq.mean()),
                             "raw_per_column": float(np.mean([r[1] for r in res])), "columns": 64, "depths": nd,
                             "counted_per_column": [int(x) for x in c],
                             "workload": "bench.py hequilibrium leg (hequilibrium_problem 128 x 128), first 64 "
                                         "columns; march iterations plus the final pass"}
    print("hequilibrium", table["hequilibrium"], flush=True)
    table["method"] = ("oracle/opcount_chem.cc: snapi_oracle_chem.c's chemeq_point and orc_hequilibrium with the "
                       "op-counting double of oracle/opcount.cc; add, sub, mul, div, sqrt, exp, log, log10 = 1, "
                       "pow = 3; each (operation, operands) counted once per column")
    table["command"] = "python oracle/opcount.py chem"
    with open(os.path.join(REPO, "profiles", "chem_flops.json"), "w") as f:
        json.dump(table, f, indent=1)


ALI_BIN = os.path.join(HERE, "_ref", "opcount_ali")


def build_ali():
    os.makedirs(os.path.dirname(ALI_BIN), exist_ok=True)
    src = os.path.join(HERE, "opcount_ali.cc")
    deps = [src, os.path.join(HERE, "opcount.cc"), os.path.join(HERE, "snapi_oracle_ali.c"),
            os.path.join(HERE, "snapi_oracle.c"), os.path.join(HERE, "snapi_oracle.h")]
    if os.path.exists(ALI_BIN) and all(os.path.getmtime(ALI_BIN) >= os.path.getmtime(d) for d in deps):
        return ALI_BIN
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wno-unknown-pragmas", "-o", ALI_BIN, src],
                   check=True)
    return ALI_BIN


def count_ali(p, cols):
    """Per column of `cols` of an ALI problem (workloads.ali_problem): (counted
    operations, raw operations, acc [nd][ntrans][4])."""
    import ctypes
    exe = build_ali()
    g = p["grid"]
    lam, lamw, fudge, sigma, theta, wray = g._keep
    atom = p["atom"]
    cs = p["colstate"]
    nd = cs.shape[2]
    nbg = p["bgop"].shape[1]
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "in.bin")
        with open(path, "wb") as f:
            f.write(struct.pack("<7i", len(cols), nd, g.nlam, g.nray, g.bg_per_ray, g.tau_grid, nbg))
            f.write(struct.pack("<d", g.boundary))
            f.write(ctypes.string_at(ctypes.addressof(atom), ctypes.sizeof(atom)))
            for a in (lam, lamw, fudge, sigma, theta, wray):
                f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
            for c in cols:
                for k in ("colstate", "vproj", "gamma", "bgop", "bgem", "pops"):
                    f.write(np.ascontiguousarray(p[k][c], dtype=np.float64).tobytes())
        out = subprocess.run([exe, path], capture_output=True, text=True, check=True).stdout.split("\n")
    res = []
    nt = atom.ntrans
    for line in out:
        if not line.strip():
            continue
        w = line.split()
        res.append((int(w[1]), int(w[2]), np.array([float.fromhex(x) for x in w[3:]]).reshape(nd, nt, 4)))
    return res


def ali_main():
    sys.path.insert(0, REPO)
    from snapi_amd import workloads
    p = workloads.ali_problem(64, 64)
    ncs = 8
    res = count_ali(p, list(range(ncs)))
    c = np.array([r[0] for r in res], dtype=np.float64)
    nsolve = p["nlam"] * p["nray"]
    table = {"ali": {"flops_per_formal_solve": float(c.mean() / nsolve),
                     "raw_per_formal_solve": float(np.mean([r[1] for r in res]) / nsolve),
                     "columns": ncs, "depths": int(p["colstate"].shape[2]), "wavelengths": int(p["nlam"]),
                     "rays": int(p["nray"]), "counted_per_column": [int(x) for x in c],
                     "workload": "bench.py ALI leg (ali_problem 64 x 64) at its starting populations, first 8 "
                                 "columns; one accumulation pass"},
             "method": ("oracle/opcount_ali.cc: snapi_oracle_ali.c's orc_ali_accumulate (with snapi_oracle.c's "
                        "Voigt function and scalar formal solution) with the op-counting double of "
                        "oracle/opcount.cc; add, sub, mul, div, sqrt, exp, log, log10, sin, cos = 1, pow = 3; each "
                        "(operation, operands) counted once per column"),
             "command": "python oracle/opcount.py ali"}
    print(table["ali"], flush=True)
    with open(os.path.join(REPO, "profiles", "ali_flops.json"), "w") as f:
        json.dump(table, f, indent=1)


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "chem":
        return chem_main()
    if len(sys.argv) > 1 and sys.argv[1] == "ali":
        return ali_main()
    table = {}
    for name, make in (("cfg0", cfg0_case), ("cube512", cube512_sample)):
        tb, ray, st, lam = make()
        res = count(tb, ray, st, lam)
        c = np.array([r[0] for r in res], dtype=np.float64)
        raw = np.array([r[1] for r in res], dtype=np.float64)
        table[name] = {"flops_per_lambda_point": float(c.mean() / lam.size),
                       "raw_per_lambda_point": float(raw.mean() / lam.size),
                       "columns": int(st.shape[0]), "depths": int(st.shape[2]), "wavelengths": int(lam.size),
                       "per_column_min": float(c.min() / lam.size), "per_column_max": float(c.max() / lam.size),
                       "counted_per_column": [int(x) for x in c]}
        print(name, table[name], flush=True)
    table["method"] = ("oracle/opcount.cc: snapi_oracle.c's orc_synth with an op-counting double; add, sub, mul, "
                       "div, sqrt, exp, log, log10, sin, cos = 1, pow = 3, FMA = 2; each (operation, operands) "
                       "counted once per column")
    table["command"] = "python oracle/opcount.py"
    with open(os.path.join(REPO, "profiles", "synth_flops.json"), "w") as f:
        json.dump(table, f, indent=1)


if __name__ == "__main__":
    main()
