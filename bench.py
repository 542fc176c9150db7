"""Throughput of the HIP hot path on one MI355X.

Workload (one "step"): polarized LTE Stokes synthesis of the 512 x 512
column cube, 71 depths, 301 wavelengths, mu = 1, two Fe I Zeeman lines plus
the hydrogen lines and continuum of the reference atom set, in the
mixed-precision path; inputs resident in HBM before the timed region.  The
cube's thermodynamics are solved on the GPU before timing: hydrostatic
equilibrium of the drawn temperatures, the 12 x 512 x 512 x 71 float32 model
cube, its column states, the equation of state.
Value: lambda-point formal solves per second (column x wavelength x ray).

Also reported: the FP64 GPU path on the same cube, the accuracy of the
mixed path against the FP64 CPU oracle on a sample of columns, the roofline
fraction of the synthesis kernel, and the oracle's own throughput on the
host cores (cpu_baseline, FP64 port, OpenMP).

Second leg ("lm"): batched Levenberg-Marquardt inversion of 256 x 256 pixels
drawn as the cube (observations: FP64 GPU synthesis plus Gaussian noise of
1e-3 Ic; per-pixel masks keeping 120-301 wavelengths; 7 kinematic/magnetic
nodes started +/-20 % from truth; at most 30 iterations), mixed precision
with tangent-sweep Jacobians, against the FP64 path with central-difference
Jacobians on 8192 of the pixels (and the mixed adjoint-sweep Jacobian on the
same 8192): LM pixel-iterations per second of each.

Third leg ("ali"): NLTE populations of a Ca II 5-level + continuum atom by
ALI in 64 x 64 plane-parallel columns of 81 depths (chromospheric rise to
8000 K, +/-10 % perturbations), 220 wavelengths, 5 Gauss-Radau directions per
hemisphere, iterated to a relative change below 1e-4: formal solves per second
inside the iteration, mixed precision, against the FP64 GPU path on 512 of the
columns and the FP64 oracle on the host cores.

python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "lambda-point formal solves/s + LM pixel-iter/s at <=1e-4 Ic vs FP64 oracle"
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6


def log(msg):
    sys.stderr.write("[bench] %s\n" % msg)
    sys.stderr.flush()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--cpu-sample", type=int, default=4096, help="columns timed on the CPU oracle")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--lm-side", type=int, default=256, help="inversion leg: side of the pixel square")
    ap.add_argument("--lm-fd-pixels", type=int, default=8192, help="pixels of the FP64 difference-Jacobian run")
    ap.add_argument("--tnode-side", type=int, default=256, help="14-node inversion: side of the pixel square")
    ap.add_argument("--tnode-legs", default="mixed_tangent_analytic,mixed_tangent_analytic_first,fp64_differences",
                    help="14-node inversion runs, comma-separated: mixed_tangent_analytic, mixed_adjoint, "
                         "mixed_adjoint_analytic, fp64_differences")
    ap.add_argument("--tnode-fd-iters", type=int, default=2, help="iterations of the FP64 difference-Jacobian run")
    ap.add_argument("--no-lm", action="store_true")
    ap.add_argument("--ali-side", type=int, default=64, help="ALI leg: side of the column square")
    ap.add_argument("--no-ali", action="store_true")
    ap.add_argument("--chem-side", type=int, default=256, help="chemeq leg: side of the column square")
    ap.add_argument("--no-chem", action="store_true")
    ap.add_argument("--he-side", type=int, default=128, help="hydrostatic-equilibrium leg: side of the column square")
    ap.add_argument("--no-he", action="store_true")
    ap.add_argument("--nlteresp-cols", type=int, default=64, help="NLTE-response leg: columns of the formal pass")
    ap.add_argument("--nlteresp-systems", type=int, default=64, help="NLTE-response leg: systems solved")
    ap.add_argument("--no-nlteresp", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the hot path is per column and shards by column tiles; "
              "runs on more than one GPU are not part of this build.", file=sys.stderr)
        sys.exit(2)

    import torch
    from snapi_amd import abi, api, autotune, flops, workloads

    if not torch.cuda.is_available():
        sys.exit("no GPU visible")
    dev = torch.device("cuda:0")
    tmpl = workloads.template_cfg1()
    table, ray, lam = tmpl["table"], tmpl["ray"], tmpl["lam_vac"]
    nx, ny = args.nx, args.ny
    ncol = nx * ny
    nlam = len(lam)
    t0 = time.time()
    # the model cube's thermodynamics solved on the GPU: hydrostatic
    # equilibrium, the 12 x nx x ny x 71 float32 model cube, its column
    # states for the mu = 1 ray, the equation of state
    st64_build, _cube = workloads.cube_state_eos_dev(tmpl, nx, ny, seed=1)
    st_d = st64_build.float().contiguous()
    del st64_build, _cube
    torch.cuda.synchronize()
    nf, nd = st_d.shape[1], st_d.shape[2]
    log("cube 12x%dx%dx%d (float32, hydrostatic equilibrium + equation of state) -> %d column states of %d fields, "
        "nlambda=%d, built in %.1fs" % (nx, ny, nd, ncol, nf, nlam, time.time() - t0))
    lam_d = torch.from_numpy(np.ascontiguousarray(lam)).to(dev)
    out_d = torch.empty((ncol, nlam, 4), dtype=torch.float32, device=dev)
    stream = torch.cuda.Stream(device=dev)

    def run_steps(n, state_t, out_t, prec):
        for _ in range(n):
            api.synth_stokes_dev(table, ray, state_t, lam_d, out_t, prec, stream)

    def timed(state_t, out_t, prec, steps, warmup):
        with torch.cuda.stream(stream):
            run_steps(warmup, state_t, out_t, prec)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        with torch.cuda.stream(stream):
            run_steps(steps, state_t, out_t, prec)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps  # ms per step = per launch

    # the autotuner times the kernel's launch shapes on the first call of a
    # problem shape: done here on a slice of the cube, outside any step
    ntune = min(ncol, 16384)
    with torch.cuda.stream(stream):
        api.synth_stokes_dev(table, ray, st_d[:ntune], lam_d, out_d[:ntune], abi.SNAPI_MIXED, stream)
    torch.cuda.synchronize()
    ms = timed(st_d, out_d, abi.SNAPI_MIXED, args.steps, args.warmup)
    value = ncol * nlam / (ms * 1e-3)
    log("mixed: %.3f ms/step, %.4g lambda-point solves/s" % (ms, value))

    # algorithmic flops per lambda point: the frozen operation count of this
    # workload (profiles/synth_flops.json, counted through the FP64
    # restatement on a 64-column sample of this cube); the operation model of
    # the kernel (snapi_amd/flops.py) on the same sample is reported beside it
    rng = np.random.default_rng(5)
    samp = rng.choice(ncol, 64, replace=False)
    samp_t = torch.from_numpy(np.sort(samp)).to(dev)
    fpl_model = flops.flops_per_lambda_point(table, ray, st_d[samp_t].cpu().numpy().astype(np.float64), lam)
    fpl, fpl_src = fpl_model, "model"
    ff = os.path.join(REPO, "profiles", "synth_flops.json")
    if os.path.exists(ff):
        fc = json.load(open(ff)).get("cube512")
        if fc and (nx, ny, nd, nlam) == (512, 512, fc["depths"], fc["wavelengths"]):
            fpl, fpl_src = fc["flops_per_lambda_point"], "counted"
    launch_flops = fpl * ncol * nlam
    achieved = launch_flops / (ms * 1e-3) / 1e12
    # HBM bytes per launch from the FETCH_SIZE / WRITE_SIZE counter runs of this
    # kernel on this workload (scaled by a plain copy's reading)
    traffic = None
    tf = os.path.join(REPO, "profiles", "synth_traffic.json")
    if os.path.exists(tf):
        t = json.load(open(tf))
        if (t["columns"], t["depths"], t["wavelengths"]) == (ncol, nd, nlam):
            traffic = t["bytes_per_launch"]
    roof = {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": traffic,
            "algorithmic_bytes": int(ncol * nf * nd * 4 + ncol * nlam * 16),
            "kernel": "synth_kernel_f", "flops_per_lambda_point": round(fpl, 1), "flops_source": fpl_src,
            "flops_per_lambda_point_model": round(fpl_model, 1),
            "ms_per_launch": round(ms, 4)}

    fp64 = None
    if not args.no_fp64:
        st64_d = st_d.to(torch.float64)
        out64_d = torch.empty((ncol, nlam, 4), dtype=torch.float64, device=dev)
        with torch.cuda.stream(stream):
            api.synth_stokes_dev(table, ray, st64_d[:ntune], lam_d, out64_d[:ntune], abi.SNAPI_FP64, stream)
        torch.cuda.synchronize()
        ms64 = timed(st64_d, out64_d, abi.SNAPI_FP64, max(1, args.steps // 2), 1)
        a64 = launch_flops / (ms64 * 1e-3) / 1e12
        fp64 = {"value": ncol * nlam / (ms64 * 1e-3), "ms_per_step": round(ms64, 4),
                "achieved_tflops": round(a64, 3), "frac_fp64_peak": round(a64 / FP64_PEAK_TFLOPS, 4)}
        log("fp64: %.3f ms/step" % ms64)
        del st64_d, out64_d

    gpu_out = out_d.cpu().numpy()
    cpu = None
    acc = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(REPO, "tests"))
        import _oracle  # the FP64 CPU restatement: baseline and checker only
        nthreads = max(1, min(16, os.cpu_count() or 1))
        ns = min(args.cpu_sample, ncol)
        idx = np.arange(ns)
        st64 = st_d[:ns].cpu().numpy().astype(np.float64)
        t0 = time.time()
        ref, rc = _oracle.synth(table, ray, st64, lam, nthreads=nthreads)
        dt = time.time() - t0
        cpu = {"value": ns * nlam / dt, "unit": "lambda-point formal solves/s", "cores": nthreads, "kind": "port",
               "sample": "%d columns x %d wavelengths of the same cube (FP64 oracle, OpenMP)" % (ns, nlam),
               "seconds": round(dt, 2)}
        ic = float(np.max(ref[..., 0]))
        err = np.abs(gpu_out[idx].astype(np.float64) - ref) / ic
        acc = {"max_err_ic": float(err.max()), "mean_err_ic": float(err.mean()), "columns": int(ns),
               "reference": "FP64 CPU oracle"}
        log("cpu oracle: %.4g lambda-points/s on %d threads; mixed max err %.3e Ic" % (cpu["value"], nthreads,
                                                                                      err.max()))

    lm = None
    if not args.no_lm:
        lm = lm_leg(torch, dev, tmpl, lam_d, stream, args.lm_side, args.lm_fd_pixels, args.tnode_side,
                    tuple(x for x in args.tnode_legs.split(",") if x), args.tnode_fd_iters)

    alires = None
    if not args.no_ali:
        alires = ali_leg(torch, dev, stream, args.ali_side, not args.no_cpu)

    chem = None
    if not args.no_chem:
        chem = chemeq_leg(torch, dev, stream, args.chem_side, not args.no_cpu)

    he = None
    if not args.no_he:
        he = hequilibrium_leg(torch, dev, stream, args.he_side, not args.no_cpu)

    nresp = None
    if not args.no_nlteresp:
        nresp = nlteresp_leg(torch, dev, stream, args.nlteresp_cols, args.nlteresp_systems, not args.no_cpu)

    if lm is not None and he is not None and he.get("roofline"):
        # the inversion's dominant kernel is the hydrostatic march (the
        # kernel profile of the 14-node leg); its roofline as the
        # hequilibrium leg of this run measures it
        lm["roofline"] = dict(he["roofline"], kernel="hequilibrium_kernel",
                              note="dominant kernel of the 14-node inversion (37 % of its GPU time, with the "
                                   "equation-of-state pass 35 % and the tangent sweeps 20 %); measured in the "
                                   "hequilibrium leg of this run")
    line = {"metric": METRIC, "value": value, "unit": "lambda-point formal solves/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
            "config": {"workload": "LTE polarized Stokes synthesis, cube %dx%d columns x %d depths, %d wavelengths, "
                                   "mu=1, Fe I 6301.5/6302.5 + H lines + H- continuum, mixed FP32/FP64"
                                   % (nx, ny, nd, nlam),
                       "columns": ncol, "depths": nd, "wavelengths": nlam, "lines": table.nline},
            "roofline": roof, "autotune": autotune.report(),
            "cpu_baseline": cpu, "accuracy": acc, "fp64_gpu": fp64, "lm": lm, "ali": alires,
            "chemeq": chem, "hequilibrium": he, "nlte_responses": nresp}
    print(json.dumps(line))


def chem_roofline(key, per, units, ms, kernel):
    """FP64 VALU roofline of a chemistry kernel: the frozen operation count
    (profiles/chem_flops.json, counted through the FP64 restatement of the
    same workload) times the units of one call over the call's time."""
    ff = os.path.join(REPO, "profiles", "chem_flops.json")
    if not os.path.exists(ff):
        return None
    fpu = json.load(open(ff))[key][per]
    achieved = fpu * units / (ms * 1e-3) / 1e12
    return {"bound": "valu", "achieved": round(achieved, 4), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 5), "traffic": None, "kernel": kernel,
            "flops_per_unit": round(fpu, 1), "unit_counted": per, "flops_source": "counted", "ms_per_call": ms}


def chemeq_leg(torch, dev, stream, side, with_cpu):
    from snapi_amd import abi, chemeq, workloads
    t0 = time.time()
    p = workloads.chemeq_problem(side, side)
    tb = p["table"]
    ncol, nd = p["T"].shape
    npt = ncol * nd
    nst = abi.chem_stage_count(tb)
    log("chemeq workload: %d columns x %d depths, %d elements (%d stages), %d molecules, %d levels, built in %.1fs"
        % (ncol, nd, tb.nelement, nst, tb.nmolecule, tb.nlevel, time.time() - t0))
    T = torch.from_numpy(p["T"]).to(dev)
    Nt = torch.from_numpy(p["Nt"]).to(dev)
    ne = torch.empty((ncol, nd), dtype=torch.float64, device=dev)
    st = torch.empty((ncol, nst, nd), dtype=torch.float64, device=dev)
    mo = torch.empty((ncol, tb.nmolecule, nd), dtype=torch.float64, device=dev)
    lv = torch.empty((ncol, tb.nlevel, nd), dtype=torch.float64, device=dev)
    it = torch.empty((ncol, nd), dtype=torch.int32, device=dev)
    kr = torch.empty((ncol, nd), dtype=torch.float64, device=dev)

    def call():
        chemeq.chemeq_dev(tb, T, Nt, ne, st, mo, lv, it, stream=stream, opref_t=kr)

    call()  # kernel loading, not timed
    torch.cuda.synchronize()
    reps = 5
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        call()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    itn = it.cpu().numpy()
    res = {"points": npt, "ms_per_call": round(ms, 3), "points_per_s": npt / (ms * 1e-3),
           "mean_newton_steps": float(itn.mean()), "max_newton_steps": int(itn.max()),
           "singular": int(np.sum(itn < 0))}
    log("chemeq: %d points in %.2f ms = %.4g points/s; Newton steps mean %.2f max %d"
        % (npt, ms, res["points_per_s"], itn.mean(), itn.max()))
    roof = chem_roofline("chemeq", "flops_per_point", npt, ms, "chemeq_kernel")
    cpu = None
    if with_cpu:
        sys.path.insert(0, os.path.join(REPO, "tests"))
        import _oracle  # the FP64 CPU restatement: baseline and checker only
        nthreads = max(1, min(16, os.cpu_count() or 1))
        nc = min(ncol, 2048)
        t1 = time.time()
        ne_c, st_c, mo_c, lv_c, it_c, kr_c = _oracle.chemeq(tb, p["T"][:nc], p["Nt"][:nc], nthreads=nthreads,
                                                            with_opref=True)
        dt = time.time() - t1
        cpu = {"value": nc * nd / dt, "unit": "depth points/s", "cores": nthreads, "kind": "port",
               "sample": "%d columns x %d depths of the same workload (FP64 oracle, OpenMP)" % (nc, nd),
               "seconds": round(dt, 2)}
        g_ne = ne[:nc].cpu().numpy()
        res["ne_rel_err_vs_oracle"] = float(np.max(np.abs(g_ne / ne_c - 1.0)))
        res["steps_equal_fraction"] = float(np.mean(itn[:nc] == it_c))
        res["opref_rel_err_vs_oracle"] = float(np.max(np.abs(kr[:nc].cpu().numpy() / kr_c - 1.0)))
        log("chemeq cpu oracle: %.4g points/s on %d threads; GPU vs oracle Ne max rel %.2e, steps equal %.4f"
            % (cpu["value"], nthreads, res["ne_rel_err_vs_oracle"], res["steps_equal_fraction"]))
    return {"workload": "chemical equilibrium + LTE populations + referent continuum opacity, %d columns x %d depths, %d elements, %d molecules, "
                        "%d levels, FP64" % (ncol, nd, tb.nelement, tb.nmolecule, tb.nlevel),
            "value": res["points_per_s"], "unit": "depth points/s", "fp64": res, "roofline": roof,
            "cpu_baseline": cpu}


def hequilibrium_leg(torch, dev, stream, side, with_cpu):
    from snapi_amd import abi, chemeq, workloads
    p = workloads.hequilibrium_problem(side, side)
    tb = p["table"]
    ncol, nd = p["T"].shape
    log("hequilibrium workload: %d columns x %d depths" % (ncol, nd))
    T = torch.from_numpy(p["T"]).to(dev)
    lt = torch.from_numpy(p["logtau"]).to(dev)
    f = lambda: torch.empty((ncol, nd), dtype=torch.float64, device=dev)
    Nt, ne, kr, rho = f(), f(), f(), f()
    it = torch.empty((ncol, nd), dtype=torch.int32, device=dev)

    def call():
        chemeq.hequilibrium_dev(tb, T, lt, Nt, ne, kr, rho, iterations_t=it, stream=stream)

    call()  # kernel loading, not timed
    torch.cuda.synchronize()
    reps = 3
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        call()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    itn = it.cpu().numpy()
    solves = int(itn.sum()) + ncol * nd  # march iterations plus the final pass
    res = {"columns": ncol, "depths": nd, "ms_per_call": round(ms, 3), "columns_per_s": ncol / (ms * 1e-3),
           "chemeq_solves": solves, "solves_per_s": solves / (ms * 1e-3),
           "mean_iterations": float(itn.mean()), "max_iterations": int(itn.max())}
    log("hequilibrium: %d columns in %.2f ms = %.4g columns/s (%.4g chemical equilibria/s); iterations mean %.2f max %d"
        % (ncol, ms, res["columns_per_s"], res["solves_per_s"], itn.mean(), itn.max()))
    roof = chem_roofline("hequilibrium", "flops_per_equilibrium", solves, ms,
                         "hequilibrium_kernel + chemeq_kernel (the final pass)")
    cpu = None
    if with_cpu:
        sys.path.insert(0, os.path.join(REPO, "tests"))
        import _oracle  # the FP64 CPU restatement: baseline and checker only
        nthreads = max(1, min(16, os.cpu_count() or 1))
        nc = min(ncol, 2048)
        t1 = time.time()
        Nt_c, ne_c, kr_c, rho_c, it_c = _oracle.hequilibrium(tb, p["T"][:nc], p["logtau"][:nc], nthreads=nthreads)
        dt = time.time() - t1
        cpu = {"value": nc / dt, "unit": "columns/s", "cores": nthreads, "kind": "port",
               "sample": "%d columns x %d depths of the same workload (FP64 oracle, OpenMP over columns)" % (nc, nd),
               "seconds": round(dt, 2)}
        same = np.all(itn[:nc] == it_c, axis=1)
        g = Nt[:nc].cpu().numpy()
        res["Nt_rel_err_vs_oracle"] = float(np.max(np.abs(g[same] / Nt_c[same] - 1.0))) if same.any() else None
        res["columns_equal_iterations"] = float(same.mean())
        log("hequilibrium cpu oracle: %.4g columns/s on %d threads; GPU vs oracle Nt max rel %s, equal iterations %.4f"
            % (cpu["value"], nthreads, res["Nt_rel_err_vs_oracle"], same.mean()))
    return {"workload": "hydrostatic equilibrium (chemical equilibrium + referent opacity per iteration), "
                        "%d columns x %d depths, FP64" % (ncol, nd),
            "value": res["columns_per_s"], "unit": "columns/s", "fp64": res, "roofline": roof, "cpu_baseline": cpu}


def nlteresp_leg(torch, dev, stream, ncol, nsys, with_cpu):
    """The two numerical kernels of the NLTE population responses on the Ca II
    column recorded from the reference (tests/golden/nlteresp.npz): the
    formal solution with dI/dop and dI/dem for 2 rays x 444 wavelengths per
    column (the reference's NLTE grid), inputs perturbed per column; and the
    366 x 366 response-system solve with its 183 right-hand sides."""
    from snapi_amd import nlteresp
    d = np.load(os.path.join(REPO, "tests", "golden", "nlteresp.npz"))
    nd, npr0 = int(d["nd"]), len(d["prob"])
    rng = np.random.default_rng(12)
    nprob = 888
    pick = np.arange(nprob) % npr0
    theta = torch.from_numpy(np.ascontiguousarray(d["prob"][pick, 2])).to(dev)
    grid = torch.from_numpy(np.repeat(d["grid"][None], ncol, axis=0)).to(dev)
    f = np.exp(rng.normal(0.0, 0.1, (ncol, nprob, nd)))
    op = torch.from_numpy(np.ascontiguousarray(d["op"][pick][None] * f)).to(dev)
    em = torch.from_numpy(np.ascontiguousarray(d["em"][pick][None] * f)).to(dev)
    I = torch.empty((ncol, nprob, nd), dtype=torch.float64, device=dev)
    L = torch.empty_like(I)
    Rop = torch.empty((ncol, nprob, nd, nd), dtype=torch.float64, device=dev)
    Rem = torch.empty_like(Rop)

    def call():
        nlteresp.formal_responses_dev(grid, op, em, theta, I, L, Rop, Rem, stream=stream)

    call()
    torch.cuda.synchronize()
    reps = 5
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        call()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    probs = ncol * nprob
    # HBM bytes per problem: op, em read, I, L written, both nd x nd matrices written
    byt = probs * (4 * nd + 2 * nd * nd) * 8.0
    formal = {"problems": probs, "depths": nd, "ms_per_launch": round(ms, 3), "problems_per_s": probs / (ms * 1e-3),
              "roofline": {"bound": "hbm", "achieved": byt / (ms * 1e-3) / 1e9, "peak": 8000.0, "unit": "GB/s",
                           "frac": byt / (ms * 1e-3) / 8e12, "algorithmic_bytes": byt}}
    log("formal_with_responses: %d problems x %d depths in %.3f ms = %.4g problems/s, %.0f GB/s"
        % (probs, nd, ms, probs / (ms * 1e-3), byt / (ms * 1e-3) / 1e9))
    n = d["M"].shape[0]
    M = d["M"][None] * (1.0 + 1e-3 * rng.standard_normal((nsys, n, n)))
    B = np.repeat(d["beta"].reshape(3 * nd, n)[None], nsys, axis=0)
    nlteresp.response_solve(M[:1], B[:1])  # kernel loading, not timed
    t0 = time.perf_counter()
    X = nlteresp.response_solve(M, B)
    wall = time.perf_counter() - t0
    flop = nsys * (2.0 * n ** 3 / 3.0 + 2.0 * n * n * B.shape[1])
    solve = {"systems": nsys, "n": n, "rhs": int(B.shape[1]), "seconds_with_copies": round(wall, 4),
             "systems_per_s": nsys / wall, "fp64_gflop_per_s": flop / wall / 1e9}
    log("response solve: %d systems of %d x %d with %d right-hand sides in %.3f s (host copies included)"
        % (nsys, n, n, B.shape[1], wall))
    cpu = None
    if with_cpu:
        sys.path.insert(0, os.path.join(REPO, "tests"))
        import _oracle  # the FP64 CPU restatement: baseline and checker only
        t1 = time.time()
        nref = 32
        for p in range(nref):
            _oracle.formal_responses(d["grid"], d["op"][p % npr0], d["em"][p % npr0], d["prob"][p % npr0, 2])
        dt = time.time() - t1
        t2 = time.time()
        Xo = _oracle.response_solve(M[0], B[0])
        dts = time.time() - t2
        cpu = {"value": nref / dt, "unit": "formal_with_responses problems/s", "cores": 1, "kind": "port",
               "sample": "%d problems of the same column (FP64 oracle, one thread); one system solve %.3f s"
                         % (nref, dts), "seconds": round(dt, 3), "solve_seconds": round(dts, 3)}
        solve["rel_err_vs_oracle"] = float(np.max(np.abs(X[0] - Xo)) / np.abs(Xo).max())
        log("nlteresp cpu oracle: %.4g problems/s, one solve %.3f s; GPU solve vs oracle %.2e"
            % (nref / dt, dts, solve["rel_err_vs_oracle"]))
    system = nlte_system_leg(torch, dev, stream, ncol, with_cpu)
    return {"workload": "NLTE population responses, Ca II 5 levels + continuum in the 61-depth column: the whole "
                        "response system (profile norms, formal_with_responses, assembly, rate part, solve) for %d "
                        "columns x %d (ray, wavelength) problems; formal_with_responses alone; %d solves of the "
                        "%d x %d response system" % (ncol, nprob, nsys, n, n),
            "value": system["columns_per_s"], "unit": "NLTE response systems (columns)/s", "system": system,
            "formal": formal, "solve": solve, "cpu_baseline": system["cpu_baseline"], "cpu_baseline_kernels": cpu}


def nlte_system_leg(torch, dev, stream, ncol, with_cpu):
    """snapi_nlte_response_system_dev on ncol perturbed copies of the Ca II
    case recorded from the reference (workloads.nlte_system_problem), with
    the solve, inputs resident on the device; the oracle's assembly of one
    column timed beside it."""
    from snapi_amd import nlteresp, workloads
    atom, c, inp = workloads.nlte_system_problem(ncol)
    nd, nm = c["nd"], c["nmap"]
    n = nd * nm
    tens = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in inp.items()}
    lam_t = torch.from_numpy(c["lam"]).to(dev)
    lamw_t = torch.from_numpy(c["lamw"]).to(dev)
    M_t = torch.empty((ncol, n, n), dtype=torch.float64, device=dev)
    b_t = torch.empty((ncol, 3, nd, n), dtype=torch.float64, device=dev)
    r_t = torch.empty((ncol, 3, n, nd), dtype=torch.float64, device=dev)

    def call():
        nlteresp.response_system_dev(atom, c["theta"], c["wray"], lam_t, lamw_t, tens, M_t, b_t, r_t,
                                     tau_grid=c["tau_grid"], boundary=c["boundary"], stream=stream)

    call()
    torch.cuda.synchronize()
    reps = 3
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        call()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    nprob = c["nray"] * c["nlam"]
    res = {"columns": ncol, "depths": nd, "levels": nm, "problems_per_column": nprob, "ms_per_call": round(ms, 3),
           "columns_per_s": ncol / (ms * 1e-3)}
    log("NLTE response system: %d columns (Ca II, %d depths, %d problems each) in %.3f ms = %.4g columns/s"
        % (ncol, nd, nprob, ms, res["columns_per_s"]))
    cpu = None
    if with_cpu:
        sys.path.insert(0, os.path.join(REPO, "tests"))
        import _oracle  # the FP64 CPU restatement: baseline and checker only
        M0, b0 = M_t[0].cpu().numpy(), b_t[0].cpu().numpy()
        ck = dict(c, pops=inp["pops"][0])
        t1 = time.time()
        nref = 0
        while True:
            Mo, bo = _oracle.nlte_response_system(atom, ck, opem=inp["opem"][0])
            nref += 1
            if time.time() - t1 > 10.0:
                break
        dt = time.time() - t1
        rel = lambda a, b: float(np.max(np.abs(a - b) / np.maximum(np.abs(b).max(axis=-1, keepdims=True), 1e-300)))
        res["rel_err_vs_oracle"] = {"M": rel(M0, Mo), "beta_T": rel(b0[0], bo[0]), "beta_Nt": rel(b0[1], bo[1]),
                                    "beta_vt": rel(b0[2], bo[2])}
        cpu = {"value": nref / dt, "unit": "NLTE response systems (columns)/s", "cores": 1, "kind": "port",
               "sample": "the assembly of column 0, %d times (FP64 oracle: profile norms, formal_with_responses, "
                         "contributions and rate part, one thread; no solve)" % nref, "seconds": round(dt, 2)}
        log("NLTE response system cpu oracle: %.4g columns/s (one thread); GPU vs oracle %s"
            % (cpu["value"], res["rel_err_vs_oracle"]))
    res["cpu_baseline"] = cpu
    return res


def ali_leg(torch, dev, stream, side, with_cpu):
    from snapi_amd import abi, ali, workloads
    t0 = time.time()
    p = workloads.ali_problem(side, side)
    ncol = side * side
    nd = p["colstate"].shape[2]
    nlam, nray = p["nlam"], p["nray"]
    log("ALI workload: %d columns x %d depths, %d wavelengths, %d rays, built in %.1fs"
        % (ncol, nd, nlam, nray, time.time() - t0))
    T = {k: torch.from_numpy(np.ascontiguousarray(p[k])).to(dev)
         for k in ("colstate", "vproj", "gamma", "bgop", "bgem", "coll")}

    def run(n, precision, tol=1e-4, max_iter=300):
        pops = torch.from_numpy(p["pops"][:n].copy()).to(dev)
        it = torch.zeros(n, dtype=torch.int32, device=dev)
        rc = torch.zeros(n, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ali.solve_dev(p["atom"], p["grid"], T["colstate"][:n], T["vproj"][:n], T["gamma"][:n], T["bgop"][:n],
                      T["bgem"][:n], T["coll"][:n], pops, it, rc, tol=tol, max_iter=max_iter, precision=precision,
                      stream=stream)
        torch.cuda.synchronize()
        return time.perf_counter() - t1, pops.cpu().numpy(), it.cpu().numpy(), rc.cpu().numpy()

    # untimed: kernel loading, and the autotuner's choice of the
    # accumulation's occupancy for a batch this large (one iteration)
    run(ncol, abi.SNAPI_MIXED, max_iter=1)
    sec, pm, itm, rcm = run(ncol, abi.SNAPI_MIXED)
    nsol = float(itm.sum()) * nray * nlam
    mixed = {"columns": ncol, "seconds": round(sec, 3), "column_iterations": int(itm.sum()),
             "formal_solves_per_s": nsol / sec, "column_iter_per_s": float(itm.sum()) / sec,
             "mean_iterations": float(itm.mean()), "not_converged": int(np.sum((rcm < 0) | (rcm >= 1e-4)))}
    log("ALI mixed: %d column-iterations in %.2fs = %.4g formal solves/s; mean iterations %.1f"
        % (itm.sum(), sec, nsol / sec, itm.mean()))
    n64 = min(512, ncol)
    sec64, p64, it64, rc64 = run(n64, abi.SNAPI_FP64)
    fp64 = {"columns": n64, "seconds": round(sec64, 3), "formal_solves_per_s": float(it64.sum()) * nray * nlam / sec64,
            "mean_iterations": float(it64.mean())}
    err = float(np.max(np.abs(pm[:n64] / p64 - 1.0)))
    log("ALI fp64: %d columns in %.2fs; mixed vs fp64 populations max rel %.2e" % (n64, sec64, err))
    cpu = None
    if with_cpu:
        sys.path.insert(0, os.path.join(REPO, "tests"))
        import _oracle  # the FP64 CPU restatement: baseline only
        nthreads = max(1, min(16, os.cpu_count() or 1))
        nc = 16
        q = {k: np.ascontiguousarray(p[k][:nc]) for k in ("colstate", "vproj", "gamma", "bgop", "bgem", "coll")}
        q.update(atom=p["atom"], grid=p["grid"], nd=nd, ntrans=p["atom"].ntrans, nlevel=p["atom"].nlevel)
        t1 = time.time()
        opops, itc, _ = _oracle.ali_solve(q, p["pops"][:nc], 1e-4, 300, ncol=nc, nthreads=nthreads)
        dt = time.time() - t1
        cpu = {"value": float(itc.sum()) * nray * nlam / dt, "unit": "ALI formal solves/s", "cores": nthreads,
               "kind": "port", "sample": "%d columns of the same workload to 1e-4 (FP64 oracle, OpenMP)" % nc,
               "seconds": round(dt, 2)}
        cpu["fp64_gpu_populations_rel_err_vs_oracle"] = float(np.max(np.abs(p64[:nc] / opops - 1.0)))
        cpu["mixed_populations_rel_err_vs_oracle"] = float(np.max(np.abs(pm[:nc] / opops - 1.0)))
        cpu["iterations_equal_fp64_gpu"] = bool(np.array_equal(it64[:nc], itc))
        log("ALI cpu oracle: %.4g formal solves/s on %d threads; populations vs oracle: fp64 GPU %.2e, mixed %.2e"
            % (cpu["value"], nthreads, cpu["fp64_gpu_populations_rel_err_vs_oracle"],
               cpu["mixed_populations_rel_err_vs_oracle"]))
    syn = ca8542_leg(torch, dev, stream, p, pm, with_cpu)
    # roofline of the accumulation (ali_acc_kernel: opacity, formal solution
    # and atom::add sums per (column, ray, wavelength), float formal solution
    # in the mixed path): the frozen operation count of one formal solution
    # of this workload (profiles/ali_flops.json) times the whole iteration's
    # rate, rate solves and convergence checks included, so a lower bound
    roof = None
    ff = os.path.join(REPO, "profiles", "ali_flops.json")
    if os.path.exists(ff):
        fa = json.load(open(ff))["ali"]
        if (fa["depths"], fa["wavelengths"], fa["rays"]) == (nd, nlam, nray):
            ach = fa["flops_per_formal_solve"] * mixed["formal_solves_per_s"] / 1e12
            roof = {"bound": "valu", "achieved": round(ach, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(ach / FP32_PEAK_TFLOPS, 4), "traffic": None, "kernel": "ali_acc_kernel",
                    "flops_per_formal_solve": round(fa["flops_per_formal_solve"], 1), "flops_source": "counted",
                    "note": "rate of whole ALI iterations (rate solves and convergence checks included)"}
    return {"workload": "NLTE Ca II 5 levels + continuum by ALI, %d columns x %d depths, %d wavelengths, %d rays, "
                        "to relative change 1e-4" % (ncol, nd, nlam, nray),
            "value": mixed["formal_solves_per_s"], "unit": "ALI formal solves/s (column x ray x wavelength)",
            "mixed": mixed, "fp64_gpu": fp64, "populations_rel_err_mixed_vs_fp64": err, "roofline": roof,
            "cpu_baseline": cpu, "synthesis_8542": syn}


def ca8542_leg(torch, dev, stream, p, pops, with_cpu):
    """Polarized Ca II 8542 A synthesis (|B| <= 500 G) from the ALI leg's
    converged mixed-precision populations on every column, mixed precision,
    HIP events on the launch stream; the FP64 oracle on 16 columns checks it."""
    from snapi_amd import abi, api, workloads
    tb, ray, st, lam = workloads.ca8542_synthesis(p, pops)
    ncol, nf, nd = st.shape
    st_t = torch.from_numpy(st.astype(np.float32)).to(dev)
    lam_t = torch.from_numpy(lam).to(dev)
    out = torch.zeros((ncol, len(lam), 4), dtype=torch.float32, device=dev)
    api.synth_stokes_dev(tb, ray, st_t, lam_t, out, abi.SNAPI_MIXED, stream=stream)  # kernel loading, not timed
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(stream):
        e0.record(stream)
        for _ in range(5):
            api.synth_stokes_dev(tb, ray, st_t, lam_t, out, abi.SNAPI_MIXED, stream=stream)
        e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 5
    res = {"workload": "Ca II 8542 A polarized synthesis from the converged populations, %d columns x %d depths, "
                       "%d wavelengths, |B| <= 500 G, mixed" % (ncol, nd, len(lam)),
           "ms_per_launch": round(ms, 3), "lambda_point_solves_per_s": ncol * len(lam) / (ms * 1e-3)}
    log("8542 synthesis: %d columns x %d wavelengths in %.3f ms = %.4g lambda-point solves/s"
        % (ncol, len(lam), ms, res["lambda_point_solves_per_s"]))
    if with_cpu:
        import _oracle
        nc = 16
        ref, _ = _oracle.synth(tb, ray, st[:nc], lam)
        g = out[:nc].cpu().numpy().astype(np.float64)
        res["max_err_ic_vs_oracle"] = float(np.abs(g - ref).max() / ref[..., 0].max())
        res["max_abs_V_over_Ic"] = float(np.abs(ref[..., 3]).max() / ref[..., 0].max())
        log("8542 synthesis vs oracle on %d columns: %.2e Ic (max |V|/Ic %.2e)"
            % (nc, res["max_err_ic_vs_oracle"], res["max_abs_V_over_Ic"]))
    return res


def lm_leg(torch, dev, tmpl, lam_d, stream, side, nfd, tnode_side=48, tnode_legs=("mixed_tangent_analytic",),
           tnode_fd_iters=2):
    from snapi_amd import abi, api, invert, workloads
    table, ray = tmpl["table"], tmpl["ray"]
    nlam = lam_d.shape[0]
    t0 = time.time()
    prob = workloads.inversion_problem(tmpl, side, side, seed=3)
    npix = prob["state"].shape[0]
    nn = len(prob["nodes"])
    model = invert.node_model(prob["nodes"])
    ws = (1.0, 4.0, 4.0, 2.0)
    st64 = torch.from_numpy(prob["state"]).to(dev)
    obs64 = torch.empty((npix, nlam, 4), dtype=torch.float64, device=dev)
    api.synth_stokes_dev(table, ray, st64, lam_d, obs64, abi.SNAPI_FP64, stream)
    torch.cuda.synchronize()
    ic = obs64[:, :, 0].amax(dim=1)
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    noise = torch.randn(obs64.shape, dtype=torch.float64, device=dev, generator=g) * (1e-3 * ic)[:, None, None]
    obs64 += noise
    mask64 = torch.from_numpy(prob["mask"]).to(dev)
    start = torch.from_numpy(prob["start"]).to(dev)
    # chi^2 expected from the noise alone: sum over kept points of (w_s sigma)^2
    wsq = sum(w * w for w in ws)
    chi2_noise = (mask64.sum(dim=1) * wsq * (1e-3 * ic) ** 2)
    log("inversion workload: %d pixels, %d nodes, built in %.1fs" % (npix, nn, time.time() - t0))

    def run(n, precision, jacobian, label):
        dt_ = torch.float32 if precision == abi.SNAPI_MIXED else torch.float64
        st_t = st64[:n].to(dt_).contiguous()
        ob_t = obs64[:n].to(dt_).contiguous()
        mk_t = mask64[:n].to(dt_).contiguous()
        nodes = torch.empty((n, nn), dtype=torch.float64, device=dev)
        chi2 = torch.empty(n, dtype=torch.float64, device=dev)
        iters = torch.empty(n, dtype=torch.int32, device=dev)
        # untimed: kernel loading, and the autotuner's choices for the launch
        # shapes of a batch this large (one iteration of up to 16384 pixels)
        nw = min(n, 16384)
        invert.invert_dev(table, ray, model, st_t[:nw], lam_d, ob_t[:nw], mk_t[:nw], start[:nw], nodes[:nw],
                          chi2[:nw], iters[:nw], wstokes=ws, max_iter=1, precision=precision, jacobian=jacobian,
                          stream=stream)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        total = invert.invert_dev(table, ray, model, st_t, lam_d, ob_t, mk_t, start[:n], nodes, chi2, iters,
                                  wstokes=ws, max_iter=30, precision=precision, jacobian=jacobian, stream=stream)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t1
        it = iters.cpu().numpy()
        red = (chi2 / chi2_noise[:n]).cpu().numpy()
        nd_ = nodes.cpu().numpy()
        tr = prob["truth"][:n]
        strong = tr[:, 3] > 300.0
        berr = np.median(np.abs(nd_[strong, 3] - tr[strong, 3]) / tr[strong, 3])
        out = {"pixels": n, "pixel_iterations": int(total), "seconds": round(sec, 3),
               "pixel_iter_per_s": total / sec, "mean_iterations": float(it.mean()),
               "at_iteration_limit": float(np.mean(it >= 30)), "median_chi2_over_noise": float(np.median(red)),
               "median_B_rel_err_above_300G": float(berr)}
        log("%s: %d pixel-iterations in %.2fs = %.4g /s; mean iterations %.1f; median chi2/noise %.3g"
            % (label, total, sec, total / sec, it.mean(), np.median(red)))
        return out

    tnodes = tnode_inversion(torch, dev, tmpl, lam_d, stream, ws, side=tnode_side, legs=tnode_legs,
                             fd_iters=tnode_fd_iters)
    mixed = run(npix, abi.SNAPI_MIXED, 2, "mixed tangent")
    fd = run(min(nfd, npix), abi.SNAPI_FP64, 1, "fp64 differences")
    adj = run(min(nfd, npix), abi.SNAPI_MIXED, 0, "mixed adjoint")
    # node-Jacobian error of the mixed tangent sweep against the FP64 one
    # (8 pixels at their true node values, every 10th wavelength), relative to
    # the largest |dS/dnode| of each pixel and node
    lam_h = lam_d.cpu().numpy()[::10]
    st8, v8 = prob["state"][:8], prob["truth"][:8]
    _, j64 = invert.node_jacobian(table, ray, model, st8, v8, lam_h, abi.SNAPI_FP64, 2)
    _, j32 = invert.node_jacobian(table, ray, model, st8.astype(np.float32), v8, lam_h, abi.SNAPI_MIXED, 2)
    sc = np.maximum(np.abs(j64).max(axis=1, keepdims=True), 1e-300)
    jerr = float(np.max(np.abs(j32.astype(np.float64) - j64) / sc))
    kin = {"workload": "batched LM inversion, %d pixels of the cube, 71 depths, %d wavelengths masked to "
                       "120-%d per pixel, %d nodes (2 vLOS, 1 vmicro, 2 B, 1 incl, 1 azi) with the thermodynamic "
                       "state held fixed, noise 1e-3 Ic, start +/-20 %%, <=30 iterations" % (npix, nlam, nlam, nn),
           "value": mixed["pixel_iter_per_s"], "unit": "LM pixel-iterations/s",
           "mixed_tangent": mixed, "fp64_differences": fd, "mixed_adjoint": adj,
           "speedup_vs_fp64_differences": mixed["pixel_iter_per_s"] / fd["pixel_iter_per_s"],
           "jacobian_rel_err_mixed_vs_fp64": jerr}
    head = tnodes.get("mixed_tangent_analytic") or next(v for v in tnodes.values() if isinstance(v, dict))
    return {"workload": tnodes["workload"] + "; mixed tangent Jacobian, analytic node maps",
            "value": head["pixel_iter_per_s"], "unit": "LM pixel-iterations/s", "nodes": 14,
            "pixels": head["pixels"], "with_temperature_nodes": tnodes, "kinematic_7_nodes": kin}


def tnode_inversion(torch, dev, tmpl, lam_d, stream, ws, side=48, legs=("mixed_tangent_analytic",), fd_iters=2):
    """The 14-node model (7 temperature nodes): every iteration rebuilds the
    thermodynamics (hydrostatic equilibrium, equation of state) and the
    temperature-node responses (two rebuilds per node), adjoint Jacobian in
    mixed precision."""
    from snapi_amd import abi, api, invert, workloads
    table, ray = tmpl["table"], tmpl["ray"]
    nlam = lam_d.shape[0]
    t0 = time.time()
    prob = workloads.inversion_problem_tnodes(tmpl, side, side, seed=5)
    chem, chem_he, smap = workloads.eos_tables("cfg1_template", ray.lam_mid)
    npix = prob["base"].shape[0]
    nn = len(prob["nodes"])
    model = invert.node_model(prob["nodes"])
    st64 = torch.from_numpy(prob["base"]).to(dev)
    truth = torch.from_numpy(prob["truth"]).to(dev)
    invert.build_from_nodes_dev(model, chem, chem_he, smap, st64, truth, stream)
    obs64 = torch.empty((npix, nlam, 4), dtype=torch.float64, device=dev)
    api.synth_stokes_dev(table, ray, st64, lam_d, obs64, abi.SNAPI_FP64, stream)
    torch.cuda.synchronize()
    ic = obs64[:, :, 0].amax(dim=1)
    g = torch.Generator(device=dev)
    g.manual_seed(9)
    obs64 += torch.randn(obs64.shape, dtype=torch.float64, device=dev, generator=g) * (1e-3 * ic)[:, None, None]
    mask = torch.from_numpy(prob["mask"]).to(dev)
    start = torch.from_numpy(prob["start"]).to(dev)
    wsq = sum(w * w for w in ws)
    chi2_noise = mask.sum(dim=1) * wsq * (1e-3 * ic) ** 2
    log("inversion with temperature nodes: %d pixels, %d nodes, built in %.1fs" % (npix, nn, time.time() - t0))
    nodes = torch.empty((npix, nn), dtype=torch.float64, device=dev)
    chi2 = torch.empty(npix, dtype=torch.float64, device=dev)
    iters = torch.empty(npix, dtype=torch.int32, device=dev)

    def run(analytic, jacobian, precision=abi.SNAPI_MIXED, max_iter=30, label="", nslot=0):
        dt_ = torch.float32 if precision == abi.SNAPI_MIXED else torch.float64
        st_p = st64.to(dt_).contiguous()
        ob_p = obs64.to(dt_).contiguous()
        mk_p = mask.to(dt_).contiguous()
        # untimed: kernel loading, and the autotuner's choices for the launch
        # shapes of a batch this large (one iteration of up to 16384 pixels)
        nw = min(npix, nslot if nslot > 0 else 16384, 16384)
        invert.invert_eos_dev(table, ray, model, chem, chem_he, smap, st_p[:nw], lam_d, ob_p[:nw], mk_p[:nw],
                              start[:nw], nodes[:nw], chi2[:nw], iters[:nw], wstokes=ws, max_iter=1,
                              precision=precision, stream=stream, analytic_maps=analytic, jacobian=jacobian)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        total = invert.invert_eos_dev(table, ray, model, chem, chem_he, smap, st_p, lam_d, ob_p, mk_p, start, nodes,
                                      chi2, iters, wstokes=ws, max_iter=max_iter, nslot=nslot, precision=precision,
                                      stream=stream, analytic_maps=analytic, jacobian=jacobian)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t1
        it = iters.cpu().numpy()
        red = (chi2 / chi2_noise).cpu().numpy()
        dT = np.abs(nodes.cpu().numpy()[:, 2:6] - prob["truth"][:, 2:6])
        log("14-node inversion (%s): %d pixel-iterations in %.2fs = %.4g /s; mean iterations %.1f; median chi2/noise "
            "%.3g; median |dT| (log tau -3..0) %.0f K" % (label, total, sec, total / sec, it.mean(), np.median(red),
                                                          np.median(dT)))
        return {"pixels": npix, "pixel_iterations": int(total), "seconds": round(sec, 3),
                "pixel_iter_per_s": total / sec, "mean_iterations": float(it.mean()),
                "at_iteration_limit": float(np.mean(it >= max_iter)), "iteration_cap": max_iter,
                "median_chi2_over_noise": float(np.median(red)),
                "median_abs_T_err_logtau_-3_to_0_K": float(np.median(dT))}

    out = {"workload": "batched LM inversion with the 14-node model (7 T, 2 vLOS, 1 vmicro, 2 B, 1 incl, 1 azi), "
                       "%d pixels, 71 depths, %d wavelengths masked to 120-%d, noise 1e-3 Ic, start +/-20 %%, "
                       "<=30 iterations" % (npix, nlam, nlam)}
    for key in legs:
        if key == "mixed_tangent_analytic":
            # the fast path: forward-mode Jacobian, node responses by exact
            # derivatives and the hydrostatic tangent
            out[key] = run(True, abi.SNAPI_JAC_TANGENT, label="mixed tangent, analytic maps")
        elif key == "mixed_adjoint":
            # the reference's node responses (two hydrostatic rebuilds per
            # temperature node) with the adjoint Jacobian
            out[key] = run(False, abi.SNAPI_JAC_ADJOINT, label="mixed adjoint, reference node responses")
        elif key == "mixed_adjoint_analytic":
            out[key] = run(True, abi.SNAPI_JAC_ADJOINT, label="mixed adjoint, analytic maps")
        elif key == "mixed_tangent_analytic_first":
            out[key] = run(True, abi.SNAPI_JAC_TANGENT, max_iter=fd_iters,
                           label="mixed tangent, analytic maps, first %d iterations" % fd_iters)
        elif key == "fp64_differences":
            # 2n + 1 node sets per slot, each with its own hydrostatic march
            # and equation of state: 8192 slots at a time through the queue
            out[key] = run(False, abi.SNAPI_JAC_DIFFERENCES, abi.SNAPI_FP64, fd_iters,
                           label="fp64 differences, first %d iterations" % fd_iters, nslot=min(8192, npix))
    if "mixed_tangent_analytic_first" in out and "fp64_differences" in out:
        # the same iterations (every one with a Jacobian) of the same pixels
        out["speedup_vs_fp64_differences_same_iterations"] = (
            out["mixed_tangent_analytic_first"]["pixel_iter_per_s"] / out["fp64_differences"]["pixel_iter_per_s"])
    if "mixed_tangent_analytic" in out and "fp64_differences" in out:
        out["speedup_vs_fp64_differences"] = (out["mixed_tangent_analytic"]["pixel_iter_per_s"] /
                                              out["fp64_differences"]["pixel_iter_per_s"])
    # node-Jacobian error of the mixed tangent sweep against the FP64 one at
    # the truth of 8 pixels (every 10th wavelength), relative to the largest
    # |dS/dnode| of each pixel and node
    lam_h = lam_d.cpu().numpy()[::10]
    st8 = invert.build_from_nodes(model, chem, chem_he, smap, prob["base"][:8], prob["truth"][:8])
    v8 = prob["truth"][:8]
    _, j64 = invert.node_jacobian_eos(table, ray, model, chem, chem_he, smap, st8, v8, lam_h, abi.SNAPI_FP64,
                                      jacobian=abi.SNAPI_JAC_TANGENT)
    _, j32 = invert.node_jacobian_eos(table, ray, model, chem, chem_he, smap, st8, v8, lam_h, abi.SNAPI_MIXED,
                                      jacobian=abi.SNAPI_JAC_TANGENT)
    sc = np.maximum(np.abs(j64).max(axis=1, keepdims=True), 1e-300)
    out["jacobian_rel_err_mixed_vs_fp64"] = float(np.max(np.abs(j32.astype(np.float64) - j64) / sc))
    log("14-node tangent Jacobian, mixed vs FP64: %.2e of each node's largest entry"
        % out["jacobian_rel_err_mixed_vs_fp64"])
    return out


if __name__ == "__main__":
    main()
