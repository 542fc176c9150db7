"""Times the compiled reference (oracle/_ref/ref_driver, built by oracle/ref.mk
from /root/reference) on this machine's cores, the way the reference deploys:
one single-threaded worker process per core.

  atmosphere::obs_stokes on the cfg0 column (61 depths, 402 wavelengths) and
  the cube's column shape (71 depths, 301 wavelengths);
  atmosphere::stokes_lm_fit with the 14-node model on the fit fixture's
  pixels (tests/golden/fit.npz: 71 depths, 101 wavelengths), 8 iterations.

Runs where /root/reference is (the atom files come from its cfg/):
    python oracle/time_reference.py [nproc]
"""
import math
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "tests", "golden"))
import make_golden as mg  # noqa: E402


def time_synth(nproc, tag, nd, lam, reps=3):
    d = "/tmp/snapi_time_ref_%s" % tag
    os.makedirs(d, exist_ok=True)
    mg.setup_dir(d, mg.cfg0_atmosphere(nd), lam)
    ps = [subprocess.Popen([mg.DRIVER, "time_synth", "atmos.cfg", "lambda.txt", repr(math.pi), "0.0", str(reps)],
                           cwd=d, env=dict(os.environ, SNAPI_REF_HE="1"), stdout=subprocess.PIPE,
                           stderr=subprocess.DEVNULL, text=True) for _ in range(nproc)]
    per = [float(p.communicate()[0].split("seconds_per_call")[1].split()[0]) for p in ps]
    s = float(np.mean(per))
    print("obs_stokes, %d depths x %d wavelengths: %.4f s per call per process; %d processes: %.4g lambda-point "
          "solves/s" % (nd, len(lam), s, nproc, nproc * len(lam) / s), flush=True)


def time_fit(nproc, kmax=8):
    d = "/tmp/snapi_time_ref_fit"
    os.makedirs(d, exist_ok=True)
    lam = mg.fit_lambda()
    mg.setup_dir(d, mg.cfg0_atmosphere(71), lam)
    f = np.load(os.path.join(os.path.dirname(HERE), "tests", "golden", "fit.npz"))
    for k in range(2):
        path = os.path.join(d, "fit_%d.txt" % k)
        mg.write_model(path, f["start%d" % k])
        with open(path, "a") as fh:
            ws = f["ws"]
            fh.write("%.17g %.17g %.17g %.17g %.17g %d\n" % (ws[0], ws[1], ws[2], ws[3], 1.0, kmax))
            fh.write(" ".join("%.17g" % m for m in f["mask%d" % k]) + "\n")
            for row in f["obs%d" % k]:
                fh.write(" ".join("%.17g" % v for v in row) + "\n")
    t0 = time.time()
    ps = [subprocess.Popen([mg.DRIVER, "fit", "atmos.cfg", "lambda.txt", repr(math.pi), "0.0", "fit_%d.txt" % (i % 2),
                            "last"], cwd=d, env=dict(os.environ, SNAPI_REF_HE="0"), stdout=subprocess.DEVNULL,
                           stderr=subprocess.DEVNULL) for i in range(nproc)]
    for p in ps:
        p.wait()
    wall = time.time() - t0
    print("stokes_lm_fit, 14 nodes, 71 depths x %d wavelengths: %d processes x %d iterations in %.1f s = %.3g "
          "pixel-iterations/s" % (len(lam), nproc, kmax, wall, nproc * kmax / wall), flush=True)


if __name__ == "__main__":
    n = int(sys.argv[1]) if len(sys.argv) > 1 else (os.cpu_count() or 1)
    time_synth(n, "cfg0", 61, mg.cfg0_lambda())
    time_synth(n, "cfg1", 71, mg.cfg1_lambda())
    time_fit(n)
