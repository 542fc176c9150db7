"""Builds the HIP library snapi_amd/libsnapi_gpu.so for gfx950 with hipcc.

Every source is compiled to its own object (in parallel, under
snapi_amd/_build/), then the objects are linked into the shared library.

python -m snapi_amd.build
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsnapi_gpu.so")
SOURCES = ["runtime.hip", "synth.hip", "responses.hip", "tangent.hip", "lm.hip", "formal_scalar.hip", "invert.hip",
           "ali.hip", "chemeq.hip",
           "nlte_resp.hip", "nlte_system.hip", "instrument.hip", "cube.hip", "atomdata.hip", "autotune.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         # packed-FP32 instructions stay off: hipcc would otherwise form
         # v_pk_* FP32 ops from scalar float code
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


def _jobs():
    n = os.cpu_count() or 4
    return max(1, min(len(SOURCES), n, 16))


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    missing = [p for p in srcs if not os.path.exists(p)]
    if missing:
        raise FileNotFoundError("HIP sources missing: %s" % missing)
    os.makedirs(OBJ, exist_ok=True)
    objs = [os.path.join(OBJ, os.path.splitext(s)[0] + ".o") for s in SOURCES]

    def compile_one(pair):
        src, obj = pair
        cmd = [HIPCC] + FLAGS + ["-c", "-o", obj, src]
        if verbose:
            print(" ".join(cmd), flush=True)
        return src, subprocess.run(cmd, capture_output=True, text=True)

    with ThreadPoolExecutor(max_workers=_jobs()) as ex:
        results = list(ex.map(compile_one, zip(srcs, objs)))
    failed = [(s, r) for s, r in results if r.returncode != 0]
    if failed:
        for s, r in failed:
            sys.stderr.write("%s:\n%s%s" % (s, r.stdout, r.stderr))
        raise RuntimeError("hipcc failed on %s" % ", ".join(os.path.basename(s) for s, _ in failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc link failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=False))
