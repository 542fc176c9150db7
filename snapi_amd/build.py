"""Builds the HIP library snapi_amd/libsnapi_gpu.so for gfx950 with hipcc.

python -m snapi_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsnapi_gpu.so")
SOURCES = ["runtime.hip", "synth.hip", "responses.hip", "tangent.hip", "lm.hip", "formal_scalar.hip", "invert.hip",
           "ali.hip", "chemeq.hip",
           "nlte_resp.hip", "instrument.hip", "cube.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         # packed-FP32 instructions stay off: hipcc would otherwise form
         # v_pk_* FP32 ops from scalar float code
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    missing = [p for p in srcs if not os.path.exists(p)]
    if missing:
        raise FileNotFoundError("HIP sources missing: %s" % missing)
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
