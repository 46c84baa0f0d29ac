"""Builds liblbcuda_amd.so (HIP, gfx950) in-tree with hipcc.

Packed FP32 instructions are disabled for all device code
(-target-feature -packed-fp32-ops).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "liblbcuda_amd.so")
SOURCES = ["lbm_fused.hip", "lbm_fluid.hip", "lbm_particles.hip", "lbm_mlp.hip", "lbm_api.cpp"]
ARCH = os.environ.get("LBM_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for p in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(p) or p == "hipcc":
            return p


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    common = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              "-Wall", "-Wno-unused-function", "-Wno-unused-result"]
    cmds = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(OUT_DIR, os.path.splitext(src)[0] + ".o")
        cmds.append(common + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        out = "\n".join(x for x in r.stdout.splitlines() if "not a recognized feature" not in x)
        if out.strip():
            print(out, flush=True)
        return r.returncode

    with ThreadPoolExecutor(max_workers=min(len(cmds), 8)) as ex:
        codes = list(ex.map(run, cmds))
    if any(codes):
        raise RuntimeError("hipcc failed")
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
