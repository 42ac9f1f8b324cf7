"""Builds libhpgl.so (gfx950) into the geo_bsd package directory.

    python hpgl-reborn_amd/build.py

* plan.cpp, variogram_plan.cpp and sgs_plan.cpp are compiled with g++
  -ffp-contract=off: they produce the covariance keys that order the search
  template, the variogram's tunnel and lag decisions and the simulation's
  deviates and normal-score transforms, which must round exactly as the
  reference's own g++-built code does.
* The .hip sources are compiled with hipcc for gfx950 only, with packed FP32
  instructions disabled.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "geo_bsd")
BUILD = os.path.join(HERE, "build")
LIB = os.path.join(OUT_DIR, "libhpgl.so")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for cpp in ("plan.cpp", "variogram_plan.cpp", "sgs_plan.cpp"):
        o = os.path.join(BUILD, cpp.replace(".cpp", ".o"))
        _run(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, cpp), "-o", o])
        objs.append(o)
    # the HIP translation units compile in parallel
    procs = []
    units = ["krige.hip", "solve_wide.hip", "loo.hip", "api.hip", "inc_io.hip", "variogram.hip", "sgs.hip", "loo_m0.hip",
             "loo_m1.hip", "loo_m2.hip"]
    units += [f"solve_n{n}_k{k}.hip" for n in (32, 24, 16, 8) for k in (0, 1, 2)]
    for src in units:
        o = os.path.join(BUILD, src.replace(".hip", ".o"))
        cmd = [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", o]
        print("+", " ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd)))
        objs.append(o)
    failed = [src for src, pr in procs if pr.wait() != 0]
    if failed:
        raise RuntimeError("hipcc failed for " + ", ".join(failed))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    shutil.move(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
