"""Builds libhpgl.so (gfx950) into the geo_bsd package directory.

    python hpgl-reborn_amd/build.py

* plan.cpp, variogram_plan.cpp and sgs_plan.cpp are compiled with g++
  -ffp-contract=off: they produce the covariance keys that order the search
  template, the variogram's tunnel and lag decisions and the simulation's
  deviates and normal-score transforms, which must round exactly as the
  reference's own g++-built code does.
* The .hip sources are compiled with hipcc for gfx950 only, with packed FP32
  instructions disabled.

    python hpgl-reborn_amd/build.py --incremental

recompiles only the objects older than their source or than any header of
csrc/ and include/.
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


def _stale(src, obj, newest_header):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return os.path.getmtime(src) > t or newest_header > t


def build(verbose=True, incremental=False):
    os.makedirs(BUILD, exist_ok=True)
    headers = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    inc = os.path.join(os.path.dirname(HERE), "include")
    headers += [os.path.join(inc, f) for f in os.listdir(inc) if f.endswith(".h")]
    newest_header = max(os.path.getmtime(h) for h in headers)
    objs = []
    for cpp in ("plan.cpp", "variogram_plan.cpp", "sgs_plan.cpp"):
        o = os.path.join(BUILD, cpp.replace(".cpp", ".o"))
        if not incremental or _stale(os.path.join(CSRC, cpp), o, newest_header):
            _run(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, cpp), "-o", o])
        objs.append(o)
    # the HIP translation units compile in parallel
    procs = []
    units = ["krige.hip", "solve_wide.hip", "loo.hip", "api.hip", "inc_io.hip", "variogram.hip", "sgs.hip", "loo_m0.hip",
             "loo_m1.hip", "loo_m2.hip"]
    units += [f"solve_n{n}_k{k}.hip" for n in (32, 24, 16, 8) for k in (0, 1, 2)]
    for src in units:
        o = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(o)
        if incremental and not _stale(os.path.join(CSRC, src), o, newest_header):
            continue
        cmd = [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", o]
        print("+", " ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd)))
    failed = [src for src, pr in procs if pr.wait() != 0]
    if failed:
        raise RuntimeError("hipcc failed for " + ", ".join(failed))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    shutil.move(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv[1:])
    sys.exit(0)
