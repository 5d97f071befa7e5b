"""Builds libmpm3d_hip.so in-tree for gfx950.

    python mpm3d-f90_amd/build.py [--jobs N]

HIP sources are compiled with hipcc --offload-arch=gfx950; the deck reader
(host only) with g++.  -ffp-contract=off keeps the FP64 path free of FMA
contraction so its per-particle arithmetic rounds like the reference's; the
FP32-bulk kernels are a second object of the same source with contraction
allowed.
Packed-FP32 instruction selection is switched off for all device code.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
INC = os.path.join(ROOT, "include")
OUT = os.path.join(HERE, "libmpm3d_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument", "-I" + INC, "-I" + CSRC]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-I" + INC, "-I" + CSRC]

# (source, object, extra flags): kernels.hip is compiled once per precision.
# The FP32-bulk object contracts to FMA and takes FP32 division and square
# root to within a few ulp (v_rcp_f32 / v_sqrt_f32 based) instead of the
# correctly rounded sequences; the FP64 object keeps the reference's roundings.
HIP_SOURCES = [("kernels.hip", "kernels_f64.o", ["-DMPMK_PART=1"]),
               ("kernels.hip", "kernels_f32.o", ["-DMPMK_PART=2", "-ffp-contract=fast",
                                                 "-fno-hip-fp32-correctly-rounded-divide-sqrt"]),
               ("engine.hip", "engine.o", []),
               ("radix.hip", "radix.o", [])]
CXX_SOURCES = ["deck.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build step failed: %s\n%s" % (" ".join(cmd), err[-6000:]))
    return err


def build(jobs=4, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs_list = []
    objs = []
    for s, obj, extra in HIP_SOURCES:
        o = os.path.join(BUILD, obj)
        jobs_list.append([HIPCC] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in CXX_SOURCES:
        o = os.path.join(BUILD, s.replace(".cpp", ".o"))
        jobs_list.append(["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for msg in ex.map(_run, jobs_list):
            if verbose and msg.strip():
                print(msg)
    tmp = OUT + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, OUT)
    # Fortran host (ISO_C_BINDING) when a Fortran compiler is present
    if shutil.which("amdflang") or os.path.exists("/opt/rocm/bin/amdflang"):
        env = dict(os.environ, PATH=os.environ.get("PATH", "") + ":/opt/rocm/bin")
        r = subprocess.run(["make", "-s", "-B", "-C", os.path.join(HERE, "fortran")], capture_output=True,
                           text=True, env=env)
        if r.returncode != 0:
            raise RuntimeError("Fortran host build failed:\n" + r.stdout[-3000:] + r.stderr[-3000:])
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    a = ap.parse_args()
    print(build(a.jobs, verbose=True))
    sys.exit(0)
