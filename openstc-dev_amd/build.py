"""Builds openstc-dev_amd/libostc.so for gfx950 with hipcc (in-tree).

    python openstc-dev_amd/build.py [-j N]

Every translation unit is compiled with hipcc (host C++ and HIP alike) and
linked into one shared library exporting the C ABI of include/ostc.h.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libostc.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["setup_io.cpp", "chemkin_io.cpp", "mesh.cpp", "plt_io.cpp", "api.cpp", "flux.hip", "update.hip", "implicit.hip", "sst.hip", "mlp.hip", "surrogate.hip", "grid.hip", "flux_sst.hip"]

# The k-omega SST sweeps: at -O2 / -O3 the FP64 k-direction sweep of the
# 9-species + SST case gives run-to-run varying results in the k row; -O1
# builds of the same source are reproducible and match the CPU oracle.
OPT_LEVEL = {"flux_sst.hip": "-O1"}

COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
          "-Wno-unused-command-line-argument"]


def compile_one(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + ["-c", os.path.join(SRC, src), "-o", obj]
    if src.endswith(".hip"):
        cmd[1:1] = ["-x", "hip"]
    if src in OPT_LEVEL:
        cmd = [OPT_LEVEL[src] if c == "-O3" else c for c in cmd]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args()
    try:
        print(build(a.j))
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
