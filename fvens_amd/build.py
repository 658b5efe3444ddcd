"""Builds libfvx.so (HIP, gfx950) in-tree from fvens_amd/csrc.

Every HIP source is compiled to its own object (the flux kernel once per numerical flux) in
parallel, then linked into fvens_amd/libfvx.so. Invoked by __graft_entry__.build().
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfvx.so")
ARCH = "gfx950"

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = ["-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
          # packed FP32 instructions are excluded from all device code
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          # multiply-adds are fused only within one source expression, so a formula rounds the
          # same in every kernel it is inlined into (the two forms of the flux pass give
          # bitwise equal residuals)
          "-ffp-contract=on",
          "-I", os.path.join(HERE, "..", "include"), "-Wall", "-Wno-unused-result"]
NFLUX = 7


def _units():
    units = []
    for k in range(NFLUX):
        units.append((os.path.join(CSRC, "flux_inst.hip"), os.path.join(OBJ, f"flux_{k}.o"),
                      [f"-DFVX_FLUX_ID={k}"]))
    for name in sorted(os.listdir(CSRC)):
        if name.endswith(".hip") and name != "flux_inst.hip":
            units.append((os.path.join(CSRC, name), os.path.join(OBJ, name[:-4] + ".o"), []))
        elif name.endswith(".cpp"):
            units.append((os.path.join(CSRC, name), os.path.join(OBJ, name[:-4] + ".o"),
                          ["-x", "hip"]))
    return units


def _compile(unit):
    src, obj, extra = unit
    cmd = [HIPCC] + COMMON + extra + ["-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{p.stdout}\n{p.stderr}")
    return obj


def build(verbose: bool = False, jobs: int | None = None) -> str:
    os.makedirs(OBJ, exist_ok=True)
    units = _units()
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, units))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{p.stdout}\n{p.stderr}")
    os.replace(tmp, LIB)
    # C++ case driver over the host adapter (host code only, links libfvx.so)
    drv = os.path.join(HERE, "host", "fvx_case")
    cmd = [HIPCC, "-std=c++17", "-O2", "-Wall", os.path.join(HERE, "host", "fvx_case.cpp"),
           "-I", os.path.join(HERE, "..", "include"), "-L", HERE, "-lfvx",
           "-Wl,-rpath,$ORIGIN/..", "-o", drv + ".tmp"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{p.stdout}\n{p.stderr}")
    os.replace(drv + ".tmp", drv)
    if verbose:
        print(f"built {LIB} and {drv}", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
