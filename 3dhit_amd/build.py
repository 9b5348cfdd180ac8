"""Builds libhit3d_hip.so (gfx950) in-tree from csrc/: one hipcc process per
translation unit, then a shared link. Usage: python 3dhit_amd/build.py [-j N]"""

from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "hit3d")
LIB = os.path.join(OUT_DIR, "libhit3d_hip.so")
OBJ_DIR = os.path.join(HERE, "_build")

ARCH = os.environ.get("HIT3D_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # no packed-FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

# (object name, source, extra flags)
UNITS = [
    ("passes_f32.o", "passes.hip", ["-DH3_REAL=float"]),
    ("passes_f64.o", "passes.hip", ["-DH3_REAL=double"]),
    ("misc.o", "misc.hip", []),
    ("closure.o", "closure.hip", []),
    ("capi.o", "capi.hip", []),
]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(unit):
    name, src, extra = unit
    cmd = [_hipcc(), *FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", os.path.join(OBJ_DIR, name)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = [l for l in (r.stderr or "").splitlines() if "packed-fp32-ops" not in l]
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n" + "\n".join(noise[-60:]))
    return name


def build(jobs: int | None = None, verbose: bool = False) -> str:
    os.makedirs(OBJ_DIR, exist_ok=True)
    subprocess.run([sys.executable, os.path.join(CSRC, "gen_twiddles.py")], check=True)
    jobs = jobs or min(len(UNITS), max(1, len(os.sched_getaffinity(0))))
    with ThreadPoolExecutor(jobs) as ex:
        for name in ex.map(_compile, UNITS):
            if verbose:
                print("compiled", name, flush=True)
    objs = [os.path.join(OBJ_DIR, u[0]) for u in UNITS]
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args()
    print(build(a.j, verbose=True))
