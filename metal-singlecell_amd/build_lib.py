"""Build libmsc_hip.so (gfx950) in-tree: ``python build_lib.py [--jobs N]``.

Each ``csrc/*.hip`` file is compiled to an object with hipcc, then linked into
``metalsinglecell/_lib/libmsc_hip.so``.  Flags:

* ``--offload-arch=gfx950``: CDNA4 only.
* ``-ffp-contract=off``: no implicit fused multiply-add; the exact-distance and fixed-point
  arithmetic of the path is defined with separately rounded operations.
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 instructions.
* ``knn.hip`` only: ``-mno-amdgpu-ieee -fno-honor-nans`` (see ``FILE_FLAGS``).
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "metalsinglecell", "_lib")
OBJ_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libmsc_hip.so")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}"]


# knn.hip: min/max without NaN canonicalisation (the select's gate takes the minimum of 16
# MFMA results per M-block; distances are never NaN).  Arithmetic is unchanged.
FILE_FLAGS = {"knn.hip": ["-mno-amdgpu-ieee", "-fno-honor-nans"]}


def _compile(src):
    obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, *FILE_FLAGS.get(os.path.basename(src), []), "-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(line for line in p.stderr.splitlines()
                    if "packed-fp32-ops" not in line and "not a recognized feature" not in line)
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{err}")
    return obj, err


def build(jobs=None, verbose=False):
    os.makedirs(OBJ_DIR, exist_ok=True)
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    jobs = jobs or min(8, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(_compile, srcs))
    for _, err in results:
        if verbose and err.strip():
            print(err, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed:\n{p.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))
