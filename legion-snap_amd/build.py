"""Build libsnap_hip.so for gfx950 in-tree (legion-snap_amd/snap_amd/).

    python legion-snap_amd/build.py [--jobs N]

Flags:
  -ffp-contract=off   FP64 kernels round every multiply and add separately,
                      as the reference's CPU arithmetic does; FP32 kernels opt
                      into fused multiply-adds explicitly.
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed-FP32 instructions (v_pk_*_f32) are generated.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "snap_amd")
LIB = os.path.join(OUT_DIR, "libsnap_hip.so")
OBJ_DIR = os.path.join(HERE, "_build")
SOURCES = ["snap_common.hip", "snap_xs.hip", "snap_source.hip", "snap_outer.hip",
           "snap_sweep.hip", "snap_sweep_seq.hip", "snap_adjoint.hip", "snap_dropin.hip", "snap_solver.hip",
           "snap_batch.hip", "snap_mms.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SNAP_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    "-I", os.path.join(REPO, "include"), "-I", CSRC,
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {src}")
    noise = "packed-fp32-ops"
    msgs = [ln for ln in (r.stdout + r.stderr).splitlines() if ln.strip() and noise not in ln]
    if msgs:
        sys.stderr.write("\n".join(msgs) + "\n")
    return obj


def build(jobs: int = 0) -> str:
    os.makedirs(OBJ_DIR, exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=0)
    a = ap.parse_args()
    print(build(a.jobs))
