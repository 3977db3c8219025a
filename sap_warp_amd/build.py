"""Build ``libsapx.so`` (gfx950) in-tree with hipcc.

Every translation unit is compiled for gfx950 only, with packed-fp32 vector
ops disabled (``-packed-fp32-ops``) and IEEE-rounded float32 division and
square root.  The collision unit additionally disables FMA contraction so its
float32 results are bit-identical to the CPU restatement in ``oracle/``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsapx.so")
BUILD = os.path.join(HERE, "_build")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
UNITS = {
    "sapx_step.hip": [],
    "sapx_collide.hip": ["-ffp-contract=off"],
    "sapx_group.hip": [],
    "sapx_api.hip": [],
    "sapx_wave16.hip": [],
    "sapx_wave32.hip": [],
    "sapx_wave48.hip": [],
    "sapx_wave64.hip": [],
    "sapx_batch16.hip": [],
    "sapx_batch32.hip": [],
    "sapx_batch64.hip": [],
    "sapx_order.hip": [],
    "sapx_learned.hip": [],
    "sapx_mlp.hip": [],
    "sapx_chol.hip": [],
    "sapx_large.hip": [],
    "sapx_bigworld.hip": [],
}


def _compile(unit: str, extra: list) -> str:
    obj = os.path.join(BUILD, unit.replace(".hip", ".o"))
    cmd = [HIPCC, *COMMON, *extra, "-c", os.path.join(CSRC, unit), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {unit}:\n{res.stdout}\n{res.stderr}")
    if res.stderr.strip():
        sys.stderr.write(res.stderr)
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(UNITS)) as ex:
        objs = list(ex.map(lambda kv: _compile(*kv), UNITS.items()))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
