"""Build libecsmhd.so in-tree with hipcc for gfx950.

The library is split into one translation unit per grid size (csrc/grid_n*.hip,
each instantiating the kernels and drivers of one N) plus the host/C-ABI unit
(csrc/ecsmhd.hip); they compile in parallel into objects under _build/ and are
linked into the shared object.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libecsmhd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
UNITS = ["ecsmhd.hip"] + [f"grid_n{n}.hip" for n in (16, 32, 64, 128, 256)]

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed-FP32 instructions stay off (FP64 code never needs them)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile the HIP library; returns the path of the shared object."""
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(UNITS), os.cpu_count() or 1)
    objs, pending, failed = [], [], []
    for unit in UNITS:
        obj = os.path.join(BUILD, unit.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC, *FLAGS, "-c", "-o", obj, os.path.join(CSRC, unit)]
        pending.append((unit, cmd))
    running = []
    while pending or running:
        while pending and len(running) < jobs:
            unit, cmd = pending.pop(0)
            running.append((unit, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        unit, proc = running.pop(0)
        out, err = proc.communicate()
        if proc.returncode != 0:
            failed.append(unit)
            sys.stderr.write(out + err)
        elif verbose:
            sys.stderr.write(err)
    if failed:
        raise RuntimeError("hipcc failed on " + ", ".join(failed))
    tmp = OUT + ".tmp"
    link = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC",
            "-Wl,--version-script=" + os.path.join(CSRC, "exports.map"), "-o", tmp, *objs]
    res = subprocess.run(link,
                         capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed to link libecsmhd.so")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
