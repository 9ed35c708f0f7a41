"""Compile libqd_hip.so for gfx950 in-tree (hipcc; no JIT cache).

The translation units compile in parallel: the C-ABI host code, the MLP
kernels, and qd_dispatch.hip once per kernel family (-DQD_PART=0..10).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libqd_hip.so")
DISPATCH_PARTS = 11

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _units():
    units = [("qd_api.hip", []), ("qd_mlp.hip", [])]
    units += [("qd_dispatch.hip", ["-DQD_PART=%d" % p]) for p in range(DISPATCH_PARTS)]
    return units


def build(verbose: bool = False, jobs: int = 0) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    jobs = jobs or max(1, min(len(_units()), os.cpu_count() or 1, 16))
    tmpdir = tempfile.mkdtemp(prefix="qd_build_")
    try:
        def compile_one(i_unit):
            i, (src, defs) = i_unit
            obj = os.path.join(tmpdir, "u%d.o" % i)
            cmd = [hipcc] + HIPCC_FLAGS + defs + ["-c", os.path.join(CSRC, src), "-o", obj]
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError("hipcc failed on %s %s:\n%s" % (src, " ".join(defs), r.stderr[-4000:]))
            return obj

        with ThreadPoolExecutor(jobs) as ex:
            objs = list(ex.map(compile_one, enumerate(_units())))
        tmp = OUT + ".tmp"
        r = subprocess.run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs,
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("hipcc link failed:\n" + r.stderr[-4000:])
        os.replace(tmp, OUT)
    finally:
        shutil.rmtree(tmpdir, ignore_errors=True)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
