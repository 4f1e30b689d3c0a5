"""Builds libseqvp_hip.so (gfx950) in-tree with hipcc.

Every translation unit is compiled for gfx950 only. The prescreen and certificate units are compiled
without floating-point contraction: the prescreen's proof bound assumes every +, -, * is rounded
separately, the interval certificate widens each separately rounded result by one ulp, and the pyramid
resampling reproduces the separately rounded (1 - w) a + w b of the Q1 prolongation.
Packed FP32 instructions are disabled for all units.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libseqvp_hip.so"
OBJ = HERE / "_build"

SOURCES = ("q1_sweeps.hip", "prescreen.hip", "certificate.hip", "spline.hip", "dst.hip", "vec.hip", "preopt.hip",
           "resample.hip", "ceiling.hip", "graph.hip", "abi.hip")
COMMON = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
PER_FILE = {"prescreen.hip": ["-ffp-contract=off"], "certificate.hip": ["-ffp-contract=off"],
            "resample.hip": ["-ffp-contract=off"]}


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    return "hipcc"


def _compile(src):
    out = OBJ / (Path(src).stem + ".o")
    cmd = [hipcc(), *COMMON, *PER_FILE.get(src, []), "-c", str(CSRC / src), "-o", str(out)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{proc.stderr}")
    return out, proc.stderr


def build(verbose=False):
    """Compile every HIP source for gfx950 and link the shared library (always rebuilds)."""
    OBJ.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as pool:
        results = list(pool.map(_compile, SOURCES))
    if verbose:
        for out, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *(str(o) for o, _ in results)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("link failed:\n" + proc.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
