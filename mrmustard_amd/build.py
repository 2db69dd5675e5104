"""Build libmmhip.so (the HIP kernels and their C ABI) for gfx950, in-tree.

The library is several translation units (core sweeps and VJPs, one per
LDS sweep kernel instantiation, the photon-number-resolved strategies, the
matrix-core product, autoshape), compiled in parallel and linked into one
shared object.
"""

from __future__ import annotations

import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

_HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(_HERE, "csrc")
SOURCES = [
    "hermite.hip",
    "lds_3f.hip",
    "lds_3.hip",
    "lds_8f.hip",
    "lds_8.hip",
    "compact.hip",
    "zgemm.hip",
    "autoshape.hip",
]
SRC = os.path.join(CSRC, "hermite.hip")
OUT_DIR = os.path.join(_HERE, "_build")
OUT = os.path.join(OUT_DIR, "libmmhip.so")

# -ffp-contract=off: the forward recurrence must round like the reference
# (numba compiles it without fastmath, so no multiply-add is ever fused).
# -packed-fp32-ops: never emit packed FP32 instructions.
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}): {' '.join(cmd[-3:])}\n{res.stderr[-4000:]}")
    return res.stderr


def build_library(verbose: bool = False) -> str:
    """Compile csrc/*.hip into _build/libmmhip.so (always recompiles)."""
    os.makedirs(OUT_DIR, exist_ok=True)
    cc = hipcc()
    objs = [os.path.join(OUT_DIR, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    cmds = [[cc, *HIPCC_FLAGS, "-c", "-o", o, os.path.join(CSRC, s)] for s, o in zip(SOURCES, objs)]
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        logs = list(ex.map(_run, cmds))
    tmp = OUT + ".tmp"
    logs.append(_run([cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]))
    if verbose:
        for lg in logs:
            if lg:
                print(lg[-2000:])
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build_library(verbose=True))
