"""Build the gfx950 HIP library in-tree: mettleq_amd/libmettleq_sv.so.

    python -m mettleq_amd.build

hipcc cross-compiles for gfx950 without a GPU. Flags:
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 VALU ops
  -mllvm -simplifycfg-sink-common=false              keeps the register-resident
      amplitude arrays of the tile kernel out of scratch memory (sinking the
      per-target butterfly cases into one block turns register indices into
      pointer phis, which forces the array into scratch).
  -mllvm -instcombine-max-copied-from-constant-users=4096
      keeps the sweep kernels' by-value argument blocks out of scratch memory:
      the frontend copies a by-value kernel argument whose address is taken
      into private memory, and InstCombine forwards that copy back to the
      kernel argument segment only while it has at most this many users (the
      default, 300, is below the use count of the fused first sweeps, which
      then read their 2.6 KiB of arguments from scratch).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "libmettleq_sv.so"
SOURCES = ["msv_kernels.hip", "msv_runtime.hip", "msv_batch.hip", "msv_preflight.hip", "msv_mps.hip"]
HEADERS = ["msv_internal.h", "msv_device.h", "msv_philox.h"]

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-simplifycfg-sink-common=false",
    "-mllvm", "-instcombine-max-copied-from-constant-users=4096",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the engine)")


def build(verbose: bool = False) -> Path:
    cmd = [hipcc(), *HIP_FLAGS, *[str(CSRC / s) for s in SOURCES], "-o", str(OUT) + ".tmp",
           "-L/opt/rocm/lib", "-lhsa-runtime64"]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    noise = "packed-fp32-ops"
    errs = "\n".join(l for l in proc.stderr.splitlines() if noise not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed ({proc.returncode}):\n{errs}")
    os.replace(str(OUT) + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
