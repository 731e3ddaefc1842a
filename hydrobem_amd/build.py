"""Build of the HIP library: hipcc for gfx950 into hydrobem_amd/_build/libhydrobem_amd.so.

Every source of hydrobem_amd/csrc is compiled in one hipcc invocation (the device code for
gfx950 only). Packed FP32 instructions are switched off on the device side
(-target-feature -packed-fp32-ops).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "csrc"
OUT = HERE / "_build"
LIB = OUT / "libhydrobem_amd.so"
ARCH = "gfx950"


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and Path(c).exists():
            return c
    raise RuntimeError("hipcc not found (ROCm is required to build hydrobem_amd)")


def sources():
    return sorted(SRC.glob("*.hip")) + sorted(SRC.glob("*.cpp"))


def build(verbose: bool = False) -> Path:
    """Compile the library (always; the caller decides when)."""
    OUT.mkdir(parents=True, exist_ok=True)
    tmp = OUT / "libhydrobem_amd.so.tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-result", "-I", str(SRC), "-o", str(tmp)] + [str(s) for s in sources()]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building hydrobem_amd")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
