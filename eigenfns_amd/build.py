"""Build libeigenfns_hip.so in-tree with hipcc for gfx950.

    python -m eigenfns_amd.build        # or __graft_entry__.build()

The library is a plain C-ABI shared object (include/eigenfns_hip.h) loaded by
eigenfns_amd._lib through ctypes; no torch headers are involved.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libeigenfns_hip.so"
SOURCES = [CSRC / "capi.hip"]
HEADERS = sorted(CSRC.glob("*.hpp")) + [ROOT / "include" / "eigenfns_hip.h"]

ARCH = os.environ.get("EFN_OFFLOAD_ARCH", "gfx950")
# Packed FP32 instructions are disabled for all device code: some of their
# forms have given wrong results on this hardware without an error.
HIPCC_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
# rocSOLVER's batched Hermitian eigensolver (small dense FP64 algebra) and the
# rocBLAS handle it runs on
LINK_FLAGS = ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build libeigenfns_hip.so)")


def build(verbose: bool = True, extra_flags=()) -> Path:
    """Compile the library (always recompiles); returns its path."""
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, *extra_flags, "-o", str(tmp), *map(str, SOURCES), *LINK_FLAGS]
    if verbose:
        print("[eigenfns_amd.build]", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, cwd=str(ROOT), capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with exit code {res.returncode}")
    os.replace(tmp, LIB)
    if verbose:
        warn = [l for l in res.stderr.splitlines() if "warning" in l and "packed-fp32" not in l]
        for line in warn[:20]:
            print(line)
        print(f"[eigenfns_amd.build] wrote {LIB} ({LIB.stat().st_size / 1e6:.1f} MB)", flush=True)
    return LIB


if __name__ == "__main__":
    build()
