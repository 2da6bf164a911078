"""Compile csrc/*.hip into libvc_hip.so for gfx950 (in-tree, so the library
travels with the source tree)."""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
SRC = [PKG / "csrc" / f for f in ("vc_abi.hip", "vc_lgn.hip", "vc_lgn_adjoint.hip", "vc_v1.hip",
                               "vc_v1_adjoint.hip", "vc_v1_delay.hip")]
OUT = PKG / "libvc_hip.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off on host and device: the kernels and the CPU oracle must
# evaluate the same binary32 operation sequences (no fused multiply-add).
# Packed FP32 instructions are disabled (a precaution on this hardware pool).
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-ffp-contract=off", "-fno-fast-math",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose: bool = False) -> Path:
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, *FLAGS, "-o", str(tmp), *map(str, SRC)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    if verbose and r.stderr.strip():
        print(r.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
