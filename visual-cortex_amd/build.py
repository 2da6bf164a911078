"""Compile csrc/*.hip into libvc_hip.so for gfx950 (in-tree, so the library
travels with the source tree).  Each source is compiled to an object in its
own hipcc process, in parallel, then the objects are linked."""
from __future__ import annotations

import os
import subprocess
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
SRC = [PKG / "csrc" / f for f in ("vc_abi.hip", "vc_lgn.hip", "vc_lgn_adjoint.hip", "vc_v1.hip",
                               "vc_v1_adjoint.hip", "vc_v1_delay.hip", "vc_v1_init.hip",
                               "vc_v1_learn.hip")]
OUT = PKG / "libvc_hip.so"
OBJ_DIR = PKG / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off on host and device: the kernels and the CPU oracle must
# evaluate the same binary32 operation sequences (no fused multiply-add).
# Packed FP32 instructions are disabled (a precaution on this hardware pool).
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-ffp-contract=off", "-fno-fast-math",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _compile(src: Path) -> Path:
    obj = OBJ_DIR / (src.stem + ".o")
    tmp = obj.with_suffix(".o.tmp")
    r = subprocess.run([HIPCC, *FLAGS, "-c", "-o", str(tmp), str(src)],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n" + r.stdout + r.stderr)
    os.replace(tmp, obj)
    return obj


def build(verbose: bool = False) -> Path:
    OBJ_DIR.mkdir(exist_ok=True)
    jobs = min(len(SRC), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, SRC))
    tmp = OUT.with_suffix(".so.tmp")
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp),
                        *map(str, objs)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    if verbose and r.stderr.strip():
        print(r.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
