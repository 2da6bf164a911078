"""The C-ABI library loads (without a GPU) and exports every entry point
include/vc_abi.h declares; the build compiles on this host."""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "vc_abi.h").read_text()
    return sorted(set(re.findall(r"^(?:int|uint64_t|uint32_t|const char \*)\s*(vc_\w+)\s*\(", text,
                                 re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("vc_v1_step", "vc_v1_run", "vc_v1_rec_apply", "vc_lgn_precompute",
              "vc_lgn_frame_response", "vc_lgn_steps", "vc_lgn_init", "vc_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol(pkg):
    from visual_cortex_amd import abi
    assert abi.LIB_PATH.exists(), "libvc_hip.so not built"
    lib = ctypes.CDLL(str(abi.LIB_PATH))
    for n in declared_functions():
        assert hasattr(lib, n), n
    assert set(declared_functions()) <= set(abi.SIGNATURES), "abi.py misses a declared function"


def test_struct_sizes_match_header_layout(pkg):
    """ctypes mirrors have the C layout: compile a probe with the host compiler."""
    import subprocess
    import tempfile
    from visual_cortex_amd import abi
    src = r'''
#include <stdio.h>
#include "vc_abi.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %d\n", sizeof(vc_v1_params),
         sizeof(vc_v1_state), sizeof(vc_ff_conn), sizeof(vc_rec_conn), sizeof(vc_spikes),
         sizeof(vc_lgn_setup), sizeof(vc_lgn_params), sizeof(vc_lgn_state), sizeof(vc_v1_adjoint),
         sizeof(vc_v1_init), sizeof(vc_v1_spinit), VC_SNAP_ROWS);
  return 0; }
'''
    with tempfile.TemporaryDirectory() as d:
        c = Path(d) / "p.c"
        c.write_text(src)
        exe = Path(d) / "p"
        subprocess.run(["gcc", "-I", str(ROOT / "include"), "-o", str(exe), str(c)], check=True)
        sizes = list(map(int, subprocess.run([str(exe)], capture_output=True, text=True,
                                             check=True).stdout.split()))
    mine = [ctypes.sizeof(t) for t in (abi.V1Params, abi.V1State, abi.FFConn, abi.RecConn,
                                       abi.Spikes, abi.LgnSetup, abi.LgnParams, abi.LgnState,
                                       abi.V1Adjoint, abi.V1Init, abi.V1SpInit)] + [abi.VC_SNAP_ROWS]
    assert sizes == mine


def test_product_refuses_without_gpu(pkg):
    """The product path has no CPU fallback."""
    import pytest
    import torch
    from visual_cortex_amd import abi
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    abi._LIB = None
    with pytest.raises(abi.VCError):
        abi.load_library()
