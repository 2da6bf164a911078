"""visual-cortex_amd: an MI355X (gfx950) implementation of the patchV1
simulation step -- LGN convolution and spike draw, LGN->V1 and V1->V1
conductance gathers, conductance-based AdEx/LIF RK2 update and spike
bookkeeping -- as hand-written HIP kernels behind a C ABI (include/vc_abi.h),
driven from Python with PyTorch-ROCm owning device memory.

Modules:
  abi      ctypes mirror of include/vc_abi.h, loader of libvc_hip.so
  network  model specification and synthetic network / stimulus generation
  sim      the GPU simulation driver (GpuSim)
  config   the reference's .cfg format (Boost program_options semantics)
  build    compiles csrc/*.hip into libvc_hip.so for gfx950
"""
from . import network  # noqa: F401

__all__ = ["network"]
