"""ctypes mirror of include/vc_abi.h and the loader of libvc_hip.so.

The product path loads only libvc_hip.so (built from csrc/ by build.py).
If the library is missing, or no GPU is visible, load_library() raises: there
is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libvc_hip.so"

VC_MAX_NTYPE = 2
VC_MAX_NG = 2
VC_OK, VC_E_INVALID, VC_E_UNSUPPORTED, VC_E_HIP = 0, -1, -2, -3
VC_LGN_GEOM = 8            # floats of one region's sample geometry (vc_lgn_sample_geometry)
VC_SNAP_ROWS = 14          # snapshot rows: v, tBack, gFF[2], hFF[2], gE[2], hE[2], gI[2], hI[2]

f32 = C.c_float
u32 = C.c_uint32
i32 = C.c_int32
vp = C.c_void_p
fp = C.POINTER(C.c_float)


def _arr(t, n):
    return t * n


class V1Params(C.Structure):
    _fields_ = [
        ("nV1", u32), ("nblock", u32), ("blockSize", u32), ("nE", u32),
        ("nType", u32), ("typeAcc", _arr(u32, VC_MAX_NTYPE)),
        ("ngFF", u32), ("ngE", u32), ("ngI", u32),
        ("iModel", i32), ("dt", f32),
        ("vR", _arr(f32, VC_MAX_NTYPE)), ("vThres", _arr(f32, VC_MAX_NTYPE)),
        ("gL", _arr(f32, VC_MAX_NTYPE)), ("C", _arr(f32, VC_MAX_NTYPE)),
        ("tRef", _arr(f32, VC_MAX_NTYPE)), ("vT", _arr(f32, VC_MAX_NTYPE)),
        ("deltaT", _arr(f32, VC_MAX_NTYPE)),
        ("tau_w", _arr(f32, VC_MAX_NTYPE)), ("a", _arr(f32, VC_MAX_NTYPE)),
        ("b", _arr(f32, VC_MAX_NTYPE)),
        ("pFF", _arr(f32, VC_MAX_NTYPE * VC_MAX_NG)),
        ("pE", _arr(f32, VC_MAX_NTYPE * VC_MAX_NG)),
        ("pI", _arr(f32, VC_MAX_NTYPE * VC_MAX_NG)),
        ("riseFF", _arr(f32, VC_MAX_NG)), ("decayFF", _arr(f32, VC_MAX_NG)), ("dodFF", _arr(f32, VC_MAX_NG)),
        ("riseE", _arr(f32, VC_MAX_NG)), ("decayE", _arr(f32, VC_MAX_NG)), ("dodE", _arr(f32, VC_MAX_NG)),
        ("riseI", _arr(f32, VC_MAX_NG)), ("decayI", _arr(f32, VC_MAX_NG)), ("dodI", _arr(f32, VC_MAX_NG)),
        ("bgStrength", _arr(f32, VC_MAX_NTYPE)), ("bgLambda", _arr(f32, VC_MAX_NTYPE)),
        ("bgP0", _arr(f32, VC_MAX_NTYPE)),
        ("pBG", _arr(f32, VC_MAX_NTYPE * VC_MAX_NG)),
        ("seed_lo", u32), ("seed_hi", u32), ("trial", u32),
        ("blockTrial", vp), ("trialV1Off", vp), ("trialSeed", vp),
        ("stepWG", u32), ("gatherTile", u32),
        ("synFailFF", _arr(f32, VC_MAX_NTYPE)), ("synPerConFF", _arr(f32, VC_MAX_NTYPE)),
        ("synFail", _arr(f32, VC_MAX_NTYPE * VC_MAX_NTYPE)),
        ("synPerCon", _arr(f32, VC_MAX_NTYPE * VC_MAX_NTYPE)),
        ("noisyDep", _arr(f32, VC_MAX_NTYPE)), ("tauNoise", f32), ("noiseOnTonic", u32),
    ]


class V1State(C.Structure):
    _fields_ = [("v", vp), ("w", vp), ("tBack", vp), ("gFF", vp), ("hFF", vp),
                ("gE", vp), ("hE", vp), ("gI", vp), ("hI", vp), ("tonicDep", vp),
                ("gapS", vp), ("gap", vp), ("lastNoise", vp)]


class FFConn(C.Structure):
    _fields_ = [("nLGN", vp), ("sLGN", vp), ("idLGN", vp), ("maxLGN", u32)]


class FFEvents(C.Structure):
    _fields_ = [("revOff", vp), ("revTgt", vp), ("mask", vp), ("nV1", u32)]


class RecConn(C.Structure):
    _fields_ = [("maxNear", u32), ("nNear", vp), ("nbr", vp), ("rowMask", vp),
                ("rowOff", vp), ("w", vp), ("wDense", vp),
                ("maxGap", u32), ("nGap", vp), ("gapPre", vp), ("gapG", vp)]


class FarConn(C.Structure):
    _fields_ = [("maxFar", u32), ("nFar", vp), ("pre", vp), ("w", vp), ("dist", vp),
                ("preMask", vp), ("outOff", vp), ("outPost", vp), ("outJ", vp), ("hit", vp)]


class Spikes(C.Structure):
    _fields_ = [("mask", vp), ("og", vp), ("oh", vp)]


class V1Init(C.Structure):
    _fields_ = [
        ("nV1", u32), ("nblock", u32), ("blockSize", u32), ("nType", u32),
        ("typeAcc", _arr(u32, VC_MAX_NTYPE)), ("ngFF", u32), ("ngE", u32), ("ngI", u32),
        ("iModel", i32), ("useV0", u32),
        ("vR", _arr(f32, VC_MAX_NTYPE)), ("vT", _arr(f32, VC_MAX_NTYPE)),
        ("w0", _arr(f32, VC_MAX_NTYPE * 2)),
        ("gFF0", _arr(f32, VC_MAX_NTYPE * VC_MAX_NG * 4)),
        ("gE0", _arr(f32, VC_MAX_NTYPE * VC_MAX_NG * 4)),
        ("gI0", _arr(f32, VC_MAX_NTYPE * VC_MAX_NG * 4)),
    ]


class V1SpInit(C.Structure):
    _fields_ = [("sp0", _arr(f32, VC_MAX_NTYPE * 2)), ("SCsplit", u32), ("depthRef", u32)]


class LearnFFParams(C.Structure):
    _fields_ = [("tauLTP", f32), ("tauLTD", f32), ("tauTrip", f32), ("tauAvg", f32),
                ("A_LTP", f32), ("A_ratio", f32), ("targetFR", f32), ("gmax", f32),
                ("gmin", f32), ("applyHomeo", u32), ("symmetricHomeo", u32),
                ("expHomeo", f32), ("rebound", u32)]


class LearnFFState(C.Structure):
    _fields_ = [("sLGN", vp), ("vLTD", vp), ("vTrip", vp), ("vAvg", vp), ("totalFF", vp),
                ("totalFFinf", vp)]


class LgnSetup(C.Structure):
    _fields_ = [
        ("nLGN", u32), ("nSample1D", u32), ("nKernelSample", u32),
        ("kernelSampleInterval", u32), ("iKernelSampleT0", u32), ("maxSample1D", u32),
        ("dt", f32), ("nsig", f32), ("normViewDistance", f32), ("x0", f32), ("y0", f32),
        ("polar", vp), ("ecc", vp), ("rw", vp), ("rh", vp), ("orient", vp), ("k", vp),
        ("tauR", vp), ("tauD", vp), ("delay", vp), ("ratio", vp), ("nR", vp), ("nD", vp),
        ("covariant", vp),
    ]


class LgnParams(C.Structure):
    _fields_ = [
        ("nLGN", u32), ("nKernelSample", u32), ("kernelSampleInterval", u32),
        ("iKernelSampleT0", u32), ("nRetrace", u32), ("ntPerFrame", u32), ("maxFrame", u32),
        ("dt", f32), ("convolRatio", f32), ("frRatio", f32), ("tRef", f32),
        ("seed_lo", u32), ("seed_hi", u32), ("trial", u32),
        ("cellTrial", vp), ("trialLgnOff", vp), ("trialSeed", vp), ("denorm", u32),
    ]


class LgnState(C.Structure):
    _fields_ = [("TW", vp), ("kC", vp), ("kS", vp), ("maxConvol", vp),
                ("c50", vp), ("sharpness", vp), ("logA", vp), ("logB", vp),
                ("leftTimeRate", vp), ("lastNegLogRand", vp)]


class V1Adjoint(C.Structure):
    _fields_ = [("snap", vp), ("spk", vp), ("sInfo", vp), ("lgnStride", C.c_uint64),
                ("lgnSInfo", vp), ("gOut", vp), ("lam", vp), ("lamTsp", vp), ("dsLGN", vp),
                ("dw", vp)]


# exported symbols (name -> argtypes); tests check that the .so exports each
SIGNATURES = {
    "vc_v1_step": [vp, C.POINTER(V1Params), C.POINTER(V1State), C.POINTER(FFConn),
                   C.POINTER(RecConn), C.POINTER(Spikes), C.POINTER(Spikes), vp, vp, u32,
                   vp, vp, vp],
    "vc_v1_run": [vp, C.POINTER(V1Params), C.POINTER(V1State), C.POINTER(FFConn),
                  C.POINTER(RecConn), C.POINTER(Spikes), u32, vp, C.c_uint64, vp, C.c_uint64,
                  u32, u32, vp, C.c_uint64],
    "vc_v1_run_persistent": [vp, C.POINTER(V1Params), C.POINTER(V1State), C.POINTER(FFConn),
                             C.POINTER(RecConn), C.POINTER(Spikes), u32, vp, C.c_uint64, vp,
                             C.c_uint64, u32, u32, vp, C.c_uint64, vp, u32, vp],
    "vc_v1_persistent_scratch_bytes": [C.POINTER(V1Params)],
    "vc_v1_run_persistent_noalloc": [vp, C.POINTER(V1Params), C.POINTER(V1State),
                                     C.POINTER(FFConn), C.POINTER(RecConn), C.POINTER(Spikes),
                                     u32, vp, C.c_uint64, vp, C.c_uint64, u32, u32, vp,
                                     C.c_uint64, vp],
    "vc_v1_run_persistent_noalloc_max_v1": [],
    "vc_v1_dep_current": [vp, C.POINTER(V1Params), C.POINTER(V1State), vp],
    "vc_v1_gap": [vp, C.POINTER(V1Params), C.POINTER(V1State), C.POINTER(RecConn), vp, u32,
                  u32, vp],
    "vc_v1_rec_apply": [vp, C.POINTER(V1Params), C.POINTER(V1State), C.POINTER(RecConn),
                        C.POINTER(Spikes)],
    "vc_v1_rec_apply_at": [vp, C.POINTER(V1Params), C.POINTER(V1State), C.POINTER(RecConn),
                           C.POINTER(Spikes), u32],
    "vc_spikes_clear": [vp, u32, C.POINTER(Spikes)],
    "vc_v1_initial_state": [C.c_uint64, C.POINTER(V1Init), vp, vp, vp, vp, vp, vp, vp, vp,
                            vp, vp],
    "vc_v1_initial_spikes": [vp, C.POINTER(V1Params), C.POINTER(V1State), vp,
                             C.POINTER(V1SpInit), vp, vp, u32, u32],
    "vc_spikes_compact": [vp, C.POINTER(V1Params), vp, C.POINTER(Spikes)],
    "vc_v1_rec_distances": [vp, C.POINTER(V1Params), C.POINTER(RecConn), vp, vp],
    "vc_v1_rec_delay_table": [vp, C.c_uint64, vp, vp, C.c_float, C.c_float, vp],
    "vc_v1_rec_delay": [vp, C.POINTER(V1Params), C.POINTER(V1State), C.POINTER(RecConn), vp,
                        vp, vp, u32, u32, u32, vp],
    "vc_v1_delay_run": [vp, C.POINTER(V1Params), C.POINTER(V1State), C.POINTER(FFConn),
                        C.POINTER(RecConn), vp, C.POINTER(FarConn), vp, vp, u32, u32, u32, vp,
                        C.c_uint64, vp, C.c_uint64, u32, u32, C.c_float, vp, vp, vp, vp,
                        C.c_uint64, vp],
    "vc_v1_rec_far": [vp, C.POINTER(V1Params), C.POINTER(V1State), C.POINTER(FarConn), vp,
                      vp, u32, u32, C.c_float, vp],
    "vc_v1_step_snapshot": [vp, C.POINTER(V1Params), C.POINTER(V1State), C.POINTER(FFConn),
                            C.POINTER(RecConn), C.POINTER(Spikes), C.POINTER(Spikes), vp, vp,
                            u32, vp, vp],
    "vc_v1_adjoint_run": [vp, C.POINTER(V1Params), C.POINTER(V1State), C.POINTER(FFConn),
                          C.POINTER(RecConn), C.POINTER(V1Adjoint), u32, u32],
    "vc_lgn_precompute": [vp, C.POINTER(LgnSetup), vp, vp, vp, vp, vp],
    "vc_lgn_frame_response": [vp, vp, u32, u32, u32, u32, vp, vp, vp, vp, vp, C.c_uint64, vp],
    "vc_lgn_frame_response_cells": [vp, vp, u32, u32, u32, u32, u32, u32, vp, vp, vp, vp, vp,
                                    C.c_uint64, vp],
    "vc_lgn_steps": [vp, C.POINTER(LgnParams), C.POINTER(LgnState), vp, u32, u32, vp, vp,
                     C.POINTER(FFEvents)],
    "vc_lgn_steps_learn": [vp, C.POINTER(LgnParams), C.POINTER(LgnState), vp, u32, u32, vp,
                           vp, C.POINTER(FFEvents), C.c_float, vp, vp, vp, vp],
    "vc_v1_learn_ff": [vp, C.POINTER(V1Params), C.POINTER(LearnFFParams), C.POINTER(FFConn),
                       C.POINTER(LearnFFState), vp, vp, vp, vp, vp, u32],
    "vc_lgn_init": [vp, C.POINTER(LgnParams), vp, vp],
    "vc_lgn_frame_responses": [vp, vp, u32, u32, u32, u32, u32, vp, vp, vp, vp, vp],
    "vc_lgn_sample_geometry": [vp, C.POINTER(LgnSetup), vp],
    "vc_lgn_frame_response_geom": [vp, vp, u32, u32, u32, u32, u32, u32, vp, C.c_float,
                                   C.c_float, C.c_float, vp, vp, vp, C.c_uint64, vp],
    "vc_lgn_temporal_weights": [vp, C.POINTER(LgnSetup), vp],
    "vc_lgn_rates": [vp, C.POINTER(LgnParams), C.POINTER(LgnState), vp, u32, u32, vp],
    "vc_lgn_rates_adjoint": [vp, C.POINTER(LgnParams), C.POINTER(LgnState), vp, u32, u32, vp,
                             vp, vp, vp],
    "vc_lgn_frame_responses_adjoint": [vp, vp, u32, u32, u32, u32, u32, vp, vp, vp, vp, vp,
                                       vp],
    "vc_lgn_temporal_adjoint": [vp, C.POINTER(LgnSetup), vp, vp, vp],
    "vc_last_error": [],
    "vc_version": [],
}


class VCError(RuntimeError):
    pass


_LIB = None


def load_library(require_gpu: bool = True):
    """Load libvc_hip.so.  Raises if it is missing (build it first) or, when
    require_gpu, if torch sees no GPU: the product has no CPU path."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise VCError(f"{LIB_PATH} not built: run __graft_entry__.build() first")
    if require_gpu:
        import torch
        if not torch.cuda.is_available():
            raise VCError("no GPU visible: the HIP path has no CPU fallback")
    lib = C.CDLL(str(LIB_PATH))
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = (C.c_char_p if name == "vc_last_error" else
                      C.c_uint64 if name.endswith("_bytes") else
                      C.c_uint32 if name.endswith("_max_v1") else C.c_int)
    _LIB = lib
    return lib


def check(rc: int, what: str):
    if rc != 0:
        msg = _LIB.vc_last_error().decode() if _LIB is not None else ""
        raise VCError(f"{what} failed ({rc}): {msg}")


def ptr(t) -> int | None:
    """Device (or host) address of a torch tensor, None for None."""
    if t is None:
        return None
    return t.data_ptr()
