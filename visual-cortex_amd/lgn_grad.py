"""Differentiable LGN stage: the LGN rates of a run of steps as a function of
the stimulus frames and six parameters per cell, with the forward and the
reverse mode both in libvc_hip.so (HIP kernels, no torch autograd tracing).

    fr = LgnStage(spec, lgn, nsteps).rates(frames, theta)    # [nsteps, nLGN] Hz

frames [nF, 3, H, W] float32 (nF = ceil(nsteps / ntPerFrame)) are the stimulus
frames in the reference's layout; frame k is on screen for steps
[k ntPerFrame, (k+1) ntPerFrame) and earlier frames are the uniform
pre-stimulus frame.  theta [nLGN, 6] float32 holds, per cell, THETA_NAMES:
the centre and surround gains of the difference-of-Gaussians and the rise
and decay time constants of the centre and surround temporal kernels
(src/discrete_input_convol.cu:14-33, 230-305 temporal kernel;
:1003-1340 convolution; :1637-1680 and src/LGN_props.cuh:94-100 rate).

The normalisation maxConvol is computed once from the initial parameters
(as the reference does at set-up) and held constant.
"""
from __future__ import annotations

import ctypes as C
import math
from types import SimpleNamespace

import numpy as np
import torch

from . import abi
from .abi import check, ptr
from .network import ModelSpec, max_sample_1d, timing, view_geometry
from .params import lgn_params_struct

THETA_NAMES = ("kC", "kS", "tauR_C", "tauD_C", "tauR_S", "tauD_S")


class LgnStage:
    """Set-up of the LGN layer `lgn` (a network.lgn_parameters dict) for
    runs of `nsteps` steps on `device`."""

    def __init__(self, spec: ModelSpec, lgn: dict, nsteps: int, device="cuda",
                 init_lms=(0.5, 0.5, 0.5)):
        self.lib = abi.load_library()
        self.spec, self.T = spec, int(nsteps)
        self.dev = torch.device(device)
        if self.dev.type == "cuda" and self.dev.index is None:
            self.dev = torch.device("cuda", torch.cuda.current_device())
        d = self.dev
        tm = timing(spec)
        if tm["denorm"] != 1:
            raise NotImplementedError("the LGN adjoint needs frames of a whole number of steps "
                                      "(denorm 1)")
        self.tm = tm
        self.npf = tm["ntPerFrame"]
        self.nF = -(-self.T // self.npf)
        # frames before the first: far enough back for the first step's taps
        self.lead = -(-tm["nRetrace"] // self.npf) + 1
        self.t = {k: torch.as_tensor(np.ascontiguousarray(lgn[k], dtype=np.float32), device=d)
                  for k in ("polar", "ecc", "rw", "rh", "orient", "k", "tauR", "tauD", "delay",
                            "ratio", "nR", "nD", "covariant", "c50", "sharpness", "logA",
                            "logB")}
        self.coneType = torch.as_tensor(np.ascontiguousarray(lgn["coneType"], dtype=np.uint32)
                                        .view(np.int32), device=d)
        x0, y0, nvd = view_geometry(spec)
        q = abi.LgnSetup()
        q.nLGN, q.nSample1D, q.nKernelSample = spec.nLGN, spec.nSpatialSample1D, spec.nKernelSample
        q.kernelSampleInterval, q.iKernelSampleT0 = tm["kernelSampleInterval"], tm["iKernelSampleT0"]
        q.maxSample1D = max_sample_1d(SimpleNamespace(spec=spec, lgn=lgn))
        q.dt, q.nsig, q.normViewDistance, q.x0, q.y0 = spec.dt, spec.nsig, nvd, x0, y0
        for k in ("polar", "ecc", "rw", "rh", "orient", "k", "tauR", "tauD", "delay", "ratio",
                  "nR", "nD", "covariant"):
            setattr(q, k, ptr(self.t[k]))
        self.q = q
        nS = spec.nSpatialSample1D ** 2
        f32 = dict(dtype=torch.float32, device=d)
        self.TW0 = torch.zeros(spec.nLGN, 2, spec.nKernelSample, **f32)
        self.SW = torch.zeros(nS, **f32)
        self.SCx = torch.zeros(spec.nLGN, 2, nS, **f32)
        self.SCy = torch.zeros(spec.nLGN, 2, nS, **f32)
        self.maxConvol = torch.zeros(spec.nLGN, **f32)
        check(self.lib.vc_lgn_precompute(self._h(), C.byref(q), ptr(self.TW0), ptr(self.SW),
                                         ptr(self.SCx), ptr(self.SCy), ptr(self.maxConvol)),
              "vc_lgn_precompute")
        p = lgn_params_struct(SimpleNamespace(spec=spec))
        p.maxFrame = self.lead + self.nF          # every frame of the run is held: no wrap
        self.p = p
        init = torch.empty(3, spec.height, spec.width, **f32)
        for ch in range(3):
            init[ch] = float(np.float32(init_lms[ch]))
        self.R_init = torch.zeros(1, spec.nLGN, 2, **f32)
        check(self.lib.vc_lgn_frame_responses(self._h(), ptr(init), 1, spec.width, spec.height,
                                              spec.nLGN, spec.nSpatialSample1D, ptr(self.SCx),
                                              ptr(self.SCy), ptr(self.SW), ptr(self.coneType),
                                              ptr(self.R_init)),
              "vc_lgn_frame_responses")
        self.theta0 = torch.stack([self.t["k"][0], self.t["k"][1], self.t["tauR"][0],
                                   self.t["tauD"][0], self.t["tauR"][1], self.t["tauD"][1]], 1)

    def _h(self):
        return C.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)

    def rates(self, frames: torch.Tensor, theta: torch.Tensor) -> torch.Tensor:
        return LgnRates.apply(frames, theta, self)

    # ---------------------------------------------------------------- forward
    def _setup_theta(self, theta: torch.Tensor):
        """Temporal weights and the state struct for the parameters theta."""
        sp = self.spec
        tauR = torch.stack([theta[:, 2], theta[:, 4]]).contiguous()
        tauD = torch.stack([theta[:, 3], theta[:, 5]]).contiguous()
        kC = theta[:, 0].contiguous()
        kS = theta[:, 1].contiguous()
        q = abi.LgnSetup.from_buffer_copy(self.q)
        q.tauR, q.tauD = ptr(tauR), ptr(tauD)
        TW = torch.empty(sp.nLGN, 2, sp.nKernelSample, dtype=torch.float32, device=self.dev)
        check(self.lib.vc_lgn_temporal_weights(self._h(), C.byref(q), ptr(TW)),
              "vc_lgn_temporal_weights")
        s = abi.LgnState(ptr(TW), ptr(kC), ptr(kS), ptr(self.maxConvol), ptr(self.t["c50"]),
                         ptr(self.t["sharpness"]), ptr(self.t["logA"]), ptr(self.t["logB"]),
                         None, None)
        return dict(q=q, s=s, keep=(tauR, tauD, kC, kS, TW))

    def _check_inputs(self, frames, theta):
        sp = self.spec
        if frames.dtype != torch.float32 or frames.dim() != 4 or \
                tuple(frames.shape[1:]) != (3, sp.height, sp.width) or frames.shape[0] < self.nF:
            raise ValueError(f"frames must be float32 [>= {self.nF}, 3, {sp.height}, {sp.width}]")
        if theta.dtype != torch.float32 or tuple(theta.shape) != (sp.nLGN, 6):
            raise ValueError(f"theta must be float32 [{sp.nLGN}, 6]")
        if frames.device != self.dev or theta.device != self.dev:
            raise ValueError("frames and theta must be on the stage's device")

    def forward(self, frames: torch.Tensor, theta: torch.Tensor):
        sp = self.spec
        self._check_inputs(frames, theta)
        frames = frames[:self.nF].contiguous()
        th = self._setup_theta(theta)
        R = torch.empty(self.lead + self.nF, sp.nLGN, 2, dtype=torch.float32, device=self.dev)
        R[:self.lead] = self.R_init
        check(self.lib.vc_lgn_frame_responses(self._h(), ptr(frames), self.nF, sp.width,
                                              sp.height, sp.nLGN, sp.nSpatialSample1D,
                                              ptr(self.SCx), ptr(self.SCy), ptr(self.SW),
                                              ptr(self.coneType), ptr(R[self.lead:])),
              "vc_lgn_frame_responses")
        fr = torch.empty(self.T, sp.nLGN, dtype=torch.float32, device=self.dev)
        check(self.lib.vc_lgn_rates(self._h(), C.byref(self.p), C.byref(th["s"]), ptr(R),
                                    self.lead * self.npf, self.T, ptr(fr)), "vc_lgn_rates")
        return fr, R, th

    # --------------------------------------------------------------- backward
    def backward(self, frames: torch.Tensor, R: torch.Tensor, th: dict, gFr: torch.Tensor):
        sp = self.spec
        frames = frames[:self.nF].contiguous()
        gFr = gFr.contiguous().float()
        dR = torch.zeros_like(R)
        dTW = torch.zeros(sp.nLGN, 2, sp.nKernelSample, dtype=torch.float32, device=self.dev)
        dK = torch.zeros(sp.nLGN, 2, dtype=torch.float32, device=self.dev)
        check(self.lib.vc_lgn_rates_adjoint(self._h(), C.byref(self.p), C.byref(th["s"]),
                                            ptr(R), self.lead * self.npf, self.T, ptr(gFr),
                                            ptr(dR), ptr(dTW), ptr(dK)), "vc_lgn_rates_adjoint")
        dFrames = torch.zeros_like(frames)
        check(self.lib.vc_lgn_frame_responses_adjoint(
            self._h(), ptr(frames), self.nF, sp.width, sp.height, sp.nLGN, sp.nSpatialSample1D,
            ptr(self.SCx), ptr(self.SCy), ptr(self.SW), ptr(self.coneType),
            ptr(dR[self.lead:]), ptr(dFrames)), "vc_lgn_frame_responses_adjoint")
        dTauR = torch.empty(2, sp.nLGN, dtype=torch.float32, device=self.dev)
        dTauD = torch.empty(2, sp.nLGN, dtype=torch.float32, device=self.dev)
        check(self.lib.vc_lgn_temporal_adjoint(self._h(), C.byref(th["q"]), ptr(dTW),
                                               ptr(dTauR), ptr(dTauD)),
              "vc_lgn_temporal_adjoint")
        dTheta = torch.stack([dK[:, 0], dK[:, 1], dTauR[0], dTauD[0], dTauR[1], dTauD[1]], 1)
        return dFrames, dTheta


class LgnRates(torch.autograd.Function):
    """fr = rates(frames, theta); backward through the HIP adjoint kernels."""

    @staticmethod
    def forward(ctx, frames, theta, stage: LgnStage):
        fr, R, th = stage.forward(frames, theta)
        ctx.stage, ctx.th, ctx.nfr = stage, th, frames.shape[0]
        ctx.save_for_backward(frames, R)
        return fr

    @staticmethod
    def backward(ctx, gFr):
        frames, R = ctx.saved_tensors
        stage = ctx.stage
        dF, dTheta = stage.backward(frames, R, ctx.th, gFr)
        if ctx.nfr > stage.nF:                       # frames past the run get no gradient
            dF = torch.cat([dF, dF.new_zeros((ctx.nfr - stage.nF,) + tuple(dF.shape[1:]))])
        return dF, dTheta, None
