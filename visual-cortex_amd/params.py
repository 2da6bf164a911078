"""Packing of a Network's parameters into the C-ABI structs of
include/vc_abi.h, and the initial V1 state.  No device code here."""
from __future__ import annotations

import math

import numpy as np

from . import abi
from .network import Network, timing


def v1_params_struct(net: Network) -> abi.V1Params:
    sp = net.spec
    p = abi.V1Params()
    p.nV1, p.nblock, p.blockSize, p.nE = sp.nV1, sp.nblock, sp.blockSize, sp.nE
    p.nType = 2
    p.typeAcc[0], p.typeAcc[1] = sp.nE, sp.blockSize
    p.ngFF, p.ngE, p.ngI = len(sp.riseTimeFF), len(sp.riseTimeE), len(sp.riseTimeI)
    p.iModel = sp.iModel
    p.dt = sp.dt
    for t in range(2):
        for name in ("vR", "vThres", "gL", "C", "tRef", "vT", "deltaT", "tau_w", "a", "b"):
            getattr(p, name)[t] = getattr(sp, name)[t]
        for ig in range(p.ngFF):
            p.pFF[t * p.ngFF + ig] = sp.pFF[t][ig]
            p.pBG[t * p.ngFF + ig] = sp.pBG[t][ig]
        for ig in range(p.ngE):
            p.pE[t * p.ngE + ig] = sp.pE[t][ig]
        for ig in range(p.ngI):
            p.pI[t * p.ngI + ig] = sp.pI[t][ig]
        lam = np.float32(sp.bgRate * sp.dt / 1000.0)
        p.bgLambda[t] = float(lam)
        p.bgP0[t] = float(np.float32(math.exp(-float(lam))))
        p.bgStrength[t] = sp.bgStrength[t]
        p.noisyDep[t] = sp.noisyDep[t]
        p.synFailFF[t] = sp.synFailFF[t]
        p.synPerConFF[t] = sp.synPerConFF[t]
        for u in range(2):
            p.synFail[t * 2 + u] = sp.synFail[t * 2 + u]
            p.synPerCon[t * 2 + u] = sp.synPerCon[t * 2 + u]
    p.tauNoise = sp.tau_noise
    p.noiseOnTonic = 1 if sp.noiseOnTonic else 0

    def shape(prefix, rise, decay):
        r = getattr(p, "rise" + prefix)
        d = getattr(p, "decay" + prefix)
        o = getattr(p, "dod" + prefix)
        for ig in range(abi.VC_MAX_NG):
            rr = np.float32(rise[min(ig, len(rise) - 1)])
            dd = np.float32(decay[min(ig, len(decay) - 1)])
            r[ig], d[ig] = float(rr), float(dd)
            o[ig] = float(rr / (dd - rr))          # ConductanceShape dod, float32
    shape("FF", sp.riseTimeFF, sp.decayTimeFF)
    shape("E", sp.riseTimeE, sp.decayTimeE)
    shape("I", sp.riseTimeI, sp.decayTimeI)
    p.seed_lo = sp.seed & 0xffffffff
    p.seed_hi = (sp.seed >> 32) & 0xffffffff
    p.trial = sp.trial
    return p


def lgn_params_struct(net: Network) -> abi.LgnParams:
    sp = net.spec
    tm = timing(sp)
    q = abi.LgnParams()
    q.nLGN = sp.nLGN
    q.nKernelSample = sp.nKernelSample
    q.kernelSampleInterval = tm["kernelSampleInterval"]
    q.iKernelSampleT0 = tm["iKernelSampleT0"]
    q.nRetrace = tm["nRetrace"]
    q.ntPerFrame = tm["ntPerFrame"]
    q.maxFrame = tm["maxFrame"]
    q.dt = sp.dt
    q.convolRatio = sp.convolRatio
    q.frRatio = sp.frRatioLGN
    q.tRef = 0.0
    q.seed_lo = sp.seed & 0xffffffff
    q.seed_hi = (sp.seed >> 32) & 0xffffffff
    q.trial = sp.trial
    return q


def initial_state(net: Network) -> dict:
    """Initial V1 state (host arrays): v uniform in [vR, vT), w, tBack, g = 0.
    A batched network concatenates its trials' own initial states."""
    if getattr(net, "batch", None) is not None:
        parts = []
        for sp_t in net.trial_specs:
            parts.append(_initial_state_spec(sp_t))
        return {k: np.concatenate([q[k] for q in parts], axis=-1) for k in parts[0]}
    return _initial_state_spec(net.spec)


def _initial_state_spec(sp) -> dict:
    rng = np.random.default_rng(sp.seed + 7)
    local = np.arange(sp.nV1) % sp.blockSize
    t = (local >= sp.nE).astype(int)
    vR = np.array(sp.vR)[t]
    vT = np.array(sp.vT)[t]
    v = (vR + (vT - vR) * rng.random(sp.nV1) * 0.8).astype(np.float32)
    ngFF, ngE, ngI = len(sp.riseTimeFF), len(sp.riseTimeE), len(sp.riseTimeI)
    z = lambda *s: np.zeros(s, np.float32)
    return dict(v=v, w=z(sp.nV1), tBack=z(sp.nV1), gFF=z(ngFF, sp.nV1), hFF=z(ngFF, sp.nV1),
                gE=z(ngE, sp.nV1), hE=z(ngE, sp.nV1), gI=z(ngI, sp.nV1), hI=z(ngI, sp.nV1))
