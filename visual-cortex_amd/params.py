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
    q.denorm = tm["denorm"]
    q.dt = sp.dt
    q.convolRatio = sp.convolRatio
    q.frRatio = sp.frRatioLGN
    q.tRef = 0.0
    q.seed_lo = sp.seed & 0xffffffff
    q.seed_hi = (sp.seed >> 32) & 0xffffffff
    q.trial = sp.trial
    return q


def initial_state(net: Network) -> dict:
    """Initial V1 state (host arrays).  With the reference's initial
    distributions in the spec (cfg keys w0, gFF0, gE0, gI0, use_v0) the
    reference's own draws (vc_v1_initial_state); otherwise the synthetic
    state: v uniform in [vR, vR + 0.8 (vT - vR)), w, tBack, g = 0.  A batched
    network concatenates its trials' own initial states."""
    if getattr(net, "batch", None) is not None:
        parts = []
        for sp_t in net.trial_specs:
            if sp_t.init_gFF0 is not None:
                raise NotImplementedError("the reference's initial-state draws with batched "
                                          "trials are not built")
            parts.append(_initial_state_spec(sp_t))
        return {k: np.concatenate([q[k] for q in parts], axis=-1) for k in parts[0]}
    if net.spec.init_gFF0 is not None:
        return reference_initial_state(net)
    return _initial_state_spec(net.spec)


def init_struct(sp) -> abi.V1Init:
    """vc_v1_init of a spec holding the reference's initial distributions."""
    q = abi.V1Init()
    q.nV1, q.nblock, q.blockSize, q.nType = sp.nV1, sp.nblock, sp.blockSize, 2
    q.typeAcc[0], q.typeAcc[1] = sp.nE, sp.blockSize
    q.ngFF, q.ngE, q.ngI = len(sp.riseTimeFF), len(sp.riseTimeE), len(sp.riseTimeI)
    q.iModel = sp.iModel
    q.useV0 = 1 if sp.use_v0 else 0
    for t in range(2):
        q.vR[t], q.vT[t] = sp.vR[t], sp.vT[t]
    for name in ("w0", "gFF0", "gE0", "gI0"):
        vals = getattr(sp, "init_" + name)
        if vals is not None:
            arr = getattr(q, name)
            for i, x in enumerate(vals):
                arr[i] = x
    return q


def reference_initial_state(net: Network) -> dict:
    """The reference's initial V1 state (src/patch.cu:4405-4570), drawn by the
    library's host function with the reference's generator and order, seeded
    as the reference seeds it when the LGN parameters come from fLGN: the
    cfg seed + 1 (logRand_init) + nV1 (the per-neuron engines),
    src/patch.cu:2686, 4224-4228, 4406."""
    sp = net.spec
    lib = abi.load_library(require_gpu=False)
    n = sp.nV1
    ngFF, ngE, ngI = len(sp.riseTimeFF), len(sp.riseTimeE), len(sp.riseTimeI)
    z = lambda *s: np.zeros(s, np.float32)
    st = dict(v=z(n), w=z(n), tBack=z(n), gFF=z(ngFF, n), hFF=z(ngFF, n), gE=z(ngE, n),
              hE=z(ngE, n), gI=z(ngI, n), hI=z(ngI, n))
    tonic = np.ascontiguousarray(net.tonicDep, np.float32)
    nLGN = np.ascontiguousarray(net.nLGN_of, np.uint32)
    q = init_struct(sp)
    seed = (int(sp.seed) + 1 + n) & 0xffffffffffffffff
    P = lambda a: a.ctypes.data
    abi.check(lib.vc_v1_initial_state(seed, q, P(tonic), P(nLGN), P(st["v"]), P(st["w"]),
                                      P(st["gFF"]), P(st["hFF"]), P(st["gE"]), P(st["hE"]),
                                      P(st["gI"]), P(st["hI"])), "vc_v1_initial_state")
    return st


def spinit_struct(net: Network, depth_ref: int) -> abi.V1SpInit:
    """vc_v1_spinit: the cfg's initial spike rates as chances per step
    (rate * dt / 1000, src/patch.cu:811-837), [type][simple, complex]."""
    sp = net.spec
    q = abi.V1SpInit()
    rates = list(sp.spE0) + list(sp.spI0)
    for i, r in enumerate(rates):
        q.sp0[i] = float(np.float32(r) * (np.float32(sp.dt) / np.float32(1000.0)))
    q.SCsplit = sp.SCsplit
    q.depthRef = depth_ref
    return q


def has_initial_spikes(sp) -> bool:
    return any(x > 0 for x in tuple(sp.spE0) + tuple(sp.spI0))


def reference_train_depth(net: Network) -> int:
    """The reference's trainDepth, ceil(maxDistance / speedOfThought / dt) + 1
    over the connected near pairs (src/patch.cu:3807-3813), whatever noDelay
    is: rand_spInit fills that many steps."""
    from .network import train_depth
    sp = net.spec
    nnz = net.n_rec_synapses_csr()
    dmax = float(net.rec_distances()[:nnz].max()) if nnz else 0.0
    return train_depth(dmax, sp.speedOfThought, sp.dt)


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


def learn_ff_struct(net: Network) -> abi.LearnFFParams:
    """The FF_E rule's constants (learnFF_pre / learnFF_post, src/condShape.h
    :94-155, with the E entries, src/patch.cu:1288-1330), in the reference's
    float32 / float64 operation order."""
    sp = net.spec
    f = np.float32
    ratio = f(sp.sRatioLGN[0])
    lp = abi.LearnFFParams()
    lp.tauLTP, lp.tauLTD, lp.tauTrip = f(sp.tauLTP), f(sp.tauLTD), f(sp.tauTrip)
    lp.tauAvg = f(sp.tauAvg[0])
    A_LTP = f(f(sp.A_LGN) * ratio)
    lp.A_LTP = A_LTP
    A_ratio = f(f(sp.r_LTD) * A_LTP)                     # = A_LTD
    target = f(sp.targetFR[0])
    if target > 0:
        # tauTrip/(tauLTD*(tauAvg_s*tauAvg_s)) in float, then /1000.0/targetFR
        # in double, the product rounded once to float
        ts = f(float(f(sp.tauAvg[0])) / 1000.0)
        x = f(f(sp.tauTrip) / f(f(sp.tauLTD) * f(ts * ts)))
        A_ratio = f(float(A_ratio) * (float(x) / 1000.0 / float(target)))
    lp.A_ratio = A_ratio
    lp.targetFR = target
    lp.gmax = f(f(sp.gmaxLGN) * ratio)
    lp.gmin = f(f(sp.gminLGN) * ratio)
    lp.applyHomeo = int(sp.applyHomeo)
    lp.symmetricHomeo = int(sp.symmetricHomeo)
    # exp_homeo = exponential(-dt/hdFF) on the host (src/patch.cu:1330)
    lp.expHomeo = f(math.exp(float(f(-f(sp.dt) / f(sp.decayTimeHomeoFF)))))
    lp.rebound = int(sp.rebound)
    return lp


def learn_ff_totals(net: Network) -> tuple[np.ndarray, np.ndarray]:
    """(totalFF, totalFF_inf) [nV1]: each neuron's total LGN strength summed
    in float32 in list order, and FF_InfRatio x it (src/patch.cu:4336-4349)."""
    sp = net.spec
    s = np.asarray(net.sLGN, np.float32)
    n = np.asarray(net.nLGN_of, np.int64)
    tot = np.zeros(sp.nV1, np.float32)
    for j in range(s.shape[0]):
        tot = np.where(j < n, tot + s[j], tot).astype(np.float32)
    return tot, (tot * np.float32(sp.FF_InfRatio)).astype(np.float32)
