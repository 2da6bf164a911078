"""Spike-time adjoint of the V1 dynamics: the V1 output of a run as a
differentiable function of the LGN->V1 strengths and the recurrent weights,
with the forward and the reverse sweep in libvc_hip.so.

    run = V1Run(net, nsteps)                      # conductance-LIF network
    sInfo = run.spikes(sLGN, w)                   # [nsteps, nV1], autograd-aware
    loss = soft_count_loss(sInfo, target, ...)
    loss.backward()                               # -> sLGN.grad, w.grad

sInfo is the simulator's per-step output (src/coredynamics.cu:1603-1609):
1 + tsp/dt on a step where the neuron spikes at tsp, else its membrane
potential.  Gradients are exact for the discrete dynamics at the spike
pattern of the forward (which neurons spike on which steps): they flow
through the interpolated spike times, the refractory back-up that depends on
them, and the conductance increments they cause.

The forward keeps the state every `ckpt` steps; the backward recomputes each
segment from its checkpoint (the same kernel, bitwise the same spikes),
keeping a snapshot of every step, then sweeps it backwards.  The LGN stage
depends on the stimulus only, so its output for the whole run is computed
once and held.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import abi
from .abi import check, ptr
from .network import Network
from .sim import GpuSim

STATE_ROWS = abi.VC_SNAP_ROWS


class V1Run:
    def __init__(self, net: Network, nsteps: int, device: str = "cuda", ckpt: int = 100):
        sp = net.spec
        if sp.iModel != 0 or any(x != 0 for x in sp.deltaT) or any(x != 0 for x in sp.a):
            raise ValueError("the V1 adjoint covers the conductance LIF model "
                             "(iModel 0, deltaT = a = 0): use network.lif(spec)")
        self.sim = GpuSim(net, device=device, record=False, ff_events=False, dense_weights=False)
        self.net, self.T, self.ckpt = net, int(nsteps), int(ckpt)
        self.lib = self.sim.lib
        sim = self.sim
        d = sim.dev
        self.dev = d
        nV1, nLGN = sp.nV1, sp.nLGN
        # LGN output of every step, computed once
        npf = sim.tm["ntPerFrame"]
        self.lgn = torch.empty(self.T, nLGN, dtype=torch.float32, device=d)
        cur = torch.cuda.current_stream(d)
        for k in range(-(-self.T // npf)):
            sim._lgn_period(k)
            cur.wait_stream(sim.s_lgn)
            n = min(npf, self.T - k * npf)
            self.lgn[k * npf:k * npf + n].copy_(sim.lgn_sInfo2[k % 2][:n])
            sim.s_lgn.wait_stream(cur)
        self.init = {k: v.clone() for k, v in sim.state.items()}
        self.nseg = -(-self.T // self.ckpt)
        f32 = dict(dtype=torch.float32, device=d)
        self.ck_state = [None] * self.nseg
        self.ck_spk = [None] * self.nseg
        self.sInfo = torch.empty(self.T, nV1, **f32)
        # per-segment recompute buffers: snapshots and the spikes of every step
        nT = nV1 // 64
        L = self.ckpt
        self.snap = torch.empty(L, STATE_ROWS, nV1, **f32)
        self.seg_mask = torch.zeros(L + 1, nT, dtype=torch.int64, device=d)
        self.seg_og = torch.zeros(L + 1, abi.VC_MAX_NG, nV1, **f32)
        self.seg_oh = torch.zeros(L + 1, abi.VC_MAX_NG, nV1, **f32)
        self.seg_spk = (abi.Spikes * (L + 1))(*[
            abi.Spikes(ptr(self.seg_mask[i]), ptr(self.seg_og[i]), ptr(self.seg_oh[i]))
            for i in range(L + 1)])
        self.row = torch.empty(nV1, **f32)

    def _h(self):
        return C.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)

    def spikes(self, sLGN: torch.Tensor, w: torch.Tensor) -> torch.Tensor:
        return V1Spikes.apply(sLGN, w, self)

    # ---------------------------------------------------------------- forward
    def _set_weights(self, sLGN, w):
        sim = self.sim
        if sLGN.shape != sim.sLGN.shape or w.shape != sim.wrec.shape:
            raise ValueError(f"weights must be sLGN {tuple(sim.sLGN.shape)} and w "
                             f"{tuple(sim.wrec.shape)}")
        sim.sLGN.copy_(sLGN.detach())
        sim.wrec.copy_(w.detach())

    def _restore(self, state: dict, spk):
        sim = self.sim
        for k, v in state.items():
            sim.state[k].copy_(v)
        if spk is None:
            for t in (sim.masks[0], sim.og[0], sim.oh[0]):
                t.zero_()
        else:
            sim.masks[0].copy_(spk[0])
            sim.og[0].copy_(spk[1])
            sim.oh[0].copy_(spk[2])
        sim.cur = 0

    def forward(self, sLGN: torch.Tensor, w: torch.Tensor) -> torch.Tensor:
        sim, sp = self.sim, self.net.spec
        self._set_weights(sLGN, w)
        self._restore(self.init, None)
        for c in range(self.nseg):
            it0 = c * self.ckpt
            n = min(self.ckpt, self.T - it0)
            self.ck_state[c] = {k: v.clone() for k, v in sim.state.items()}
            q = sim.cur
            self.ck_spk[c] = (sim.masks[q].clone(), sim.og[q].clone(), sim.oh[q].clone())
            check(self.lib.vc_v1_run(self._h(), C.byref(sim.p), C.byref(sim.s), C.byref(sim.ff),
                                     C.byref(sim.rec), sim.spk2, sim.cur,
                                     ptr(self.lgn[it0]), sp.nLGN, None, 0, it0, n,
                                     ptr(self.sInfo[it0]), sp.nV1), "vc_v1_run")
            sim.cur ^= (n & 1)
        return self.sInfo.clone()

    # --------------------------------------------------------------- backward
    def backward(self, gOut: torch.Tensor, verify: bool = False):
        """(dL/dsLGN, dL/dw) for gOut = dL/dsInfo of the last forward.  With
        verify, every recomputed step is checked to reproduce the forward's
        output bitwise."""
        sim, sp = self.sim, self.net.spec
        nV1 = sp.nV1
        gOut = gOut.contiguous().float()
        f32 = dict(dtype=torch.float32, device=self.dev)
        lam = torch.zeros(STATE_ROWS, nV1, **f32)
        lamTsp = torch.zeros(2, nV1, **f32)
        dsLGN = torch.zeros_like(sim.sLGN)
        dw = torch.zeros_like(sim.wrec)
        for c in reversed(range(self.nseg)):
            it0 = c * self.ckpt
            n = min(self.ckpt, self.T - it0)
            self._restore(self.ck_state[c], None)
            self.seg_mask[0].copy_(self.ck_spk[c][0])
            self.seg_og[0].copy_(self.ck_spk[c][1])
            self.seg_oh[0].copy_(self.ck_spk[c][2])
            for k in range(n):
                it = it0 + k
                check(self.lib.vc_v1_step_snapshot(
                    self._h(), C.byref(sim.p), C.byref(sim.s), C.byref(sim.ff), C.byref(sim.rec),
                    C.byref(self.seg_spk[k]), C.byref(self.seg_spk[k + 1]), ptr(self.lgn[it]),
                    None, it, ptr(self.row), ptr(self.snap[k])), "vc_v1_step_snapshot")
                if verify and not torch.equal(self.row, self.sInfo[it]):
                    raise RuntimeError(f"recomputed step {it} differs from the forward")
            a = abi.V1Adjoint(ptr(self.snap), C.cast(self.seg_spk, C.c_void_p),
                              ptr(self.sInfo[it0]), sp.nLGN, ptr(self.lgn[it0]), ptr(gOut[it0]),
                              ptr(lam), ptr(lamTsp), ptr(dsLGN), ptr(dw))
            check(self.lib.vc_v1_adjoint_run(self._h(), C.byref(sim.p), C.byref(sim.s),
                                             C.byref(sim.ff), C.byref(sim.rec), C.byref(a),
                                             it0, n), "vc_v1_adjoint_run")
        return dsLGN, dw


class V1Spikes(torch.autograd.Function):
    """sInfo = V1 output(sLGN, w); backward through the HIP reverse sweep.
    The backward of a call must follow that call's forward on the same V1Run
    (the checkpoints belong to the last forward)."""

    @staticmethod
    def forward(ctx, sLGN, w, run: V1Run):
        ctx.run = run
        return run.forward(sLGN, w)

    @staticmethod
    def backward(ctx, gOut):
        dsLGN, dw = ctx.run.backward(gOut)
        return dsLGN, dw, None


def soft_bin_counts(sInfo: torch.Tensor, dt: float, bin_steps: int, tau_s: float) -> torch.Tensor:
    """Spike counts in bins of `bin_steps` steps [nBins, nV1], each spike
    counted with the weight sigma((t-lo)/tau_s) - sigma((t-hi)/tau_s) for its
    own bin and the two next to it (t its time, [lo, hi) the bin): the hard
    count as tau_s -> 0, and a smooth function of the spike times."""
    T, N = sInfo.shape
    nb = T // bin_steps
    it, n = torch.nonzero(sInfo >= 1.0, as_tuple=True)
    t = it.to(sInfo.dtype) * dt + (sInfo[it, n] - 1.0) * dt
    b0 = it // bin_steps
    counts = sInfo.new_zeros(nb * N)
    for db in (-1, 0, 1):
        b = b0 + db
        ok = (b >= 0) & (b < nb)
        lo = (b * bin_steps).to(sInfo.dtype) * dt
        k = torch.sigmoid((t - lo) / tau_s) - torch.sigmoid((t - lo - bin_steps * dt) / tau_s)
        counts = counts.index_add(0, (b * N + n)[ok], k[ok])
    return counts.view(nb, N)


def soft_count_loss(sInfo, target, dt, bin_steps, tau_s):
    """mean((soft counts - target)^2)."""
    return ((soft_bin_counts(sInfo, dt, bin_steps, tau_s) - target) ** 2).mean()
