"""Batched trials and a device-side network generator.

`build_network_fast(spec, device)` draws the same kind of synthetic network
as network.build_network (3x3-neighbourhood partners without replacement,
p ~ exp(-d^2/2 sigma^2), lognormal strengths; LGN->V1 by Gaussian distance),
but on the GPU with torch, and emits the recurrent synapses directly in the
masked-CSR layout of include/vc_abi.h.  It is used for the large workloads
(tens of millions of neurons-synapses per second of generation); its random
streams are torch's, so its networks differ from build_network's for the
same seed but are reproducible on the same device.

`batch_networks(nets)` concatenates independent trials into one network
(blocks, LGN cells, synapses, stimuli), with the per-trial tables that let
the kernels draw each trial's random numbers with its own key and local ids
(vc_v1_params.blockTrial, vc_lgn_params.cellTrial): a trial gives the same
result batched as run alone.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

import numpy as np
import torch

from . import network as N


def _gumbel_topk(w: torch.Tensor, k: int, gen: torch.Generator) -> torch.Tensor:
    """Row-wise weighted sampling of k columns without replacement:
    largest log(u)/w (Efraimidis & Spirakis 2006)."""
    u = torch.rand(w.shape, generator=gen, device=w.device, dtype=torch.float32).clamp_(min=1e-30)
    keys = torch.where(w > 0, torch.log(u) / w, torch.full_like(w, -float("inf")))
    return torch.topk(keys, k, dim=1, sorted=False).indices


def build_network_fast(spec: N.ModelSpec, device="cuda", with_frames: bool = True) -> N.Network:
    dev = torch.device(device)
    gen = torch.Generator(device=dev)
    gen.manual_seed(spec.seed)
    rng = np.random.default_rng(spec.seed)
    bs, nb = spec.blockSize, spec.nblock
    cols = N._block_grid(nb)
    blk_mm = spec.sheet_mm / cols
    # V1 positions
    b_of = torch.arange(spec.nV1, device=dev) // bs
    bx = (b_of % cols).float()
    by = (b_of // cols).float()
    pos = torch.stack([(bx + torch.rand(spec.nV1, generator=gen, device=dev)) * blk_mm,
                       (by + torch.rand(spec.nV1, generator=gen, device=dev)) * blk_mm], 1)
    local = torch.arange(spec.nV1, device=dev) % bs
    isE = local < spec.nE
    maxNear, nNear, nbr = N._neighbours(nb)
    nSeg = bs // 64
    sig2 = 2.0 * (spec.sigma_um / 1000.0) ** 2
    masks, counts, ws = [], [], []
    for b in range(nb):
        nbl = [int(x) for x in nbr[b, :nNear[b]]]
        cand = torch.cat([torch.arange(j * bs, (j + 1) * bs, device=dev) for j in nbl])
        slot_of_cand = torch.repeat_interleave(torch.arange(len(nbl), device=dev), bs)
        posts = torch.arange(b * bs, (b + 1) * bs, device=dev)
        d2 = torch.cdist(pos[posts], pos[cand]).pow_(2)
        wgt = torch.exp(-d2 / sig2)
        wgt[:, 0:bs].fill_diagonal_(0.0)                    # own block is slot 0: no self-connection
        pre_list, post_list, w_list = [], [], []
        for sel_mask, K, scale in ((isE[cand], spec.KE, spec.sE), (~isE[cand], spec.KI, spec.sI)):
            if K == 0:
                continue
            cidx = torch.nonzero(sel_mask).squeeze(1)
            wsub = wgt[:, cidx]
            Kc = min(K, int((wsub > 0).sum(1).min().item()))
            pick = cidx[_gumbel_topk(wsub, Kc, gen)]        # [bs, Kc] indices into cand
            strength = scale * torch.exp(spec.w_mu + spec.w_sigma *
                                         torch.randn(pick.shape, generator=gen, device=dev))
            pre_list.append(pick.reshape(-1))
            post_list.append(torch.arange(bs, device=dev).repeat_interleave(Kc))
            w_list.append(strength.reshape(-1).float())
        if pre_list:
            cpos = torch.cat(pre_list)
            post = torch.cat(post_list)
            wv = torch.cat(w_list)
            row = (b * maxNear + slot_of_cand[cpos]) * bs + (cand[cpos] % bs)
            key = row * bs + post
            order = torch.argsort(key)
            key, post, wv = key[order], post[order], wv[order]
            seg_key = (key // bs) * nSeg + (post // 64) - b * maxNear * bs * nSeg
            m = torch.zeros(maxNear * bs * nSeg, dtype=torch.int64, device=dev)
            bits = torch.bitwise_left_shift(torch.ones_like(post, dtype=torch.int64), post % 64)
            m.index_add_(0, seg_key, bits)                   # distinct bits: sum == or
            cnt = torch.bincount(seg_key, minlength=maxNear * bs * nSeg)
            masks.append(m)
            counts.append(cnt)
            ws.append(wv)
        else:
            masks.append(torch.zeros(maxNear * bs * nSeg, dtype=torch.int64, device=dev))
            counts.append(torch.zeros(maxNear * bs * nSeg, dtype=torch.int64, device=dev))
    rowMask = torch.cat(masks)
    cnt = torch.cat(counts)
    rowOff = torch.zeros_like(cnt)
    rowOff[1:] = torch.cumsum(cnt[:-1], 0)
    if int(rowOff[-1] + cnt[-1]) >= 2 ** 32:
        raise ValueError("too many recurrent synapses for 32-bit offsets")
    w = torch.cat(ws) if ws else torch.zeros(1, device=dev)

    # LGN cells (jittered grid, checkerboard ON/OFF) -- same distributions as build_network
    ns = spec.nLGN_side
    spacing = spec.field_deg / ns
    gx, gy = np.meshgrid(np.arange(ns), np.arange(ns))
    gx = gx.ravel(); gy = gy.ravel()
    lx = (gx + 0.5) * spacing - spec.field_deg / 2 + (rng.random(ns * ns) - 0.5) * 0.6 * spacing
    ly = (gy + 0.5) * spacing - spec.field_deg / 2 + (rng.random(ns * ns) - 0.5) * 0.6 * spacing
    on = ((gx + gy) % 2 == 0)
    nL = ns * ns
    lgn = N.lgn_parameters(spec, lx, ly, on)
    # LGN -> V1: Gaussian distance in the visual field, chunked over neurons
    posn = pos.cpu().numpy().astype(np.float64)
    vx = torch.tensor((posn[:, 0] / spec.sheet_mm - 0.5) * spec.field_deg * 0.9, device=dev)
    vy = torch.tensor((posn[:, 1] / spec.sheet_mm - 0.5) * spec.field_deg * 0.9, device=dev)
    L = torch.tensor(np.stack([lx, ly], 1), device=dev)
    sig = spec.ff_sigma_cells * spacing
    ids = torch.zeros(spec.nV1, max(spec.m, 1), dtype=torch.int64, device=dev)
    if spec.m > 0:
        for s0 in range(0, spec.nV1, 8192):
            s1 = min(spec.nV1, s0 + 8192)
            d2 = (vx[s0:s1, None] - L[None, :, 0]) ** 2 + (vy[s0:s1, None] - L[None, :, 1]) ** 2
            wq = torch.exp(-d2 / (2 * sig * sig)).float()
            ids[s0:s1] = _gumbel_topk(wq, spec.m, gen)
        ids, _ = torch.sort(ids, dim=1)
    sR = torch.where(isE, spec.sRatioLGN[0], spec.sRatioLGN[1]).double()
    s_ff = (spec.sLGN * sR)[:, None] * (0.75 + 0.5 * torch.rand(ids.shape, generator=gen, device=dev,
                                                                 dtype=torch.float64))
    net = N.Network(
        spec=spec, pos=posn, tonicDep=np.where(isE.cpu().numpy(), spec.tonicDep[0],
                                              spec.tonicDep[1]).astype(np.float32),
        maxNear=maxNear, nNear=nNear, nbr=nbr, rec_pairs=None,
        nLGN_of=np.full(spec.nV1, spec.m, dtype=np.uint32),
        sLGN=np.ascontiguousarray(s_ff.T.float().cpu().numpy()),
        idLGN=np.ascontiguousarray(ids.T.cpu().numpy().astype(np.uint32)),
        lgn=lgn,
    )
    net.csr = (rowMask, rowOff.to(torch.int64), w)
    if with_frames:
        net.frames = N.grating_frames(spec)
    return net


@dataclass
class BatchInfo:
    ntrial: int
    blockTrial: np.ndarray            # [nblock] u32
    trialV1Off: np.ndarray            # [ntrial] u32
    cellTrial: np.ndarray             # [nLGN] u32
    trialLgnOff: np.ndarray           # [ntrial] u32
    trialSeed: np.ndarray             # [2*ntrial] u32
    trialBlocks: list = field(default_factory=list)
    trialNV1: list = field(default_factory=list)


def batch_networks(nets: list, device=None) -> N.Network:
    """One network holding every trial of `nets` (independent, concatenated).
    All trials share blockSize, nE, receptors, time step, stimulus size and
    frame count; blocks of trial t only connect within trial t."""
    sp0 = nets[0].spec
    for n in nets:
        s = n.spec
        if (s.blockSize, s.nE, s.dt, s.width, s.height, n.frames.shape, s.noDelay,
                s.speedOfThought, n.far_delayed()) != (
                    sp0.blockSize, sp0.nE, sp0.dt, sp0.width, sp0.height, nets[0].frames.shape,
                    sp0.noDelay, sp0.speedOfThought, nets[0].far_delayed()):
            raise ValueError("batched trials must share block size, dt, stimulus shape and "
                             "axonal-delay settings (noDelay, noFarDelay, speedOfThought)")
        noise = ("noisyDep", "noiseOnTonic", "tau_noise", "synFailFF", "synPerConFF", "synFail",
                 "synPerCon")
        if any(getattr(s, k) != getattr(sp0, k) for k in noise):
            raise ValueError("batched trials must share the noise and synaptic-failure "
                             "parameters (" + ", ".join(noise) + ")")
        learn = ("learning", "tauLTP", "tauLTD", "tauTrip", "r_LTD", "A_LGN", "tauAvg",
                 "targetFR", "gmaxLGN", "gminLGN", "applyHomeo", "symmetricHomeo",
                 "decayTimeHomeoFF", "FF_InfRatio", "rebound")
        if any(getattr(s, k) != getattr(sp0, k) for k in learn) or (
                s.learning and s.sRatioLGN[0] != sp0.sRatioLGN[0]):
            raise ValueError("batched trials must share the plasticity parameters ("
                             + ", ".join(learn) + ", sRatioLGN[0])")
    if device is None:
        c0 = nets[0].csr
        device = c0[0].device if c0 is not None and hasattr(c0[0], "device") else \
            ("cuda" if torch.cuda.is_available() else "cpu")
    maxNear = max(n.maxNear for n in nets)
    bs = sp0.blockSize
    nSeg = bs // 64
    blocks = [n.spec.nblock for n in nets]
    boff = np.concatenate([[0], np.cumsum(blocks)])
    nL = [n.spec.nLGN for n in nets]
    loff = np.concatenate([[0], np.cumsum(nL)])
    nbr = np.zeros((int(boff[-1]), maxNear), np.uint32)
    nNear = np.concatenate([n.nNear for n in nets]).astype(np.uint32)
    for t, n in enumerate(nets):
        nbr[boff[t]:boff[t + 1], :n.maxNear] = n.nbr + boff[t]
    # masked CSR: rows (b*maxNear + ib)*bs + pre; trial t's rows follow
    # trial t-1's, its offsets are shifted by the synapses before it
    masks, offs, ws = [], [], []
    base = 0
    for t, n in enumerate(nets):
        rm, ro, w = n.masked_csr_torch(device)
        nnz = n.n_rec_synapses_csr()
        rm = rm.view(n.spec.nblock, n.maxNear, bs * nSeg)
        ro = ro.view(n.spec.nblock, n.maxNear, bs * nSeg) + base
        if n.maxNear < maxNear:                      # empty neighbour slots
            pad = maxNear - n.maxNear
            rm = torch.cat([rm, torch.zeros(n.spec.nblock, pad, bs * nSeg, dtype=rm.dtype,
                                            device=rm.device)], 1)
            ro = torch.cat([ro, torch.full((n.spec.nblock, pad, bs * nSeg), base, dtype=ro.dtype,
                                           device=ro.device)], 1)
        masks.append(rm.reshape(-1))
        offs.append(ro.reshape(-1))
        ws.append(w[:nnz])
        base += nnz
    rowMask = torch.cat(masks)
    rowOff = torch.cat(offs)
    w = torch.cat(ws) if base > 0 else torch.zeros(1, device=rowMask.device)
    sLGN_max = max(n.sLGN.shape[0] for n in nets)
    nV1 = sum(n.spec.nV1 for n in nets)
    sLGN = np.zeros((sLGN_max, nV1), np.float32)
    idLGN = np.zeros((sLGN_max, nV1), np.uint32)
    v0 = 0
    for t, n in enumerate(nets):
        k = n.sLGN.shape[0]
        sLGN[:k, v0:v0 + n.spec.nV1] = n.sLGN
        idLGN[:k, v0:v0 + n.spec.nV1] = n.idLGN + loff[t]
        v0 += n.spec.nV1
    lgn = {}
    for key, val in nets[0].lgn.items():
        if isinstance(val, np.ndarray):
            lgn[key] = np.concatenate([n.lgn[key] for n in nets], axis=-1)
    spec = replace(sp0, nblock=int(boff[-1]), nLGN_override=int(loff[-1]),
                   m=max(n.spec.m for n in nets))
    net = N.Network(spec=spec, pos=np.concatenate([n.pos for n in nets]),
                    tonicDep=np.concatenate([n.tonicDep for n in nets]), maxNear=maxNear,
                    nNear=nNear, nbr=nbr, rec_pairs=None,
                    nLGN_of=np.concatenate([n.nLGN_of for n in nets]),
                    sLGN=sLGN, idLGN=idLGN, lgn=lgn,
                    frames=np.stack([n.frames for n in nets]), init_lms=nets[0].init_lms)
    net.csr = (rowMask, rowOff, w)
    if any(getattr(n, "far", None) is not None and n.n_far() > 0 for n in nets):
        # far lists: trial t's post neurons and presynaptic ids shifted by its first neuron
        offs_, pres, wf, df = [np.zeros(1, np.int64)], [], [], []
        v0 = base_f = 0
        for n in nets:
            nv = n.spec.nV1
            if getattr(n, "far", None) is None:
                offs_.append(np.full(nv, base_f, np.int64))
            else:
                o, pr, ww, dd = n.far
                offs_.append(o[1:].astype(np.int64) + base_f)
                pres.append(pr.astype(np.int64) + v0)
                wf.append(ww)
                df.append(dd)
                base_f += int(o[-1])
            v0 += nv
        cat = lambda xs, dt: np.concatenate(xs).astype(dt) if xs else np.zeros(0, dt)
        net.far = (np.concatenate(offs_).astype(np.uint32), cat(pres, np.uint32),
                   cat(wf, np.float32), cat(df, np.float32))
    seeds = np.array([[n.spec.seed & 0xffffffff, (n.spec.seed >> 32) & 0xffffffff] for n in nets],
                     np.uint32).ravel()
    net.batch = BatchInfo(
        ntrial=len(nets),
        blockTrial=np.repeat(np.arange(len(nets)), blocks).astype(np.uint32),
        trialV1Off=(boff[:-1] * bs).astype(np.uint32),
        cellTrial=np.repeat(np.arange(len(nets)), nL).astype(np.uint32),
        trialLgnOff=loff[:-1].astype(np.uint32), trialSeed=seeds,
        trialBlocks=blocks, trialNV1=[n.spec.nV1 for n in nets])
    net.trial_specs = [n.spec for n in nets]
    return net


def c3_specs(ntrial: int = 32, seed: int = 2024, blocks_lo: int = 8, blocks_hi: int = 96,
             nt: int = 4000, **kw):
    """Batched-trials workload: patch sizes uniform in [8, 96] blocks, 8
    orientations at 22.5 deg, distinct seeds, the small plain-forward
    distributions (1024-neuron blocks, 400 E + 100 I partners, 20 LGN inputs,
    256 LGN cells and 1 mm^2 per block, 128x128 px, dt 0.125 ms), 500 ms."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(blocks_lo, blocks_hi + 1, size=ntrial)
    specs = []
    for t, nb in enumerate(sizes):
        side = N._block_grid(int(nb))        # the sheet is side x ceil(nb/side) mm
        specs.append(N.spec_c1(nblock=int(nb), sheet_mm=float(side), nt=nt,
                               nLGN_side=16 * side, field_deg=3.0,
                               ori_deg=22.5 * (t % 8), seed=seed * 1000 + t, **kw))
    return specs
