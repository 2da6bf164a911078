"""Model specification and synthetic network construction for the patchV1
simulation step.

A `ModelSpec` carries the reference's cfg parameters (src/patch.cu:180-374,
same key names where they exist) plus the sizes of a synthetic workload.
`build_network(spec)` generates, with numpy and a fixed seed:

  * V1: `nblock` blocks of `blockSize` neurons (E first, src/patch.cu:695-706),
    block b covering a 1 mm x 1 mm square of a sqrt(nblock)-wide grid;
  * recurrent synapses: each neuron draws `KE` E and `KI` I partners from its
    3x3 block neighbourhood without replacement with p ~ exp(-d^2/2 sigma^2)
    (Efraimidis-Spirakis keys), lognormal strengths;
  * LGN: ON/OFF cells on a jittered grid, DoG centre/surround parameters and a
    biphasic difference-of-gamma temporal kernel (the fLGN parameter set,
    src/patch.cu:2472-2530);
  * LGN->V1: `m` inputs per neuron by Gaussian distance in the visual field;
  * a drifting sinusoidal grating (the stimulus.bin layout, grating.py:267-372).

The same Network feeds the GPU path (masked CSR, see include/vc_abi.h) and the
CPU oracle (the reference's dense conMat).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

import numpy as np

DEG2RAD = math.pi / 180.0


@dataclass
class ModelSpec:
    # sizes
    nblock: int = 4
    blockSize: int = 1024
    nE: int = 768
    KE: int = 400
    KI: int = 100
    m: int = 20
    nLGN_side: int = 32           # LGN grid is nLGN_side^2 cells
    sheet_mm: float = 2.0         # V1 sheet side
    field_deg: float = 3.0        # LGN centre field side
    width: int = 128              # stimulus pixels
    height: int = 128
    stim_deg: float = 4.0         # stimulus extent (range = stim_deg/2)
    # time
    dt: float = 0.125
    nt: int = 8000
    frameRate: int = 125
    # V1 neurons, per type [E, I]  (I is LIF: a = b = 0, deltaT = 0)
    iModel: int = 1
    vR: tuple = (-58.0, -58.0)
    vThres: tuple = (-20.0, -50.0)
    gL: tuple = (2.0, 2.8)
    C: tuple = (20.0, 15.0)
    tRef: tuple = (2.0, 1.0)
    vT: tuple = (-50.0, -50.0)
    deltaT: tuple = (1.5, 0.0)
    tau_w: tuple = (200.0, 10.0)
    a: tuple = (0.0, 0.0)
    b: tuple = (1.0, 0.0)
    tonicDep: tuple = (0.0, 0.0)
    # receptors: FF/E = (AMPA, NMDA), I = (GABA_A,)
    riseTimeFF: tuple = (1.0, 5.0)
    decayTimeFF: tuple = (3.0, 80.0)
    riseTimeE: tuple = (1.0, 5.0)
    decayTimeE: tuple = (3.0, 80.0)
    riseTimeI: tuple = (1.0,)
    decayTimeI: tuple = (5.0,)
    pFF: tuple = ((0.85, 0.15), (0.85, 0.15))
    pE: tuple = ((0.8, 0.2), (0.8, 0.2))
    pI: tuple = ((1.0,), (1.0,))
    # synaptic scales
    sRatioLGN: tuple = (1.0, 1.0)
    sLGN: float = 1.2             # base LGN synapse strength
    sE: float = 0.12              # E recurrent scale (times lognormal)
    sI: float = 2.0               # I recurrent scale
    w_mu: float = -1.0
    w_sigma: float = 0.8
    sigma_um: float = 150.0
    # axonal delays of the near-neighbour synapses (src/patch.cu:184, 236):
    # noDelay 0 delays every spike by distance / speedOfThought (mm/ms)
    noDelay: int = 1
    speedOfThought: float = 1.0
    # far (long-range) connections, the reference's V1_vec lists: delayed by
    # distance / speedOfThought when noFarDelay is 0; with noFarDelay 1 they
    # add nothing (src/coredynamics.cu:219-220).  -1: the same as noDelay.
    noFarDelay: int = -1
    # synthetic far inputs per post neuron, from blocks that are not its near
    # neighbours (strength sFar times lognormal)
    nFar: int = 0
    sFar: float = 0.05
    # gap junctions between inhibitory neurons (the reference's InhGap): on
    # when InhGap is 1; synthetic networks get nGap partners per inhibitory
    # neuron among the inhibitory neurons of its block's near neighbours
    # (strength sGap times lognormal)
    InhGap: int = 0
    nGap: int = 0
    sGap: float = 0.2
    # voltage noise (src/coredynamics.cu:1424-1445): sigma = sqrt(2 noisyDep
    # (vT - vR) r dt), r = the neuron's tonicDep when noiseOnTonic (else 1);
    # white when tau_noise is 0, Ornstein-Uhlenbeck with time scale
    # tau_noise (ms) otherwise
    noisyDep: tuple = (0.0, 0.0)
    noiseOnTonic: int = 1
    tau_noise: float = 0.0
    # synaptic failure (src/coredynamics.cu:1476-1480, 2766-2769): rate per
    # post type for LGN synapses, per [pre type][post type] for V1 ones,
    # with the synapses per connection that set the spread of the factor
    synFailFF: tuple = (0.0, 0.0)
    synPerConFF: tuple = (1.0, 1.0)
    synFail: tuple = (0.0, 0.0, 0.0, 0.0)
    synPerCon: tuple = (1.0, 1.0, 1.0, 1.0)
    # background Poisson drive (extension)
    bgRate: float = 2000.0        # Hz
    bgStrength: tuple = (0.08, 0.12)
    pBG: tuple = ((1.0, 0.0), (1.0, 0.0))
    # LGN
    nSpatialSample1D: int = 16
    nsig: float = 3.0
    tau: float = 40.0             # temporal kernel span, ms
    nKernelSample: int = 40
    sigC_deg: float = 0.05
    sigS_deg: float = 0.15
    kS_over_kC: float = 0.7
    tauR: float = 2.0
    tauD: float = 3.0
    nR: float = 6.0
    nD: float = 8.0
    ratio: float = 0.8
    spont: float = 0.1
    c50: float = 0.15
    sharpness: float = 10.0
    frRatioLGN: float = 60.0
    convolRatio: float = 1.0
    ff_sigma_cells: float = 1.5   # LGN->V1 pooling sigma in LGN grid spacings
    # grating
    sf_cpd: float = 4.0
    tf_hz: float = 8.0
    contrast: float = 0.5
    ori_deg: float = 0.0
    phase0: float = 0.0
    # seeds
    seed: int = 1234
    trial: int = 0
    # the reference's initial-state draws (cfg keys w0, gFF0, gE0, gI0,
    # use_v0; src/patch.cu:4405-4570): None keeps the synthetic initial state
    # (v uniform below vT, w and conductances 0); otherwise w0 is [type][mean,
    # std], gFF0/gE0/gI0 [type][ig][g mean, g std/mean, h mean, h std/mean]
    init_w0: tuple | None = None
    init_gFF0: tuple | None = None
    init_gE0: tuple | None = None
    init_gI0: tuple | None = None
    use_v0: int = 0
    # initial spikes of rand_spInit (src/coredynamics.cu:5-146): per type the
    # rate in Hz for neurons with more than SCsplit LGN inputs, then for the
    # others (cfg spE0, spI0, SCsplit)
    spE0: tuple = (0.0, 0.0)
    spI0: tuple = (0.0, 0.0)
    SCsplit: int = 0
    # plasticity of the LGN->V1 strengths: learning 3 is the reference's FF_E
    # triplet rule (src/patch.cu:1092-1330, cfg keys of the same names; the
    # E entries of tauAvg/targetFR and the first of the per-rule lists); 0 =
    # off.  Homeostasis of each neuron's total FF strength: applyHomeo 1
    # (multiplicative) or 2 (additive), relaxing over decayTimeHomeoFF towards
    # FF_InfRatio x the initial total.  learnData_FF 1 writes the strengths
    # every sampleInterval_LGN_V1 steps (the reference's sLGN file).
    learning: int = 0
    tauLTP: float = 16.8
    tauLTD: float = 33.7
    tauTrip: float = 114.0
    r_LTD: float = 1.0
    A_LGN: float = 0.0
    tauAvg: tuple = (1000.0, 1000.0)
    targetFR: tuple = (0.0, 0.0)
    gmaxLGN: float = 1.0
    gminLGN: float = 0.0
    applyHomeo: int = 0
    symmetricHomeo: int = 1
    decayTimeHomeoFF: float = 250.0
    FF_InfRatio: float = 1.0
    rebound: int = 0
    learnData_FF: int = 0
    sampleInterval_LGN_V1: int = 100
    # set when the network comes from the reference's input files
    nLGN_override: int = 0
    stim_buffer_deg: float = 0.0

    @property
    def nV1(self) -> int:
        return self.nblock * self.blockSize

    @property
    def nLGN(self) -> int:
        return self.nLGN_override if self.nLGN_override else self.nLGN_side * self.nLGN_side

    @property
    def nI(self) -> int:
        return self.blockSize - self.nE


def spec_c1(**kw) -> ModelSpec:
    """Plain forward, small: 4 blocks x 1024 conductance-LIF neurons (iModel
    0, deltaT = a = b = 0 for both types), 1024 LGN, 128x128 px, 1 s."""
    return replace(lif(ModelSpec()), **kw)


def spec_c2(**kw) -> ModelSpec:
    """Plain forward, full patch: 64 blocks, 16384 LGN, m = 30, 800 E + 200 I
    partners, 256x256 px, 2 s."""
    base = ModelSpec(nblock=64, KE=800, KI=200, m=30, nLGN_side=128, sheet_mm=8.0,
                     width=256, height=256, nt=16000, sE=0.06, sI=1.0)
    return replace(base, **kw)


def spec_c5(**kw) -> ModelSpec:
    """Differentiation of the V1 dynamics: 16 blocks = 16,384 conductance-LIF
    neurons (iModel 0, deltaT = a = b = 0 for both types), 20 LGN + 400 E +
    100 I synapses per neuron, 4096 LGN cells, 250 ms (2000 steps)."""
    base = ModelSpec(nblock=16, sheet_mm=4.0, nLGN_side=64, nt=2000, iModel=0,
                     vThres=(-50.0, -50.0), deltaT=(0.0, 0.0), a=(0.0, 0.0), b=(0.0, 0.0))
    return replace(base, **kw)


def lif(spec: ModelSpec) -> ModelSpec:
    """The conductance-LIF variant of a spec (the model the V1 adjoint covers)."""
    return replace(spec, iModel=0, vThres=(-50.0, -50.0), deltaT=(0.0, 0.0), a=(0.0, 0.0),
                   b=(0.0, 0.0))


def spec_tiny(**kw) -> ModelSpec:
    """A small case that the oracle finishes in well under a second per 100
    steps: 2 blocks of 256 (192 E), 64 LGN cells, 32x32 px."""
    base = ModelSpec(nblock=2, blockSize=256, nE=192, KE=100, KI=25, m=8, nLGN_side=8,
                     sheet_mm=1.0, field_deg=2.0, width=32, height=32, nt=400,
                     sE=0.8, sI=3.0, bgStrength=(0.15, 0.15))
    return replace(base, **kw)


# ---------------------------------------------------------------------------
@dataclass
class Network:
    spec: ModelSpec
    # V1
    pos: np.ndarray                 # [nV1, 2] mm
    tonicDep: np.ndarray            # [nV1]
    maxNear: int
    nNear: np.ndarray               # [nblock] u32
    nbr: np.ndarray                 # [nblock, maxNear] u32
    rec_pairs: list                 # per (b, ib): (pre_local u32, post_local u32, w f32), sorted by (pre, post)
    # LGN -> V1 (transposed ELL)
    nLGN_of: np.ndarray             # [nV1] u32
    sLGN: np.ndarray                # [maxLGN, nV1] f32
    idLGN: np.ndarray               # [maxLGN, nV1] u32
    # LGN parameters, [2, nLGN] (centre, surround)
    lgn: dict = field(default_factory=dict)
    # stimulus
    frames: np.ndarray | None = None  # [nFrame, 3, H, W] f32 ([ntrial, ...] when batched)
    init_lms: tuple = (0.5, 0.5, 0.5)
    # recurrent synapses already in masked-CSR form (rowMask, rowOff, w), as
    # produced by the device generator; used instead of rec_pairs when set
    csr: tuple | None = None
    batch: object | None = None         # batch.BatchInfo for batched trials
    # distances of the rec_pairs (per (b, ib), aligned), when read from a
    # V1_delayMat file; otherwise they are computed from pos
    rec_dist: list | None = None
    trial_specs: list | None = None
    # far connections as CSR over post neurons, in the V1_vec file's order:
    # (off u32 [nV1+1], pre u32, w f32 already * sRatioV1, dist f32 mm)
    far: tuple | None = None
    # gap junctions between inhibitory neurons of near-neighbour blocks, per
    # (b, ib) like rec_pairs: (pre u32, post u32 -- indices among the nI
    # inhibitory neurons of the blocks --, g f32 already * gapRatio, > 0),
    # sorted by (pre, post); gap_dist: their distances (mm) when read from a
    # V1_delayMat file, else computed from pos
    gap: list | None = None
    gap_dist: list | None = None
    # far gap junctions (V1_gapVec) as CSR over the mI inhibitory neurons:
    # (off u32 [mI+1], pre u32 global id, g f32 already * gapRatio, dist f32)
    gap_far: tuple | None = None

    def n_far(self) -> int:
        return 0 if self.far is None else int(self.far[0][-1])

    def far_delayed(self) -> bool:
        nfd = self.spec.noFarDelay
        return (self.spec.noDelay if nfd < 0 else nfd) == 0

    def far_ell(self):
        """(maxFar, nFar u32 [nV1], pre u32, w f32, dist f32 [maxFar, nV1]):
        the far lists as a transposed ELL, connection j of post n at [j, n]."""
        off, pre, w, dist = self.far
        nV1 = self.spec.nV1
        cnt = np.diff(off.astype(np.int64))
        mf = max(int(cnt.max()) if cnt.size else 0, 1)
        P = np.zeros((mf, nV1), np.uint32)
        W = np.zeros((mf, nV1), np.float32)
        Dm = np.zeros((mf, nV1), np.float32)
        post = np.repeat(np.arange(nV1), cnt)
        j = np.arange(pre.size) - np.repeat(off[:-1].astype(np.int64), cnt)
        P[j, post], W[j, post], Dm[j, post] = pre, w, dist
        return mf, cnt.astype(np.uint32), P, W, Dm

    def far_out_index(self):
        """(outOff u32 [nV1+1], outPost u32, outJ u32): the far connections
        grouped by presynaptic neuron (ascending post neuron, then j), each as
        (post neuron, index j in its list)."""
        off, pre = self.far[0].astype(np.int64), self.far[1].astype(np.int64)
        nV1 = self.spec.nV1
        if pre.size and int(pre.max()) >= nV1:
            raise ValueError("a far connection's presynaptic id is out of range")
        cnt = np.diff(off)
        post = np.repeat(np.arange(nV1), cnt)
        j = np.arange(pre.size) - np.repeat(off[:-1], cnt)
        order = np.argsort(pre, kind="stable")
        outOff = np.zeros(nV1 + 1, np.uint32)
        outOff[1:] = np.cumsum(np.bincount(pre, minlength=nV1))
        return outOff, post[order].astype(np.uint32), j[order].astype(np.uint32)

    def far_pre_mask(self) -> np.ndarray:
        """int64 [nV1/64]: bit n%64 of word n/64 set when neuron n is the
        presynaptic neuron of a far connection."""
        nV1 = self.spec.nV1
        b = np.zeros(nV1, bool)
        b[self.far[1][self.far[1] < nV1].astype(np.int64)] = True
        return np.packbits(b.reshape(-1, 64), axis=-1, bitorder="little").reshape(-1).view(np.int64)

    # ---- gap junctions --------------------------------------------------------
    def has_gap(self) -> bool:
        """Any gap-junction strength (near or far) present."""
        near = self.gap is not None and any(p[0].size for p in self.gap)
        far = self.gap_far is not None and int(self.gap_far[0][-1]) > 0
        return near or far

    def gap_on(self) -> bool:
        """Gap currents are computed (InhGap and some near junction)."""
        return bool(self.spec.InhGap) and self.gap is not None and \
            any(p[0].size for p in self.gap)

    def _gap_post_major(self):
        """Every near junction ordered by post inhibitory neuron (global index
        b*nI + post), then neighbour slot, then partner: (post_gi, slot ib,
        pre local index among the inhibitory neurons, g, dist or None)."""
        sp = self.spec
        nI = sp.nI
        P, S, L, G, D = [], [], [], [], []
        for b in range(sp.nblock):
            for ib in range(int(self.nNear[b])):
                pre, post, g = self.gap[b * self.maxNear + ib]
                if pre.size == 0:
                    continue
                P.append(b * nI + post.astype(np.int64))
                S.append(np.full(pre.size, ib, np.int64))
                L.append(pre.astype(np.int64))
                G.append(np.asarray(g, np.float32))
                D.append(self._gap_pair_dist(b, ib, pre, post))
        if not P:
            z = np.zeros(0, np.int64)
            return z, z, z, np.zeros(0, np.float32), np.zeros(0, np.float32)
        post, ib, pre, g, d = (np.concatenate(x) for x in (P, S, L, G, D))
        order = np.lexsort((pre, ib, post))
        return post[order], ib[order], pre[order], g[order], d[order]

    def _gap_pair_dist(self, b, ib, pre, post):
        if self.gap_dist is not None:
            return np.asarray(self.gap_dist[b * self.maxNear + ib], np.float32)
        nE = self.spec.nE
        bs = self.spec.blockSize
        pos = np.asarray(self.pos, dtype=np.float64)
        gpre = int(self.nbr[b, ib]) * bs + nE + pre.astype(np.int64)
        gpost = b * bs + nE + post.astype(np.int64)
        x = (pos[gpre, 0] - pos[gpost, 0]).astype(np.float32)
        y = (pos[gpre, 1] - pos[gpost, 1]).astype(np.float32)
        return np.sqrt(x * x + y * y)

    def gap_ell(self):
        """(maxGap, nGap u32 [mI], pre u32 [maxGap, mI] global partner id,
        g f32 [maxGap, mI]): the near junctions as vc_rec_conn's ELL, in the
        reference's summation order (neighbour slot, then partner)."""
        sp = self.spec
        mI, nI, bs = sp.nblock * sp.nI, sp.nI, sp.blockSize
        post, ib, pre, g, _ = self._gap_post_major()
        cnt = np.bincount(post, minlength=mI)
        mg = max(int(cnt.max()) if cnt.size else 0, 1)
        start = np.zeros(mI, np.int64)
        np.cumsum(cnt[:-1], out=start[1:])
        j = np.arange(post.size) - start[post]
        b = post // nI
        P = np.zeros((mg, mI), np.uint32)
        G = np.zeros((mg, mI), np.float32)
        P[j, post] = self.nbr[b, ib].astype(np.int64) * bs + sp.nE + pre
        G[j, post] = g
        return mg, cnt.astype(np.uint32), P, G

    def gap_lags(self, speed: float, dt: float) -> np.ndarray:
        """u32 [maxGap, mI], aligned with gap_ell: ceil((distance / speed) /
        dt) in float32, the steps a junction's partner voltage travels with
        axonal delays (recal_G_mat, src/coredynamics.cu:2249-2256)."""
        sp = self.spec
        mI = sp.nblock * sp.nI
        post, ib, pre, g, d = self._gap_post_major()
        mg, cnt, _, _ = self.gap_ell()
        start = np.zeros(mI, np.int64)
        np.cumsum(cnt[:-1].astype(np.int64), out=start[1:])
        j = np.arange(post.size) - start[post]
        f = np.float32
        t = d.astype(f) / f(speed)
        lag = np.zeros((mg, mI), np.uint32)
        lag[j, post] = np.ceil(t / f(dt)).astype(np.uint32)
        return lag

    def gap_strength_sums(self) -> np.ndarray:
        """gapS f32 [mI]: per inhibitory neuron, 0 + its far strengths (file
        order) + its near strengths (neighbour slot, then partner), added one
        at a time in float32 (src/patch.cu:3367-3397, 3695-3721)."""
        sp = self.spec
        mI = sp.nblock * sp.nI
        acc = np.zeros(mI, np.float32)
        if self.gap_far is not None:
            off, _, g, _ = self.gap_far
            cnt = np.diff(off.astype(np.int64))
            for j in range(int(cnt.max()) if cnt.size else 0):
                has = cnt > j
                idx = off[:-1].astype(np.int64) + j
                acc[has] = acc[has] + g[idx[has]]
        if self.gap is not None:
            mg, cnt, _, G = self.gap_ell()
            for j in range(mg):
                has = cnt > j
                acc[has] = acc[has] + G[j, has]
        return acc

    def dense_gapmat(self) -> np.ndarray:
        """The reference's V1_gapMat layout [nblock, maxNear, nI pre, nI post]."""
        sp = self.spec
        M = np.zeros((sp.nblock, self.maxNear, sp.nI, sp.nI), np.float32)
        if self.gap is not None:
            for b in range(sp.nblock):
                for ib in range(self.maxNear):
                    pre, post, g = self.gap[b * self.maxNear + ib]
                    M[b, ib, pre, post] = g
        return M

    def dense_gapdist(self) -> np.ndarray:
        """The distances of the junctions in the V1_gapMat layout (the
        inhibitory-to-inhibitory part of V1_delayMat), zero elsewhere."""
        sp = self.spec
        M = np.zeros((sp.nblock, self.maxNear, sp.nI, sp.nI), np.float32)
        if self.gap is not None:
            for b in range(sp.nblock):
                for ib in range(int(self.nNear[b])):
                    pre, post, g = self.gap[b * self.maxNear + ib]
                    if pre.size:
                        M[b, ib, pre, post] = self._gap_pair_dist(b, ib, pre, post)
        return M

    # ---- recurrent formats -------------------------------------------------
    def masked_csr_torch(self, device="cuda"):
        """The masked CSR as torch tensors (int64 rowMask bits, int64 rowOff,
        float32 w) on `device`."""
        import torch
        if self.csr is not None:
            rm, ro, w = self.csr
            return (torch.as_tensor(rm, device=device), torch.as_tensor(ro, device=device).to(torch.int64),
                    torch.as_tensor(w, device=device))
        rm, ro, w = self.masked_csr()
        return (torch.as_tensor(rm.view(np.int64), device=device),
                torch.as_tensor(ro.astype(np.int64), device=device), torch.as_tensor(w, device=device))

    def n_rec_synapses_csr(self) -> int:
        if self.csr is not None:
            return int(self.csr[2].numel() if hasattr(self.csr[2], "numel") else self.csr[2].size)
        return self.n_rec_synapses()

    def masked_csr(self):
        """(rowMask u64 [nRow*nSeg], rowOff u32 [nRow*nSeg], w f32 [nnz]) with
        rows r = (b*maxNear + ib)*bs + pre (see include/vc_abi.h)."""
        if self.csr is not None:
            rm, ro, w = self.csr
            to_np = lambda x: x.cpu().numpy() if hasattr(x, "cpu") else np.asarray(x)
            return (to_np(rm).view(np.uint64), to_np(ro).astype(np.uint32),
                    to_np(w).astype(np.float32))
        sp = self.spec
        bs = sp.blockSize
        nSeg = bs // 64
        nRow = sp.nblock * self.maxNear * bs
        rowMask = np.zeros(nRow * nSeg, dtype=np.uint64)
        counts = np.zeros(nRow * nSeg, dtype=np.int64)
        ws = []
        for b in range(sp.nblock):
            for ib in range(self.maxNear):
                pre, post, w = self.rec_pairs[b * self.maxNear + ib]
                if pre.size == 0:
                    continue
                row = (b * self.maxNear + ib) * bs + pre.astype(np.int64)
                key = row * nSeg + (post.astype(np.int64) >> 6)
                bits = np.left_shift(np.uint64(1), (post & 63).astype(np.uint64))
                # entries are sorted by (pre, post) -> keys are sorted
                uk, start = np.unique(key, return_index=True)
                rowMask[uk] = np.bitwise_or.reduceat(bits, start)
                counts[uk] = np.diff(np.append(start, key.size))
                ws.append(w)
        rowOff = np.zeros(nRow * nSeg, dtype=np.int64)
        np.cumsum(counts[:-1], out=rowOff[1:])
        w = np.concatenate(ws).astype(np.float32) if ws else np.zeros(1, np.float32)
        if rowOff[-1] + counts[-1] >= 2 ** 32:
            raise ValueError("too many recurrent synapses for 32-bit offsets")
        return rowMask, rowOff.astype(np.uint32), w

    def dense_conmat(self) -> np.ndarray:
        """The reference's V1_conMat layout [nblock, maxNear, bs pre, bs post]."""
        sp = self.spec
        bs = sp.blockSize
        if self.csr is not None:                      # expand the masked CSR
            rm, ro, w = self.masked_csr()
            nSeg = bs // 64
            M = np.zeros((sp.nblock * self.maxNear * bs, bs), dtype=np.float32)
            bits = ((rm[:, None] >> np.arange(64, dtype=np.uint64)[None, :]) & np.uint64(1)).astype(bool)
            rank = np.cumsum(bits, axis=1) - 1
            idx = ro.astype(np.int64)[:, None] + rank
            seg = np.arange(rm.size)
            row, sg = seg // nSeg, seg % nSeg
            r_i, l_i = np.nonzero(bits)
            M[row[r_i], sg[r_i] * 64 + l_i] = w[idx[r_i, l_i]]
            return M.reshape(sp.nblock, self.maxNear, bs, bs)
        M = np.zeros((sp.nblock, self.maxNear, bs, bs), dtype=np.float32)
        for b in range(sp.nblock):
            for ib in range(self.maxNear):
                pre, post, w = self.rec_pairs[b * self.maxNear + ib]
                M[b, ib, pre, post] = w
        return M

    def ff_reverse(self):
        """(revOff u32 [nLGN+1], revTgt u32): for every LGN cell the synapses
        it feeds, encoded (neuron << 5) | slot (see vc_ff_events)."""
        sp = self.spec
        mx = self.sLGN.shape[0]
        if mx > 32:
            raise ValueError("event-driven LGN delivery needs <= 32 LGN inputs per neuron")
        slot = np.arange(mx)[:, None]
        valid = slot < self.nLGN_of[None, :].astype(np.int64)
        cells = self.idLGN[valid].astype(np.int64)
        tgt = ((np.broadcast_to(np.arange(sp.nV1)[None, :], valid.shape)[valid].astype(np.int64) << 5)
               | np.broadcast_to(slot, valid.shape)[valid])
        order = np.argsort(cells, kind="stable")
        cnt = np.bincount(cells, minlength=sp.nLGN)
        off = np.zeros(sp.nLGN + 1, np.int64)
        np.cumsum(cnt, out=off[1:])
        return off.astype(np.uint32), tgt[order].astype(np.uint32)

    def csr_dense_index(self) -> np.ndarray:
        """For every value of the masked CSR (in its order), the flat index of
        the same synapse in the dense conMat [nblock, maxNear, bs, bs]."""
        bs = self.spec.blockSize
        nSeg = bs // 64
        rm, ro, w = self.masked_csr()
        bits = ((rm[:, None] >> np.arange(64, dtype=np.uint64)[None, :]) & np.uint64(1)).astype(bool)
        rank = np.cumsum(bits, axis=1) - 1
        seg_i, lane_i = np.nonzero(bits)
        idx = ro.astype(np.int64)[seg_i] + rank[seg_i, lane_i]
        row, sg = seg_i // nSeg, seg_i % nSeg
        out = np.zeros(idx.size, np.int64)
        out[idx] = row.astype(np.int64) * bs + sg * 64 + lane_i
        return out

    def n_rec_synapses(self) -> int:
        return int(sum(p[0].size for p in self.rec_pairs))

    # ---- axonal delays ---------------------------------------------------------
    def rec_distances(self) -> np.ndarray:
        """Pre-post distance (mm, float32) of every strength of the masked
        CSR, in its order: x = (float)(x_pre - x_post) from the double
        positions, d = sqrtf(x*x + y*y) (src/preprocess/connect.cu:424-428,
        466; the reference records it in V1_delayMat).  vc_v1_rec_distances
        computes the same values on the device.  Distances read from a
        V1_delayMat file (rec_dist) are returned as they are."""
        if self.rec_dist is not None and self.csr is None:
            d = np.concatenate([np.asarray(x, np.float32) for x in self.rec_dist])
            return d if d.size else np.zeros(1, np.float32)
        bs = self.spec.blockSize
        nSeg = bs // 64
        rm, ro, w = self.masked_csr()
        bits = ((rm[:, None] >> np.arange(64, dtype=np.uint64)[None, :]) & np.uint64(1)).astype(bool)
        rank = np.cumsum(bits, axis=1) - 1
        seg_i, lane_i = np.nonzero(bits)
        idx = ro.astype(np.int64)[seg_i] + rank[seg_i, lane_i]
        row, sg = seg_i // nSeg, seg_i % nSeg
        b = row // (self.maxNear * bs)
        ib = (row // bs) % self.maxNear
        pre = self.nbr[b, ib].astype(np.int64) * bs + row % bs
        post = b * bs + sg * 64 + lane_i
        pos = np.asarray(self.pos, dtype=np.float64)
        x = (pos[pre, 0] - pos[post, 0]).astype(np.float32)
        y = (pos[pre, 1] - pos[post, 1]).astype(np.float32)
        d = np.zeros(max(int(w.size), 1), np.float32)
        d[idx] = np.sqrt(x * x + y * y)
        return d

    def dense_delaymat(self) -> np.ndarray:
        """The distances in the reference's V1_delayMat layout [nblock,
        maxNear, bs pre, bs post] (zero where there is no synapse)."""
        bs = self.spec.blockSize
        M = np.zeros(self.spec.nblock * self.maxNear * bs * bs, np.float32)
        n = self.n_rec_synapses_csr()
        if n:
            M[self.csr_dense_index()] = self.rec_distances()[:n]
        return M.reshape(self.spec.nblock, self.maxNear, bs, bs)


def train_depth(max_distance: float, speed: float, dt: float) -> int:
    """Steps of spike history the delays need: ceil((maxDistance /
    speedOfThought) / dt) + 1 in float32 (src/patch.cu:3813)."""
    f = np.float32
    return int(np.ceil((f(max_distance) / f(speed)) / f(dt))) + 1


# ---------------------------------------------------------------------------
def _block_grid(nblock: int) -> int:
    """Number of block columns of the sheet: blocks are laid out row-major on
    a ceil(sqrt(nblock))-wide grid whose last row may be partial."""
    return max(1, int(math.ceil(math.sqrt(nblock) - 1e-9)))


def _neighbours(nblock: int):
    """3x3 neighbourhood of every block, self first then ascending id."""
    side = _block_grid(nblock)
    rows = (nblock + side - 1) // side
    lists = []
    for b in range(nblock):
        bx, by = b % side, b // side
        nb = [b]
        for oy in (-1, 0, 1):
            for ox in (-1, 0, 1):
                x, y = bx + ox, by + oy
                if (ox or oy) and 0 <= x < side and 0 <= y < rows and y * side + x < nblock:
                    nb.append(y * side + x)
        lists.append([nb[0]] + sorted(nb[1:]))
    maxNear = max(len(l) for l in lists)
    nNear = np.array([len(l) for l in lists], dtype=np.uint32)
    nbr = np.zeros((nblock, maxNear), dtype=np.uint32)
    for b, l in enumerate(lists):
        nbr[b, :len(l)] = l
    return maxNear, nNear, nbr


def _weighted_topk(keysrc_w: np.ndarray, k: int, rng: np.random.Generator) -> np.ndarray:
    """Row-wise weighted sampling of k columns without replacement
    (Efraimidis & Spirakis 2006: largest log(u)/w)."""
    u = rng.random(keysrc_w.shape)
    with np.errstate(divide="ignore"):
        keys = np.log(u) / keysrc_w
    keys[~(keysrc_w > 0)] = -np.inf
    idx = np.argpartition(-keys, k - 1, axis=1)[:, :k]
    return idx


def lgn_parameters(spec: ModelSpec, lx, ly, on) -> dict:
    """Per-cell LGN parameters [2, nLGN] (centre, surround) of the fLGN set
    (src/patch.cu:2472-2530) for cells at (lx, ly) degrees, ON where `on`:
    DoG radii r = sigma*sqrt(2) (the reference's radius parameter), surround
    gain -kS/kC, the biphasic difference-of-gamma temporal kernel, and the
    logistic nonlinearity with its precached coefficients."""
    nL = len(lx)
    ecc = np.sqrt(lx * lx + ly * ly) * DEG2RAD
    polar = np.arctan2(ly, lx)
    rc = spec.sigC_deg * math.sqrt(2) * DEG2RAD
    rs = spec.sigS_deg * math.sqrt(2) * DEG2RAD
    kC = np.where(on, 1.0, -1.0)
    kS = -spec.kS_over_kC * kC
    two = lambda a, b_: np.stack([np.broadcast_to(a, (nL,)), np.broadcast_to(b_, (nL,))]).astype(np.float32)
    spont = np.full(nL, spec.spont, np.float32)
    c50 = np.full(nL, spec.c50, np.float32)
    K = np.full(nL, spec.sharpness, np.float32)
    logA, logB = logistic_ab(spont, c50, K)
    return dict(
        polar=two(polar, polar), ecc=two(ecc, ecc), rw=two(rc, rs), rh=two(rc, rs),
        orient=two(0.0, 0.0), k=two(kC, kS),
        tauR=two(spec.tauR, spec.tauR), tauD=two(spec.tauD, spec.tauD),
        delay=two(0.0, 0.0), ratio=two(spec.ratio, spec.ratio),
        nR=two(spec.nR, spec.nR), nD=two(spec.nD, spec.nD),
        covariant=np.ones(nL, np.float32),
        coneType=np.full((2, nL), 3, dtype=np.uint32),            # (L+M)/2 both regions
        c50=c50, sharpness=K, spont=spont, logA=logA, logB=logB,
    )


def build_network(spec: ModelSpec, with_frames: bool = True) -> Network:
    rng = np.random.default_rng(spec.seed)
    bs, nb = spec.blockSize, spec.nblock
    side = _block_grid(nb)
    blk_mm = spec.sheet_mm / side
    # ---- V1 positions
    pos = np.zeros((spec.nV1, 2))
    for b in range(nb):
        bx, by = b % side, b // side
        pos[b * bs:(b + 1) * bs, 0] = (bx + rng.random(bs)) * blk_mm
        pos[b * bs:(b + 1) * bs, 1] = (by + rng.random(bs)) * blk_mm
    local = np.arange(spec.nV1) % bs
    isE = local < spec.nE
    tonic = np.where(isE, spec.tonicDep[0], spec.tonicDep[1]).astype(np.float32)

    # ---- recurrent
    maxNear, nNear, nbr = _neighbours(nb)
    sig_mm = spec.sigma_um / 1000.0
    rec_pairs = [None] * (nb * maxNear)
    for b in range(nb):
        posts = np.arange(b * bs, (b + 1) * bs)
        cand = np.concatenate([np.arange(int(j) * bs, (int(j) + 1) * bs) for j in nbr[b, :nNear[b]]])
        cand_slot = np.repeat(np.arange(nNear[b]), bs)
        cE = cand[isE[cand]]
        cI = cand[~isE[cand]]
        chosen = []
        for cset, K, scale in ((cE, spec.KE, spec.sE), (cI, spec.KI, spec.sI)):
            if K == 0:
                continue
            d2 = ((pos[posts, None, :] - pos[None, cset, :]) ** 2).sum(-1)
            wgt = np.exp(-d2 / (2 * sig_mm * sig_mm))
            wgt[posts[:, None] == cset[None, :]] = 0.0        # no self-connection
            Kc = min(K, int((wgt > 0).sum(1).min()))
            sel = _weighted_topk(wgt, Kc, rng)
            pre_g = cset[sel]                                  # [bs, Kc]
            strength = (scale * rng.lognormal(spec.w_mu, spec.w_sigma, size=pre_g.shape)).astype(np.float32)
            post_g = np.broadcast_to(posts[:, None], pre_g.shape)
            chosen.append((pre_g.ravel(), post_g.ravel(), strength.ravel()))
        if not chosen:
            chosen.append((np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0, np.float32)))
        pre_all = np.concatenate([c[0] for c in chosen])
        post_all = np.concatenate([c[1] for c in chosen])
        w_all = np.concatenate([c[2] for c in chosen])
        pre_blk = pre_all // bs
        for ib in range(maxNear):
            if ib < nNear[b]:
                sel = pre_blk == nbr[b, ib]
                pl = (pre_all[sel] % bs).astype(np.uint32)
                ql = (post_all[sel] % bs).astype(np.uint32)
                ww = w_all[sel]
                order = np.lexsort((ql, pl))
                rec_pairs[b * maxNear + ib] = (pl[order], ql[order], ww[order])
            else:
                z = np.zeros(0, np.uint32)
                rec_pairs[b * maxNear + ib] = (z, z, np.zeros(0, np.float32))
        del cand_slot

    # ---- LGN cells on a jittered grid, checkerboard ON/OFF
    ns = spec.nLGN_side
    spacing = spec.field_deg / ns
    gx, gy = np.meshgrid(np.arange(ns), np.arange(ns))
    gx = gx.ravel(); gy = gy.ravel()
    lx = (gx + 0.5) * spacing - spec.field_deg / 2 + (rng.random(ns * ns) - 0.5) * 0.6 * spacing
    ly = (gy + 0.5) * spacing - spec.field_deg / 2 + (rng.random(ns * ns) - 0.5) * 0.6 * spacing
    on = ((gx + gy) % 2 == 0)
    nL = ns * ns
    lgn = lgn_parameters(spec, lx, ly, on)

    # ---- LGN -> V1 by Gaussian distance in the visual field
    from scipy.spatial import cKDTree
    vx = (pos[:, 0] / spec.sheet_mm - 0.5) * spec.field_deg * 0.9
    vy = (pos[:, 1] / spec.sheet_mm - 0.5) * spec.field_deg * 0.9
    tree = cKDTree(np.stack([lx, ly], 1))
    kq = min(nL, max(4 * spec.m, 64))
    dist, nn = tree.query(np.stack([vx, vy], 1), k=kq)
    sig = spec.ff_sigma_cells * spacing
    wq = np.exp(-dist ** 2 / (2 * sig * sig))
    if spec.m > 0:
        sel = _weighted_topk(wq, spec.m, rng)
        ids = np.take_along_axis(nn, sel, 1)                    # [nV1, m]
    else:
        ids = np.zeros((spec.nV1, 1), dtype=np.int64)
    ids.sort(axis=1)
    sR = np.where(isE, spec.sRatioLGN[0], spec.sRatioLGN[1])
    s_ff = (spec.sLGN * sR)[:, None] * (0.75 + 0.5 * rng.random(ids.shape))
    net = Network(
        spec=spec, pos=pos, tonicDep=tonic, maxNear=maxNear, nNear=nNear, nbr=nbr,
        rec_pairs=rec_pairs,
        nLGN_of=np.full(spec.nV1, spec.m, dtype=np.uint32),
        sLGN=np.ascontiguousarray(s_ff.T.astype(np.float32)),
        idLGN=np.ascontiguousarray(ids.T.astype(np.uint32)),
        lgn=lgn,
    )
    if spec.nFar > 0:
        net.far = far_connections(spec, pos, nbr, nNear)
    if spec.nGap > 0:
        net.gap = gap_junctions(spec, pos, nbr, nNear, maxNear)
    if with_frames:
        net.frames = grating_frames(spec)
    return net


def far_connections(spec: ModelSpec, pos: np.ndarray, nbr, nNear):
    """Synthetic far lists: spec.nFar presynaptic neurons per post neuron,
    uniform over the blocks that are not among its block's near neighbours
    (over every other block when all are near), sorted by id, strength sFar
    times lognormal; distances as the reference's genCon computes them
    (float32 of the double coordinate differences, src/preprocess/connect.cu
    :424-428).  Drawn from a generator of its own, so that the rest of the
    network does not depend on nFar."""
    rng = np.random.default_rng(spec.seed + 7919)
    bs, nb, nV1 = spec.blockSize, spec.nblock, spec.nV1
    off = np.zeros(nV1 + 1, np.uint32)
    pres, ws = [], []
    for b in range(nb):
        near = set(int(x) for x in nbr[b, :nNear[b]]) | {b}
        far_b = [j for j in range(nb) if j not in near] or [j for j in range(nb) if j != b]
        if not far_b:
            pres.append(np.zeros((bs, 0), np.int64))
            continue
        cand = (np.asarray(far_b)[:, None] * bs + np.arange(bs)[None, :]).ravel()
        k = min(spec.nFar, cand.size)
        if 4 * k <= cand.size:
            # k distinct of many: redraw the rows that repeat an index
            sel = rng.integers(0, cand.size, (bs, k))
            while k > 1:
                srt = np.sort(sel, axis=1)
                dup = (srt[:, 1:] == srt[:, :-1]).any(axis=1)
                if not dup.any():
                    break
                sel[dup] = rng.integers(0, cand.size, (int(dup.sum()), k))
        else:
            sel = np.argsort(rng.random((bs, cand.size)), axis=1)[:, :k]
        pres.append(np.sort(cand[sel], axis=1))
    pre = np.concatenate([p.reshape(-1) for p in pres]).astype(np.int64)
    cnt = np.concatenate([np.full(bs, p.shape[1]) for p in pres])
    off[1:] = np.cumsum(cnt)
    w = (spec.sFar * rng.lognormal(spec.w_mu, spec.w_sigma, pre.size)).astype(np.float32)
    post = np.repeat(np.arange(nV1), cnt)
    x = (pos[pre, 0] - pos[post, 0]).astype(np.float32)
    y = (pos[pre, 1] - pos[post, 1]).astype(np.float32)
    dist = np.sqrt(x * x + y * y).astype(np.float32)
    return off, pre.astype(np.uint32), w, dist


def gap_junctions(spec: ModelSpec, pos: np.ndarray, nbr, nNear, maxNear: int) -> list:
    """Synthetic gap junctions: spec.nGap partners per inhibitory neuron
    among the inhibitory neurons of its block and its near neighbours (not
    itself), drawn by a Gaussian of the distance (sigma_um), strength sGap
    times lognormal; per (b, ib) like Network.gap.  From a generator of its
    own, so that the rest of the network does not depend on nGap."""
    rng = np.random.default_rng(spec.seed + 104729)
    bs, nb, nE, nI = spec.blockSize, spec.nblock, spec.nE, spec.nI
    sig_mm = spec.sigma_um / 1000.0
    out = [None] * (nb * maxNear)
    z = np.zeros(0, np.uint32)
    for b in range(nb):
        posts = b * bs + nE + np.arange(nI)
        slots = np.arange(int(nNear[b]))
        cand = np.concatenate([int(nbr[b, ib]) * bs + nE + np.arange(nI) for ib in slots]) \
            if slots.size else np.zeros(0, np.int64)
        cand_slot = np.repeat(slots, nI)
        k = min(spec.nGap, max(cand.size - 1, 0))
        if k == 0:
            for ib in range(maxNear):
                out[b * maxNear + ib] = (z, z, np.zeros(0, np.float32))
            continue
        d2 = ((pos[posts, None, :] - pos[None, cand, :]) ** 2).sum(-1)
        wgt = np.exp(-d2 / (2 * sig_mm * sig_mm))
        wgt[posts[:, None] == cand[None, :]] = 0.0
        k = min(k, int((wgt > 0).sum(1).min()))
        if k == 0:
            for ib in range(maxNear):
                out[b * maxNear + ib] = (z, z, np.zeros(0, np.float32))
            continue
        sel = _weighted_topk(wgt, k, rng)                    # [nI, k]
        g = (spec.sGap * rng.lognormal(0.0, 0.5, size=sel.shape)).astype(np.float32)
        post_l = np.broadcast_to(np.arange(nI)[:, None], sel.shape)
        for ib in range(maxNear):
            if ib < nNear[b]:
                m = cand_slot[sel] == ib
                pl = ((cand[sel][m] - nE) % bs).astype(np.uint32)
                ql = post_l[m].astype(np.uint32)
                gg = g[m]
                order = np.lexsort((ql, pl))
                out[b * maxNear + ib] = (pl[order], ql[order], gg[order])
            else:
                out[b * maxNear + ib] = (z, z, np.zeros(0, np.float32))
    return out


def lgn_layer(spec: ModelSpec, seed: int | None = None) -> dict:
    """The LGN cells alone (no V1): the jittered nLGN_side^2 grid with
    checkerboard ON/OFF of build_network, drawn from its own generator."""
    rng = np.random.default_rng(spec.seed if seed is None else seed)
    ns = spec.nLGN_side
    spacing = spec.field_deg / ns
    gx, gy = np.meshgrid(np.arange(ns), np.arange(ns))
    gx = gx.ravel(); gy = gy.ravel()
    lx = (gx + 0.5) * spacing - spec.field_deg / 2 + (rng.random(ns * ns) - 0.5) * 0.6 * spacing
    ly = (gy + 0.5) * spacing - spec.field_deg / 2 + (rng.random(ns * ns) - 0.5) * 0.6 * spacing
    on = ((gx + gy) % 2 == 0)
    return lgn_parameters(spec, lx, ly, on)


def white_noise_frames(spec: ModelSpec, nframe: int, sigma: float = 0.3, mean: float = 0.5,
                       seed: int = 0) -> np.ndarray:
    """Gaussian white-noise frames [nFrame, 3, H, W] float32: mean + sigma N(0, 1)
    per pixel and channel (no clipping, so every pixel is a smooth variable)."""
    rng = np.random.default_rng(seed)
    f = mean + sigma * rng.standard_normal((nframe, 3, spec.height, spec.width))
    return np.ascontiguousarray(f.astype(np.float32))


def logistic_ab(spont, c50, K):
    """The precached logistic coefficients a, b of src/LGN_props.h:163-169
    (f(0) = spont, f(1) = 1), in float64 from the float32 parameters."""
    spont, c50, K = (np.asarray(x, np.float32).astype(np.float64) for x in (spont, c50, K))
    e1 = np.exp(K * c50)
    e2 = np.exp(K * c50 - K)
    a = (1 - spont) * (1 + e1) * (1 + e2) / (e1 - e2)
    b = 1 - a / (1 + e2)
    return a.astype(np.float32), b.astype(np.float32)


def n_frames(spec: ModelSpec) -> int:
    """Stimulus frames a run of spec.nt steps reaches (frame f enters at the
    first step whose end passes f * ntPerFrame / denorm steps), plus one."""
    tm = timing(spec)
    return (spec.nt * tm["denorm"] + tm["ntPerFrame"] - 1) // tm["ntPerFrame"] + 1


def grating_frames(spec: ModelSpec, nframe: int | None = None) -> np.ndarray:
    """Drifting sinusoidal grating, L = M = S = 0.5 (1 + c sin(...)), frames
    [nFrame, 3, H, W] (the stimulus.bin frame layout, grating.py:267-372)."""
    nf = n_frames(spec) if nframe is None else nframe
    W, H = spec.width, spec.height
    xs = ((np.arange(W) + 0.5) / W - 0.5) * spec.stim_deg
    ys = ((np.arange(H) + 0.5) / H - 0.5) * spec.stim_deg
    X, Y = np.meshgrid(xs, ys)
    th = spec.ori_deg * DEG2RAD
    u = X * math.cos(th) + Y * math.sin(th)
    t = np.arange(nf) / spec.frameRate
    ph = 2 * math.pi * (spec.sf_cpd * u[None] - spec.tf_hz * t[:, None, None]) + spec.phase0
    lum = (0.5 * (1.0 + spec.contrast * np.sin(ph))).astype(np.float32)
    return np.ascontiguousarray(np.repeat(lum[:, None], 3, axis=1))


# ---------------------------------------------------------------------------
def find_denorm(u1: int, u2: int) -> tuple[int, int]:
    """(denorm, norm) of find_denorm(u1 = frameRate, u2 = stepRate, MorN =
    true) (src/util/util.h:250-278): denorm is the least number of frames
    whose length is a whole number of steps; norm/denorm the phase a frame
    advances by.  With u2 % u1 = n != 0 it takes the largest i <= n dividing
    both n and u1 (norm = n / i, denorm = u1 / i), and norm = 1, denorm = u1
    when there is none above 1."""
    if u2 % u1 == 0:
        return 1, 1
    m, n = u1, u2 - u1 * (u2 // u1)
    for i in range(n, 1, -1):
        if n % i == 0 and m % i == 0:
            return m // i, n // i
    return m, 1


def timing(spec: ModelSpec):
    """Kernel-sampling and frame bookkeeping (src/patch.cu:4949-5046).  Time
    is counted in units of dt/denorm: a frame lasts ntPerFrame of them, so
    frame f covers steps [f ntPerFrame / denorm, (f+1) ntPerFrame / denorm),
    and ntPerFrame steps hold exactly denorm frames."""
    stepRate = round(1000.0 / spec.dt)
    if abs(stepRate * spec.dt - 1000.0) > 1e-9:
        raise ValueError("1000/dt must be an integer (src/patch.cu stepRate)")
    if stepRate < spec.frameRate:
        raise NotImplementedError("frameRate above the step rate (frames shorter than dt)")
    denorm, _ = find_denorm(spec.frameRate, stepRate)
    if (stepRate * denorm) % spec.frameRate != 0:          # src/patch.cu:5027
        raise ValueError("stepRate * denorm is not a multiple of frameRate")
    ntPerFrame = stepRate * denorm // spec.frameRate
    nRetrace = int(round(spec.tau / spec.dt))
    interval = nRetrace // spec.nKernelSample
    iT0 = interval // 2 if interval % 2 == 0 else 0
    maxFrame = (nRetrace * denorm + ntPerFrame - 1) // ntPerFrame + 1   # src/patch.cu:5043
    return dict(stepRate=stepRate, ntPerFrame=ntPerFrame, nRetrace=nRetrace,
                kernelSampleInterval=interval, iKernelSampleT0=iT0, maxFrame=maxFrame,
                denorm=denorm)


def max_sample_1d(net: Network) -> int:
    """Sample grid of parvo_maxConvol (src/patch.cu:5266-5300): round(mean
    surround/centre span ratio) * nSpatialSample1D, raised to a multiple of 32
    so that the 1024-thread patches tile it (the reference asserts this)."""
    sp = net.spec
    L = net.lgn
    nL = sp.nLGN
    rwc, rhc, rws, rhs = L["rw"][0], L["rh"][0], L["rw"][1], L["rh"][1]
    dx = L["ecc"][0] * np.cos(L["polar"][0]) - L["ecc"][1] * np.cos(L["polar"][1])
    dy = L["ecc"][0] * np.sin(L["polar"][0]) - L["ecc"][1] * np.sin(L["polar"][1])
    r = np.sqrt(dx * dx + dy * dy)
    R = np.maximum(abs(sp.nsig) * rws / math.sqrt(2), abs(sp.nsig) * rhs / math.sqrt(2))
    wS = abs(sp.nsig) * rwc / math.sqrt(2)
    hS = abs(sp.nsig) * rhc / math.sqrt(2)
    r = r + np.maximum(wS, hS)
    R = np.maximum(r, R)
    ratio = np.sqrt(R * R / wS / hS)
    m1 = int(round(float(ratio.mean()))) * sp.nSpatialSample1D
    if m1 * m1 < 1024:
        return 32
    return int(math.ceil(m1 / 32.0) * 32)


def view_geometry(spec: ModelSpec):
    """Flat retina (src/patch.cu:1630-1640): x0 = y0 = 0.5 and the scale that
    maps radians to normalised texture coordinates."""
    stim_range = spec.stim_deg / 2.0
    extent = 2.0 * (stim_range + spec.stim_buffer_deg)
    max_ecc = stim_range
    normEcc = max_ecc / extent
    nvd = normEcc / max_ecc / DEG2RAD
    return 0.5, 0.5, float(nvd)
