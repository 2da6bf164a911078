"""Autotuner of the one-step V1 kernel's launch shape.

The one-step kernel (vc_v1_step / vc_v1_run) is the path of runs whose grid
cannot be resident at once, such as many batched trials.  Two shapes are free,
since every choice gives bit-identical results (the spikes are listed and
summed in the same order whatever the shape):

* stepWG, the neurons per workgroup (64, 128 or 256, dividing the block):
  smaller workgroups free a block's workgroups to start and finish
  independently, larger ones share the neighbour-spike list among more
  wavefronts;
* gatherTile, the masked-CSR strengths each lane keeps in flight in the
  recurrent gather (4, 8 or 16): more memory-level parallelism against more
  registers (4 -> 79 VGPRs, 6 wavefronts per SIMD; 8 -> 83; 16 -> 91).

tune_step times every candidate on copies of the simulator's state, spike
buffers and the LGN input of the first frame period (none of the simulator's
own buffers change), and returns the fastest.  Results are kept per workload
shape for the life of the process.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import abi
from .abi import check, ptr

WGS = (64, 128, 256)
TILES = (4, 8, 16)
_cache: dict = {}


def candidates(block_size: int):
    return [(wg, t) for wg in WGS if block_size % wg == 0 for t in TILES]


def _key(sim) -> tuple:
    sp = sim.net.spec
    return (sp.nV1, sp.nblock, sp.blockSize, int(sim.net.maxNear), int(sim.wrec.numel()),
            sim.ev is not None)


def tune_step(sim, steps: int = 64, reps: int = 2, verbose: bool = False) -> tuple[int, int]:
    """(stepWG, gatherTile) of the fastest one-step launch for `sim`'s
    network; also stored in sim.p.  Call before the run (it computes the
    first frame period's LGN input if it is not there yet)."""
    key = _key(sim)
    if key in _cache:
        wg, t = _cache[key]
        sim.p.stepWG, sim.p.gatherTile = wg, t
        return wg, t
    if sim.wDense is not None:                   # dense rows: no tile to choose
        return int(sim.p.stepWG), 0
    if getattr(sim, "delay", False) or getattr(sim, "far_only", False):
        # the gather runs in vc_v1_rec_delay / vc_v1_rec_apply
        return int(sim.p.stepWG), int(sim.p.gatherTile)
    sp = sim.net.spec
    npf = sim.tm["ntPerFrame"]
    steps = min(steps, npf)
    lib, d = sim.lib, sim.dev
    # the first frame period's LGN output, then private copies of everything
    # a V1 step writes
    if sim.lgn_to == 0:
        sim._lgn_period(0)
    torch.cuda.synchronize()
    lgn = sim.lgn_sInfo2[0].clone()
    mask0 = sim.ffMask2[0].clone() if sim.ev is not None else None
    st0 = {k: v.clone() for k, v in sim.state.items()}
    spk_t = [(sim.masks[i].clone(), sim.og[i].clone(), sim.oh[i].clone()) for i in range(2)]
    # (with the gap sums and the noise state, which the steps also write)
    st0["lastNoise"] = sim.lastNoise.clone()
    if sim.gap is not None:
        st0["gap"] = sim.gap.clone()
    st = {k: v.clone() for k, v in st0.items()}
    s = abi.V1State(*(ptr(st[k]) for k in ("v", "w", "tBack", "gFF", "hFF", "gE", "hE", "gI",
                                           "hI")), ptr(sim.tonic), ptr(sim.gapS),
                    ptr(st.get("gap")), ptr(st["lastNoise"]))
    spk = [(m.clone(), g.clone(), h.clone()) for m, g, h in spk_t]
    spk2 = (abi.Spikes * 2)(*[abi.Spikes(ptr(m), ptr(g), ptr(h)) for m, g, h in spk])
    mask = mask0.clone() if mask0 is not None else None
    out = torch.empty(steps, sp.nV1, dtype=torch.float32, device=d)
    p = abi.V1Params.from_buffer_copy(sim.p)
    h = C.c_void_p(torch.cuda.current_stream(d).cuda_stream)

    def reset():
        for k, v in st0.items():
            st[k].copy_(v)
        for (m, g, hh), (m0, g0, h0) in zip(spk, spk_t):
            m.copy_(m0)
            g.copy_(g0)
            hh.copy_(h0)
        if mask is not None:
            mask.copy_(mask0)

    def run(n):
        check(lib.vc_v1_run(h, C.byref(p), C.byref(s), C.byref(sim.ff), C.byref(sim.rec), spk2,
                            sim.cur, ptr(lgn), sp.nLGN, ptr(mask) if mask is not None else None,
                            sp.nV1, sim.it, n, ptr(out), sp.nV1), "vc_v1_run")

    best = None
    for wg, t in candidates(sp.blockSize):
        p.stepWG, p.gatherTile = wg, t
        reset()
        run(1)                                   # first use of this kernel variant
        times = []
        for _ in range(reps):
            reset()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            run(steps)
            e1.record()
            e1.synchronize()
            times.append(e0.elapsed_time(e1) / steps * 1e3)
        us = min(times)
        if verbose:
            print(f"[tune] stepWG {wg:3d} gatherTile {t:2d}: {us:8.2f} us/step", flush=True)
        if best is None or us < best[0]:
            best = (us, wg, t)
    _, wg, t = best
    _cache[key] = (wg, t)
    sim.p.stepWG, sim.p.gatherTile = wg, t
    return wg, t
