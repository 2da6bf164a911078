"""Assembling a Network from the reference's input files, and saving a
Network in those files (with a .cfg that names them), so that `patch_fast`
inputs drive this simulator and synthetic workloads can be exported.

Reads (see formats.py for the layouts and reference lines): V1_allpos,
connectome_cfg, neighborBlock, V1_conMat, V1_delayMat (noDelay = 0), V1_vec
(far connections; optional here, the reference requires it),
LGN_V1_idList / LGN_V1_sList,
fLGN and stimulus.bin.  The neuron and synapse parameters come from the cfg
keys (config.spec_from_config); sRatioV1 scales the V1 strengths on reading
and sRatioLGN the LGN ones, as the reference does.
"""
from __future__ import annotations

import math
from dataclasses import replace
from pathlib import Path

import numpy as np

from . import config as cfgmod
from . import formats as F
from .network import Network, logistic_ab


def load_network(cfg: cfgmod.Config) -> Network:
    spec = cfgmod.spec_from_config(cfg)
    nblock, bs, pos = F.read_v1_allpos(cfg.path("fV1_allpos"))
    nType, nTypeI, typeAcc = F.read_connectome_cfg(cfg.path("fConnectome_cfg"))
    if nType != 2 or nTypeI != 1:
        raise NotImplementedError("one E and one I type are supported (nType = 2, nTypeI = 1)")
    if int(typeAcc[-1]) != bs:
        raise ValueError("connectome_cfg typeAcc does not end at blockSize")
    nNear, lists = F.read_neighbor_block(cfg.path("fNeighborBlock"), nblock)
    maxNear, pairs = F.read_conmat_pairs(cfg.path("fV1_conMat"), nblock, bs)
    pairs = scale_v1(pairs, cfg.floats("sRatioV1"), int(typeAcc[0]))
    if int(nNear.max()) > maxNear:
        raise ValueError("neighborBlock has more near neighbours than V1_conMat holds")
    nbr = np.zeros((nblock, maxNear), np.uint32)
    for b in range(nblock):
        nbr[b, :nNear[b]] = lists[b][:nNear[b]]
    sRatio = cfg.floats("sRatioLGN", list(spec.sRatioLGN))
    nof, sLGN, idLGN = F.read_lgn_v1(cfg.path("fLGN_V1_ID"), cfg.path("fLGN_V1_s"), sRatio, typeAcc)
    L = F.read_flgn(cfg.path("fLGN"))
    st = F.read_stimulus(cfg.path("fStimulus"))
    nL = L["polar"].shape[1]
    a, b_ = logistic_ab(L["spont"], L["c50"], L["sharpness"])
    lgn = {k: np.ascontiguousarray(L[k], np.float32) for k in
           ("polar", "ecc", "rw", "rh", "orient", "k", "tauR", "tauD", "delay", "ratio", "nR", "nD")}
    lgn.update(covariant=np.ascontiguousarray(L["covariant"], np.float32),
               coneType=np.ascontiguousarray(L["coneType"], np.uint32),
               c50=L["c50"].astype(np.float32), sharpness=L["sharpness"].astype(np.float32),
               spont=L["spont"].astype(np.float32), logA=a, logB=b_)
    frames = st["frames"]
    spec = replace(spec, nblock=nblock, blockSize=bs, nE=int(typeAcc[0]), nLGN_override=nL,
                   width=frames.shape[3], height=frames.shape[2],
                   stim_deg=2.0 * st["range_deg"], stim_buffer_deg=st["buffer_deg"],
                   m=int(nof.max()) if nof.size else 0)
    tonic = np.where((np.arange(nblock * bs) % bs) < spec.nE, spec.tonicDep[0],
                     spec.tonicDep[1]).astype(np.float32)
    rec_dist = None
    if spec.noDelay == 0:                    # the reference's own distances
        rec_dist = F.read_delaymat_pairs(cfg.path("fV1_delayMat"), nblock, bs, pairs)
    far = None
    vec = Path(cfg.path("fV1_vec"))
    if vec.exists():
        far = F.read_v1_vec(vec, nblock * bs, bs, spec.nE, cfg.floats("sRatioV1"))
        if int(far[0][-1]) == 0:
            far = None
    return Network(spec=spec, pos=pos, tonicDep=tonic, maxNear=maxNear, nNear=nNear, nbr=nbr,
                   rec_pairs=pairs, nLGN_of=nof, sLGN=sLGN, idLGN=idLGN, lgn=lgn,
                   frames=frames, init_lms=tuple(st["init_lms"]), rec_dist=rec_dist, far=far)


def scale_v1(pairs, sratio, nE: int):
    """conMat *= sRatioV1[preType*2 + postType] as the reference does on
    reading (src/patch.cu:3655-3680; one value applies to every type pair,
    src/patch.cu:725-734), in float32; pairs whose scaled strength is not
    positive are dropped (recal_G skips zero strengths)."""
    if sratio is None:
        return pairs
    sr = [float(x) for x in sratio]
    if len(sr) == 1:
        sr = sr * 4
    if len(sr) != 4:
        raise ValueError(f"sRatioV1 has {len(sr)} values, expected 1 or nType*nType = 4")
    sr = np.asarray(sr, np.float32)
    out = []
    for pre, post, w in pairs:
        k = (pre >= nE).astype(np.int64) * 2 + (post >= nE).astype(np.int64)
        ws = (w.astype(np.float32) * sr[k]).astype(np.float32)
        keep = ws > 0
        out.append((pre[keep], post[keep], ws[keep]))
    return out


def save_network(net: Network, folder, suffix: str = "vc") -> Path:
    """Write `net` in the reference's input formats under `folder` and return
    the path of a cfg that names them (inputFolder = resourceFolder =
    outputFolder = folder)."""
    sp = net.spec
    folder = Path(folder)
    folder.mkdir(parents=True, exist_ok=True)
    suf = f"-{suffix}.bin"
    F.write_v1_allpos(folder / f"V1_allpos{suf}", sp.nblock, sp.blockSize, net.pos)
    F.write_connectome_cfg(folder / f"connectome_cfg{suf}", [sp.nE, sp.blockSize], 1)
    lists = [list(net.nbr[b, :net.nNear[b]]) for b in range(sp.nblock)]
    F.write_neighbor_block(folder / f"neighborBlock{suf}", net.nNear, lists)
    F.write_conmat(folder / f"V1_conMat{suf}", net)
    F.write_delaymat(folder / f"V1_delayMat{suf}", net)
    F.write_v1_vec(folder / f"V1_vec{suf}", sp.nV1, getattr(net, "far", None))
    # LGN_V1 lists hold the un-scaled strengths (read_LGN multiplies by sRatioLGN)
    isE = (np.arange(sp.nV1) % sp.blockSize) < sp.nE
    sR = np.where(isE, sp.sRatioLGN[0], sp.sRatioLGN[1]).astype(np.float32)
    ids = [net.idLGN[:net.nLGN_of[n], n] for n in range(sp.nV1)]
    ss = [net.sLGN[:net.nLGN_of[n], n] / sR[n] for n in range(sp.nV1)]
    F.write_lgn_v1(folder / f"LGN_V1_idList{suf}", folder / f"LGN_V1_sList{suf}", ids, ss)
    L = dict(net.lgn)
    F.write_flgn(folder / f"LGN{suf}", L)
    F.write_stimulus(folder / "stimulus.bin", net.frames, net.init_lms,
                     buffer_deg=sp.stim_buffer_deg, range_deg=sp.stim_deg / 2.0)
    t = lambda v: ", ".join(repr(float(x)) for x in v)
    lines = [
        f"inputFolder = {folder}/", f"resourceFolder = {folder}/", f"outputFolder = {folder}/",
        f"res_suffix = {suffix}", f"conV1_suffix = {suffix}", f"conLGN_suffix = {suffix}",
        f"output_suffix = {suffix}", "flat_retina = true", f"noDelay = {sp.noDelay}",
        f"speedOfThought = {sp.speedOfThought}", "learning = 0",
        f"noFarDelay = {sp.noDelay if sp.noFarDelay < 0 else sp.noFarDelay}",
        f"seed = {sp.seed}", f"dt = {sp.dt}", f"nt = {sp.nt}", f"frameRate = {sp.frameRate}",
        f"iModel = {sp.iModel}", f"tau = {sp.tau}", f"nKernelSample = {sp.nKernelSample}",
        f"nSpatialSample1D = {sp.nSpatialSample1D}", f"nsig = {sp.nsig}",
        f"frRatioLGN = {sp.frRatioLGN}", f"convolRatio = {sp.convolRatio}",
        f"sRatioLGN = {t(sp.sRatioLGN)}", "sRatioV1 = 1.0", f"bgRate = {sp.bgRate}",
        f"bgStrength = {t(sp.bgStrength)}",
    ]
    for k in ("vR", "vThres", "gL", "C", "tRef", "vT", "deltaT", "tau_w", "a", "b", "tonicDep"):
        lines.append(f"{k} = {t(getattr(sp, k))}")
    for k in ("riseTimeFF", "decayTimeFF", "riseTimeE", "decayTimeE", "riseTimeI", "decayTimeI"):
        lines.append(f"{k} = {t(getattr(sp, k))}")
    for k in ("pFF", "pE", "pI", "pBG"):
        for row in getattr(sp, k):                   # repeated keys append, one row per type
            lines.append(f"{k} = {t(row)}")
    cfg_path = folder / f"patchV1-{suffix}.cfg"
    cfg_path.write_text("\n".join(lines) + "\n")
    return cfg_path
