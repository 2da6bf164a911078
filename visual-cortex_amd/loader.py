"""Assembling a Network from the reference's input files, and saving a
Network in those files (with a .cfg that names them), so that `patch_fast`
inputs drive this simulator and synthetic workloads can be exported.

Reads (see formats.py for the layouts and reference lines): V1_allpos,
connectome_cfg, neighborBlock, V1_conMat, V1_delayMat (noDelay = 0), V1_vec
(far connections; optional here, the reference requires it), V1_gapMat and
V1_gapVec (gap junctions, times gapRatio; optional here too),
LGN_V1_idList / LGN_V1_sList, LGN_vpos, LGN_surfaceID, fLGN, conStats,
V1_feature (when the cfg asks for features) and stimulus.bin.  The neuron and
synapse parameters come from the cfg keys (config.spec_from_config); sRatioV1
scales the V1 strengths on reading and sRatioLGN the LGN ones, and the tonic
drive of every neuron follows from tonicDep, CmoreDep, minTonicRatio and
boostOri with conStats' ffRatio and V1_feature's orientation preference, as
the reference computes them.
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
    # the reference draws new LGN parameters when useNewLGN is on (its
    # default, src/patch.cu:313, 1754-2459); this build reads them from fLGN
    if cfg.bool("useNewLGN", True):
        raise NotImplementedError("useNewLGN = true (drawing new LGN parameters) is not built: "
                                  "set useNewLGN = false to read them from fLGN")
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
    # LGN_vpos gives the cell count and types the reference sizes the LGN by
    vpos = F.read_lgn_vpos(cfg.path("fLGN_vpos"))
    if vpos["nLGN"] != nL:
        raise ValueError(f"LGN_vpos holds {vpos['nLGN']} cells, fLGN {nL}")
    if "type" in L and not np.array_equal(np.asarray(L["type"], np.uint32), vpos["type"]):
        raise ValueError("LGN_vpos and fLGN disagree on the LGN cell types")
    if vpos["max_ecc"] > st["range_deg"]:            # src/patch.cu:1536-1538
        raise ValueError(f"LGN_vpos max_ecc {vpos['max_ecc']} exceeds the stimulus range "
                         f"{st['range_deg']}")
    F.read_lgn_surface_id(cfg.path("fLGN_surfaceID"), nL)
    spec = replace(spec, nblock=nblock, blockSize=bs, nE=int(typeAcc[0]), nLGN_override=nL,
                   width=frames.shape[3], height=frames.shape[2],
                   stim_deg=2.0 * st["range_deg"], stim_buffer_deg=st["buffer_deg"],
                   m=int(nof.max()) if nof.size else 0)
    ffRatio = F.read_constats(cfg.path("fConStats"), nType, nblock * bs)
    feature = None
    if feature_needed(cfg):
        feature = F.read_v1_feature(cfg.path("fV1_feature"), nblock * bs)
    tonic = tonic_drive(cfg, spec.tonicDep, typeAcc, nblock, bs, ffRatio, feature)
    gap, gap_far, gap_dist = read_gap_junctions(cfg, nblock, bs, int(typeAcc[0]), maxNear, nNear)
    rec_dist = None
    if spec.noDelay == 0:                    # the reference's own distances
        if gap is not None:
            nE = int(typeAcc[0])
            shifted = [(p[0] + nE, p[1] + nE) for p in gap]
            rec_dist, gap_dist = F.read_delaymat_pairs(cfg.path("fV1_delayMat"), nblock, bs,
                                                       pairs, extra=shifted)
        else:
            rec_dist = F.read_delaymat_pairs(cfg.path("fV1_delayMat"), nblock, bs, pairs)
    far = None
    vec = Path(cfg.path("fV1_vec"))
    if vec.exists():
        far = F.read_v1_vec(vec, nblock * bs, bs, spec.nE, cfg.floats("sRatioV1"))
        if int(far[0][-1]) == 0:
            far = None
    return Network(spec=spec, pos=pos, tonicDep=tonic, maxNear=maxNear, nNear=nNear, nbr=nbr,
                   rec_pairs=pairs, nLGN_of=nof, sLGN=sLGN, idLGN=idLGN, lgn=lgn,
                   frames=frames, init_lms=tuple(st["init_lms"]), rec_dist=rec_dist, far=far,
                   gap=gap, gap_dist=gap_dist, gap_far=gap_far)


def read_gap_junctions(cfg, nblock: int, bs: int, nE: int, maxNear: int, nNear):
    """(gap, gap_far, None) from V1_gapMat and V1_gapVec, strengths times
    gapRatio (one inhibitory type: a single value, src/patch.cu:736-744,
    3395, 3718).  Files that are absent hold no junctions here (the
    reference requires both); gapRatio must be given when a file holds a
    junction."""
    nI = bs - nE
    mI = nblock * nI
    gp, gv = Path(cfg.path("fV1_gapMat")), Path(cfg.path("fV1_gapVec"))
    gap = F.read_gapmat_pairs(gp, nblock, nI, maxNear, nNear) if gp.exists() else None
    far = F.read_gapvec(gv, mI) if gv.exists() else None
    if gap is not None and not any(p[0].size for p in gap):
        gap = None
    if far is not None and int(far[0][-1]) == 0:
        far = None
    if gap is None and far is None:
        return None, None, None
    ratio = cfg.floats("gapRatio")
    if not ratio or len(ratio) != 1:
        raise ValueError("gapRatio must hold one value (nTypeI = 1) when V1_gapMat or V1_gapVec "
                         "holds a junction (src/patch.cu:736-744)")
    r = np.float32(ratio[0])
    if r < 0:
        raise ValueError("gapRatio must be >= 0")
    if gap is not None:
        gap = [(a, b, (g * r).astype(np.float32)) for a, b, g in gap]
    if far is not None:
        far = (far[0], far[1], (far[2] * r).astype(np.float32), far[3])
    return gap, far, None


def _per_type(vals, nType, key, per=1):
    """A cfg vector of 1 or nType entries (of `per` values each) -> nType*per."""
    if len(vals) == per and nType > 1:
        vals = list(vals) * nType
    if len(vals) != nType * per:
        raise ValueError(f"{key} needs {per} or {nType * per} values, got {len(vals)}")
    return list(vals)


def _orientation_options(cfg, nType: int = 2):
    """(readFeature, fOri[nType], boostOri[2*nType] or None, nOri) after the
    reference's checks of nOri, iOri and boostOri (src/patch.cu:2843-2910)."""
    nOri = cfg.int("nOri", 0)
    read_feature = cfg.bool("readFeature", True)
    boost = cfg.floats("boostOri")
    iOri = cfg.floats("iOri")
    fOri = [0.0] * nType
    if boost is not None:
        boost = _per_type(boost, nType, "boostOri", 2)
    if nOri > 0:
        if boost is not None:
            read_feature = True
            if any(boost[2 * t + 1] != 0 for t in range(nType)) and iOri is None:
                raise ValueError("when boostOri[std] is set, iOri needs to be set")
        if iOri is not None:
            iOri = _per_type([int(x) for x in iOri], nType, "iOri")
            read_feature = True
            f32 = np.float32
            fOri = [float(np.fmod(f32(f32(i % nOri) / f32(nOri)) + f32(0.5), f32(1.0)))
                    for i in iOri]
        elif boost is not None:
            for t in range(nType):
                boost[2 * t + 1] = 0.0
    elif boost is not None:
        for t in range(nType):
            boost[2 * t + 1] = 0.0
    return read_feature, fOri, boost, nOri


def feature_needed(cfg) -> bool:
    """Whether the run reads V1_feature: readFeature (default true, or set by
    nOri > 0 with iOri/boostOri).  The reference then asserts nOri > 0 and
    both iOri and boostOri (src/patch.cu:2914-2931); so does this loader."""
    read_feature, _, boost, nOri = _orientation_options(cfg)
    if read_feature and (nOri <= 0 or cfg.floats("iOri") is None or boost is None):
        raise ValueError("readFeature needs nOri > 0, iOri and boostOri (the reference "
                         "asserts them); set readFeature = false otherwise")
    return read_feature


def tonic_drive(cfg, tonicDep, typeAcc, nblock: int, bs: int, ffRatio, feature):
    """Tonic depolarisation of every neuron (src/patch.cu:4118-4222), float32:
    tonicDep[type] * boost * (minTonicRatio + (1 - minTonicRatio) * (1 -
    ffRatio - (1 - max ffRatio)) / (max ffRatio - min ffRatio)) with CmoreDep
    (the min/max over the type's neurons in every block), tonicDep[type] *
    boost without; boost is the von Mises orientation boost of boostOri
    around iOri/nOri when the orientation features are read, else 1."""
    f32 = np.float32
    nType = len(typeAcc)
    typeAcc0 = [0] + [int(x) for x in typeAcc]
    nV1 = nblock * bs
    tonic = [f32(x) for x in tonicDep]
    more = cfg.bool("CmoreDep", True)
    mtr = cfg.floats("minTonicRatio")
    if mtr is None:
        raise ValueError("minTonicRatio needs 1 or nType values")
    mtr = [f32(x) for x in _per_type(mtr, nType, "minTonicRatio")]
    read_feature, fOri, boost, nOri = _orientation_options(cfg, nType)
    ff = np.asarray(ffRatio, f32)
    local = np.arange(nV1) % bs
    typ = np.searchsorted(np.asarray(typeAcc0[1:]), local, side="right")
    mx = np.array([max(ff.reshape(nblock, bs)[:, typeAcc0[j]:typeAcc0[j + 1]].max(), f32(0))
                   for j in range(nType)], f32)
    mn = np.array([min(ff.reshape(nblock, bs)[:, typeAcc0[j]:typeAcc0[j + 1]].min(), f32(1))
                   for j in range(nType)], f32)
    boosted = np.ones(nV1, f32)
    if boost is not None and read_feature and nOri > 0:
        op = np.asarray(feature[1], f32)
        dOri = (op - np.asarray(fOri, f32)[typ]).astype(f32)
        dOri = np.where(dOri > f32(0.5), f32(1) - dOri, dOri).astype(f32)
        dOri = np.where(dOri < f32(-0.5), f32(1) + dOri, dOri).astype(f32)
        base = np.asarray(boost[0::2], f32)[typ]
        kap = np.asarray(boost[1::2], f32)[typ]
        amp = (f32(1) - base).astype(f32)
        vm = (f32(1) - np.exp(f32(-2) * kap)).astype(f32)
        cosv = np.cos((dOri.astype(np.float64) * np.pi * 2.0).astype(f32)).astype(f32)
        vmx = np.where(vm == 0, f32(1), vm).astype(f32)
        ev = np.exp((kap * (cosv - f32(1))).astype(f32)).astype(f32)
        bv = (base + (amp * ((ev - f32(1)) + vm)) / vmx).astype(f32)   # a*(e-1+vm)/vm
        boosted = np.where(vm == 0, base, bv).astype(f32)
    td = np.asarray(tonic, f32)[typ]
    if more:
        if np.any(boosted > 1):
            raise ValueError("an orientation boost above 1 (the reference asserts boost <= 1)")
        rel = (f32(1) - ff).astype(f32)
        minT = (f32(1) - mx[typ]).astype(f32)
        maxT = (f32(1) - mn[typ]).astype(f32)
        m = np.asarray(mtr, f32)[typ]
        span = np.where(mx[typ] > mn[typ], maxT - minT, f32(1)).astype(f32)
        scaled = ((td * boosted) * (m + ((f32(1) - m) * (rel - minT)) / span)).astype(f32)
        out = np.where(mx[typ] > mn[typ], scaled, (td * boosted).astype(f32))
    else:
        out = (td * boosted).astype(f32)
    return np.ascontiguousarray(out, np.float32)


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


def lgn_positions(net: Network):
    """The LGN_vpos content of a network's cells: all ipsilateral, positions
    x = ecc cos(polar), y = ecc sin(polar) in degrees, type 0 (as fLGN is
    written), max_ecc the largest eccentricity (deg)."""
    L = net.lgn
    ecc_deg = np.asarray(L["ecc"][0], np.float64) * 180.0 / math.pi
    pol = np.asarray(L["polar"][0], np.float64)
    x = (ecc_deg * np.cos(pol)).astype(np.float32)
    y = (ecc_deg * np.sin(pol)).astype(np.float32)
    n = x.size
    return dict(nLGN_I=n, nLGN_C=0, max_ecc=float(ecc_deg.max()) if n else 0.0,
                x0=float(x.min()), xspan=float(np.ptp(x)), y0=float(y.min()),
                yspan=float(np.ptp(y)), x=x, y=y, types=np.zeros(n, np.uint32),
                polar=pol.astype(np.float32), ecc_deg=ecc_deg.astype(np.float32))


def lgn_surface_ids(n: int):
    """A square LGN surface with one texel per cell, row-major."""
    side = max(1, int(math.ceil(math.sqrt(n))))
    i = np.arange(n)
    return (i % side).astype(np.int32), (i // side).astype(np.int32)


def save_network(net: Network, folder, suffix: str = "vc", tonic: dict | None = None) -> Path:
    """Write `net` in the reference's input formats under `folder` and return
    the path of a cfg that names them (inputFolder = resourceFolder =
    outputFolder = folder).  The tonic drive is written as it is (CmoreDep =
    false, no orientation boost) unless `tonic` gives the cfg keys (CmoreDep,
    minTonicRatio, boostOri, iOri, nOri, readFeature) and the arrays ffRatio
    [nV1] and feature [nFeature][nV1] of conStats and V1_feature."""
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
    F.write_gapmat(folder / f"V1_gapMat{suf}", net)
    F.write_gapvec(folder / f"V1_gapVec{suf}", sp.nblock * sp.nI, net.gap_far)
    # LGN_V1 lists hold the un-scaled strengths (read_LGN multiplies by sRatioLGN)
    isE = (np.arange(sp.nV1) % sp.blockSize) < sp.nE
    sR = np.where(isE, sp.sRatioLGN[0], sp.sRatioLGN[1]).astype(np.float32)
    ids = [net.idLGN[:net.nLGN_of[n], n] for n in range(sp.nV1)]
    ss = [net.sLGN[:net.nLGN_of[n], n] / sR[n] for n in range(sp.nV1)]
    F.write_lgn_v1(folder / f"LGN_V1_idList{suf}", folder / f"LGN_V1_sList{suf}", ids, ss)
    L = dict(net.lgn)
    F.write_flgn(folder / f"LGN{suf}", L)
    vp = lgn_positions(net)
    F.write_lgn_vpos(folder / f"LGN_vpos{suf}", vp["nLGN_I"], vp["nLGN_C"], vp["max_ecc"],
                     vp["x0"], vp["xspan"], vp["y0"], vp["yspan"], vp["x"], vp["y"], vp["types"],
                     vp["polar"], vp["ecc_deg"])
    F.write_lgn_surface_id(folder / f"LGN_surfaceID{suf}", *lgn_surface_ids(sp.nLGN))
    tonic = tonic or {}
    ff = tonic.get("ffRatio")
    if ff is None:      # a stand-in: the share of LGN strength in each neuron's input
        fsum = net.sLGN.astype(np.float64).sum(0)
        rsum = np.zeros(sp.nV1)
        for b in range(sp.nblock):
            for ib in range(net.maxNear):
                pre, post, w = net.rec_pairs[b * net.maxNear + ib] if net.csr is None else \
                    (np.zeros(0, int), np.zeros(0, int), np.zeros(0))
                np.add.at(rsum, b * sp.blockSize + post.astype(np.int64), w)
        ff = (fsum / np.maximum(fsum + rsum, 1e-30)).astype(np.float32)
    F.write_constats(folder / f"conStats{suf}", 2, ff)
    if tonic.get("feature") is not None:
        F.write_v1_feature(folder / f"V1_feature{suf}", tonic["feature"])
    F.write_stimulus(folder / "stimulus.bin", net.frames, net.init_lms,
                     buffer_deg=sp.stim_buffer_deg, range_deg=sp.stim_deg / 2.0)
    t = lambda v: ", ".join(repr(float(x)) for x in v)
    lines = [
        f"inputFolder = {folder}/", f"resourceFolder = {folder}/", f"outputFolder = {folder}/",
        f"res_suffix = {suffix}", f"conV1_suffix = {suffix}", f"conLGN_suffix = {suffix}",
        f"output_suffix = {suffix}", "flat_retina = true", f"noDelay = {sp.noDelay}",
        f"speedOfThought = {sp.speedOfThought}", "learning = 0",
        f"InhGap = {'true' if sp.InhGap else 'false'}", "gapRatio = 1.0",
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
    keys = dict(CmoreDep="false", minTonicRatio="1.0", readFeature="false", useNewLGN="false")
    keys.update({k: v for k, v in tonic.items() if k not in ("ffRatio", "feature")})
    for k, v in keys.items():
        if isinstance(v, (list, tuple)):
            for j in range(0, len(v), 2 if k == "boostOri" else len(v)):
                lines.append(f"{k} = {t(v[j:j + (2 if k == 'boostOri' else len(v))])}")
        else:
            lines.append(f"{k} = {str(v).lower() if isinstance(v, bool) else v}")
    for k in ("riseTimeFF", "decayTimeFF", "riseTimeE", "decayTimeE", "riseTimeI", "decayTimeI"):
        lines.append(f"{k} = {t(getattr(sp, k))}")
    for k in ("pFF", "pE", "pI", "pBG"):
        for row in getattr(sp, k):                   # repeated keys append, one row per type
            lines.append(f"{k} = {t(row)}")
    cfg_path = folder / f"patchV1-{suffix}.cfg"
    cfg_path.write_text("\n".join(lines) + "\n")
    return cfg_path
