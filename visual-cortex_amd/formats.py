"""The reference simulator's binary input and output files (little-endian;
Size/PosInt = u32, Float = f32), read and written so that this package is a
drop-in for `patch_fast`'s file interface.

Inputs (readers + writers, the writers let a synthetic Network be saved in
the reference's formats):
  stimulus.bin            src/patch.cu:1393-1470, frames read at :5985-5993
  V1_allpos<res>.bin      src/patch.cu:2754-2790
  connectome_cfg<conV1>   src/patch.cu:634-645
  V1_conMat<conV1>.bin    src/patch.cu:3441-3461, 3649-3687 (dense, streamed
                          block by block into the masked-CSR rows)
  V1_delayMat<conV1>.bin  src/patch.cu:3450-3461, 3683-3684 (the same layout:
                          pre-post distance in mm of every pair, written by
                          src/preprocess/genCon.cu:511-516, 911)
  neighborBlock<conV1>    src/patch.cu:3481-3512
  V1_vec<conV1>.bin       src/patch.cu:3308-3352 (far connections: per post
                          neuron presynaptic ids, strengths, distances)
  LGN_V1_idList / sList   src/util/util.h:161-185, src/patch.h:30-77
  fLGN (old LGN setup)    src/patch.cu:2472-2530
  LGN_vpos<res>.bin       src/patch.cu:1524-1533, 1697-1714, 1758-1772
  LGN_surfaceID<res>.bin  src/patch.cu:2573-2598
  conStats<conV1>.bin     src/patch.cu:4119-4133 (ffRatio of every neuron)
  V1_feature<res>.bin     src/patch.cu:2915-2927 (OD, OP, ... per neuron)
Outputs:
  rawData<out>.bin        header src/patch.cu:4688-4716; per step sInfo
                          (:6139-6164), the dep/w/v/gFF/hFF block, then
                          gE/gI/hE/hI and gap (src/stats.cu:38-97); read back
                          with the semantics of src/readPatchOutput.py:291-351
  LGN_fr<out>.bin         src/patch.cu:4666-4686, per step at :6346
  LGN_sp<out>.bin         per step sInfo[nLGN] (src/patch.cu:6422-6424)
  patchV1_cfg<out>.bin    the run's parameters, src/patch.cu:6953-7028 (read
                          by src/readPatchOutput.py:32-88)
"""
from __future__ import annotations

import io
import struct
from pathlib import Path

import numpy as np


def _u32(f, n=1):
    a = np.fromfile(f, dtype="<u4", count=n)
    if a.size != n:
        raise ValueError(f"{getattr(f, 'name', 'file')}: truncated")
    return a if n > 1 else int(a[0])


def _f32(f, n):
    a = np.fromfile(f, dtype="<f4", count=n)
    if a.size != n:
        raise ValueError(f"{getattr(f, 'name', 'file')}: truncated")
    return a


# ------------------------------------------------------------------ stimulus
def write_stimulus(path, frames: np.ndarray, init_lms=(0.5, 0.5, 0.5), buffer_deg=0.0,
                   range_deg=2.0, neye=1):
    """frames [nFrame, 3, H, W] float32 (L, M, S)."""
    nF, c, H, W = frames.shape
    assert c == 3
    with open(path, "wb") as f:
        f.write(struct.pack("<i", -1))
        f.write(struct.pack("<3I", nF, H, W))
        f.write(struct.pack("<3f", *init_lms))
        f.write(struct.pack("<2f", buffer_deg, range_deg))
        f.write(struct.pack("<I", neye))
        f.write(np.ascontiguousarray(frames, dtype="<f4").tobytes())


def read_stimulus(path):
    """Returns dict(frames [nFrame,3,H,W], init_lms, buffer_deg, range_deg, neye)."""
    with open(path, "rb") as f:
        inputType = struct.unpack("<i", f.read(4))[0]
        if inputType != -1:
            raise ValueError(f"{path}: inputType {inputType} (only -1, an LMS movie, is supported)")
        nF, H, W = struct.unpack("<3I", f.read(12))
        init = struct.unpack("<3f", f.read(12))
        buf, rng = struct.unpack("<2f", f.read(8))
        neye = struct.unpack("<I", f.read(4))[0]
        frames = _f32(f, nF * 3 * H * W).reshape(nF, 3, H, W)
    return dict(frames=frames, init_lms=init, buffer_deg=buf, range_deg=rng, neye=neye)


# ------------------------------------------------------------------ V1 geometry
def write_v1_allpos(path, nblock, blockSize, pos_mm: np.ndarray):
    x0, y0 = float(pos_mm[:, 0].min()), float(pos_mm[:, 1].min())
    xs, ys = float(np.ptp(pos_mm[:, 0])), float(np.ptp(pos_mm[:, 1]))
    with open(path, "wb") as f:
        f.write(struct.pack("<3I", nblock, blockSize, 2))
        f.write(struct.pack("<4d", x0, xs, y0, ys))
        f.write(np.ascontiguousarray(pos_mm[:, 0], dtype="<f8").tobytes())
        f.write(np.ascontiguousarray(pos_mm[:, 1], dtype="<f8").tobytes())


def read_v1_allpos(path):
    with open(path, "rb") as f:
        nblock, blockSize, posDim = struct.unpack("<3I", f.read(12))
        if posDim > 2:
            raise ValueError("3-D positions are not supported (src/patch.cu:2762)")
        struct.unpack("<4d", f.read(32))
        n = nblock * blockSize
        x = np.fromfile(f, "<f8", n)
        y = np.fromfile(f, "<f8", n)
    return nblock, blockSize, np.stack([x, y], 1)


def write_connectome_cfg(path, typeAcc, nTypeI):
    with open(path, "wb") as f:
        f.write(struct.pack("<2I", len(typeAcc), nTypeI))
        f.write(np.asarray(typeAcc, "<u4").tobytes())


def read_connectome_cfg(path):
    with open(path, "rb") as f:
        nType, nTypeI = struct.unpack("<2I", f.read(8))
        typeAcc = _u32(f, nType) if nType > 1 else np.array([_u32(f, 1)], np.uint32)
    return int(nType), int(nTypeI), np.atleast_1d(typeAcc).astype(np.uint32)


# ------------------------------------------------------------------ connectome
def write_neighbor_block(path, nNear, nbr_lists):
    """nbr_lists[b] = all neighbour ids of block b, the first nNear[b] near."""
    with open(path, "wb") as f:
        f.write(np.asarray(nNear, "<u4").tobytes())
        f.write(np.asarray([len(l) for l in nbr_lists], "<u4").tobytes())
        for l in nbr_lists:
            f.write(np.asarray(l, "<u4").tobytes())


def read_neighbor_block(path, nblock):
    with open(path, "rb") as f:
        nNear = np.fromfile(f, "<u4", nblock)
        nAll = np.fromfile(f, "<u4", nblock)
        lists = [np.fromfile(f, "<u4", int(k)) for k in nAll]
    return nNear.astype(np.uint32), lists


def write_conmat(path, net):
    """Dense V1_conMat [nblock][nearNeighborBlock][bs][bs] of a Network, one
    block at a time."""
    sp = net.spec
    bs = sp.blockSize
    with open(path, "wb") as f:
        f.write(struct.pack("<I", net.maxNear))
        for b in range(sp.nblock):
            M = np.zeros((net.maxNear, bs, bs), "<f4")
            for ib in range(net.maxNear):
                pre, post, w = net.rec_pairs[b * net.maxNear + ib]
                M[ib, pre, post] = w
            f.write(M.tobytes())


def read_conmat_pairs(path, nblock, blockSize):
    """Stream V1_conMat block by block into per-(block, neighbour slot) sorted
    (pre, post, w) triples of the strictly positive entries (the entries the
    reference's recal_G uses, src/coredynamics.cu:2758)."""
    bs = blockSize
    pairs = []
    with open(path, "rb") as f:
        maxNear = _u32(f, 1)
        for b in range(nblock):
            M = np.fromfile(f, "<f4", maxNear * bs * bs)
            if M.size != maxNear * bs * bs:
                raise ValueError(f"{path}: truncated at block {b}")
            M = M.reshape(maxNear, bs, bs)
            for ib in range(maxNear):
                pre, post = np.nonzero(M[ib] > 0)
                pairs.append((pre.astype(np.uint32), post.astype(np.uint32),
                              M[ib, pre, post].astype(np.float32)))
    return int(maxNear), pairs


def pair_distances(net, b, ib, pre, post):
    """Distances (mm, float32) of pre-post pairs of block b's neighbour slot
    ib: x = (float)(x_pre - x_post) from the double positions, d = sqrtf(x*x +
    y*y) (src/preprocess/connect.cu:424-428)."""
    bs = net.spec.blockSize
    pos = np.asarray(net.pos, dtype=np.float64)
    gpre = int(net.nbr[b, ib]) * bs + np.asarray(pre, np.int64)
    gpost = b * bs + np.asarray(post, np.int64)
    x = (pos[gpre, 0] - pos[gpost, 0]).astype(np.float32)
    y = (pos[gpre, 1] - pos[gpost, 1]).astype(np.float32)
    return np.sqrt(x * x + y * y)


def write_delaymat(path, net):
    """Dense V1_delayMat [nblock][nearNeighborBlock][bs][bs] of a Network:
    the distance of every pair of a block and its near neighbours (the
    reference records it whether or not the pair is connected), zeros for
    neighbour slots past nNear."""
    sp = net.spec
    bs = sp.blockSize
    grid = np.arange(bs)
    pre, post = np.repeat(grid, bs), np.tile(grid, bs)
    with open(path, "wb") as f:
        f.write(struct.pack("<I", net.maxNear))
        for b in range(sp.nblock):
            M = np.zeros((net.maxNear, bs * bs), "<f4")
            for ib in range(int(net.nNear[b])):
                M[ib] = pair_distances(net, b, ib, pre, post)
            f.write(M.tobytes())


def read_delaymat_pairs(path, nblock, blockSize, pairs, extra=None):
    """Stream V1_delayMat block by block and keep the distances of the
    connected pairs, aligned with read_conmat_pairs' (pre, post, w) lists.
    extra: more (pre, post, ...) lists per (b, ib) in block-local indices
    (e.g. the gap junctions, offset by nE); their distances are returned
    as a second list."""
    bs = blockSize
    out = []
    out2 = []
    with open(path, "rb") as f:
        maxNear = _u32(f, 1)
        if len(pairs) != nblock * maxNear:
            raise ValueError(f"{path}: {maxNear} neighbour slots, V1_conMat has "
                             f"{len(pairs) // max(nblock, 1)}")
        for b in range(nblock):
            M = np.fromfile(f, "<f4", maxNear * bs * bs)
            if M.size != maxNear * bs * bs:
                raise ValueError(f"{path}: truncated at block {b}")
            M = M.reshape(maxNear, bs, bs)
            for ib in range(maxNear):
                pre, post, _ = pairs[b * maxNear + ib]
                out.append(M[ib, pre, post].astype(np.float32))
                if extra is not None:
                    pre, post = extra[b * maxNear + ib][:2]
                    out2.append(M[ib, pre, post].astype(np.float32))
    return (out, out2) if extra is not None else out


# -------------------------------------------------------------- gap junctions
def write_gapmat(path, net):
    """Dense V1_gapMat [nblock][nearNeighborBlock][nI pre][nI post] of a
    Network (header u32 nearNeighborBlock; src/patch.cu:3462-3473,
    3688-3690): the strengths as they are (the reader scales by gapRatio)."""
    sp = net.spec
    nI = sp.nI
    with open(path, "wb") as f:
        f.write(struct.pack("<I", net.maxNear))
        for b in range(sp.nblock):
            M = np.zeros((net.maxNear, nI, nI), "<f4")
            if net.gap is not None:
                for ib in range(net.maxNear):
                    pre, post, g = net.gap[b * net.maxNear + ib]
                    M[ib, pre, post] = g
            f.write(M.tobytes())


def read_gapmat_pairs(path, nblock, nI, maxNear, nNear):
    """Stream V1_gapMat block by block: per (b, ib) the (pre, post, g) of its
    nonzero entries (inhibitory indices, sorted by (pre, post)); the slots
    past a block's nNear are dropped, as the reference never reads them.
    Strengths must be >= 0 (src/patch.cu:3711)."""
    out = []
    z = np.zeros(0, np.uint32)
    with open(path, "rb") as f:
        mn = _u32(f, 1)
        if mn != maxNear:
            raise ValueError(f"{path}: {mn} neighbour slots, V1_conMat has {maxNear}")
        for b in range(nblock):
            M = np.fromfile(f, "<f4", maxNear * nI * nI)
            if M.size != maxNear * nI * nI:
                raise ValueError(f"{path}: truncated at block {b}")
            M = M.reshape(maxNear, nI, nI)
            if (M < 0).any():
                raise ValueError(f"{path}: negative gap-junction strength in block {b}")
            for ib in range(maxNear):
                if ib >= int(nNear[b]):
                    out.append((z, z, np.zeros(0, np.float32)))
                    continue
                pre, post = np.nonzero(M[ib])          # row-major: sorted by (pre, post)
                out.append((pre.astype(np.uint32), post.astype(np.uint32),
                            M[ib, pre, post].astype(np.float32)))
    return out


def write_gapvec(path, mI, gap_far=None):
    """V1_gapVec: u32 nGapVec[mI], then for every inhibitory neuron with
    junctions u32 id[n], f32 strength[n], f32 distance[n] (src/patch.cu
    :3362-3405).  gap_far: (off [mI+1], pre, g, dist) or None (no junctions)."""
    with open(path, "wb") as f:
        if gap_far is None:
            f.write(np.zeros(mI, "<u4").tobytes())
            return
        off, pre, g, dist = gap_far
        cnt = np.diff(off.astype(np.int64))
        f.write(cnt.astype("<u4").tobytes())
        for i in np.nonzero(cnt)[0]:
            a, b = int(off[i]), int(off[i + 1])
            f.write(np.asarray(pre[a:b], "<u4").tobytes())
            f.write(np.asarray(g[a:b], "<f4").tobytes())
            f.write(np.asarray(dist[a:b], "<f4").tobytes())


def read_gapvec(path, mI):
    """(off u32 [mI+1], pre u32, g f32, dist f32) of a V1_gapVec file, the
    strengths as they are."""
    with open(path, "rb") as f:
        cnt = np.fromfile(f, "<u4", mI).astype(np.int64)
        if cnt.size != mI:
            raise ValueError(f"{path}: truncated count table")
        off = np.zeros(mI + 1, np.int64)
        np.cumsum(cnt, out=off[1:])
        pre = np.zeros(int(off[-1]), np.uint32)
        g = np.zeros(int(off[-1]), np.float32)
        dist = np.zeros(int(off[-1]), np.float32)
        for i in np.nonzero(cnt)[0]:
            a, b, n = int(off[i]), int(off[i + 1]), int(cnt[i])
            pre[a:b] = np.fromfile(f, "<u4", n)
            g[a:b] = np.fromfile(f, "<f4", n)
            dist[a:b] = np.fromfile(f, "<f4", n)
    return off.astype(np.uint32), pre, g, dist


# ------------------------------------------------------------------ LGN -> V1
def write_v1_vec(path, nV1, far=None, connect_long_range=0):
    """V1_vec: int connectLongRange, nVec[nV1] (u32), a second u32[nV1]
    when connectLongRange is 1 (skipped by the reader), then for every post
    neuron with nVec > 0: ids (u32[nVec]), strengths (f32[nVec]) and
    distances in mm (f32[nVec]).  `far` = (off, pre, w, dist) as
    Network.far holds them; w is written as given (sRatioV1 applies on
    reading)."""
    if far is None:
        far = (np.zeros(nV1 + 1, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.float32),
               np.zeros(0, np.float32))
    off, pre, w, dist = far
    off = np.asarray(off, np.int64)
    cnt = np.diff(off).astype("<u4")
    with open(path, "wb") as f:
        f.write(struct.pack("<i", int(connect_long_range)))
        cnt.tofile(f)
        if connect_long_range == 1:
            np.zeros(nV1, "<u4").tofile(f)
        for n in np.nonzero(cnt)[0]:
            a, b = off[n], off[n + 1]
            np.asarray(pre[a:b], "<u4").tofile(f)
            np.asarray(w[a:b], "<f4").tofile(f)
            np.asarray(dist[a:b], "<f4").tofile(f)


def read_v1_vec(path, nV1, blockSize, nE, sRatioV1=None):
    """(off u32 [nV1+1], pre u32, w f32, dist f32) from a V1_vec file,
    strengths times sRatioV1[preType*2 + postType] in float32 as the reference
    reads them (src/patch.cu:3330-3336; one value applies to all four)."""
    sr = np.ones(4, np.float32)
    if sRatioV1 is not None:
        v = [float(x) for x in sRatioV1]
        sr = np.asarray(v * 4 if len(v) == 1 else v, np.float32)
        if sr.size != 4:
            raise ValueError(f"sRatioV1 has {len(v)} values, expected 1 or 4")
    with open(path, "rb") as f:
        hdr = f.read(4)
        if len(hdr) != 4:
            raise ValueError(f"{path}: truncated")
        clr = struct.unpack("<i", hdr)[0]
        cnt = _u32(f, nV1) if nV1 > 1 else np.atleast_1d(_u32(f, nV1))
        cnt = np.asarray(cnt, np.int64)
        if clr == 1:
            _u32(f, nV1)
        off = np.zeros(nV1 + 1, np.uint32)
        off[1:] = np.cumsum(cnt)
        n = int(off[-1])
        pre = np.zeros(n, np.uint32)
        w = np.zeros(n, np.float32)
        dist = np.zeros(n, np.float32)
        for i in np.nonzero(cnt)[0]:
            a, b, c = int(off[i]), int(off[i + 1]), int(cnt[i])
            pre[a:b] = _u32(f, c) if c > 1 else [_u32(f, 1)]
            w[a:b] = _f32(f, c)
            dist[a:b] = _f32(f, c)
    if n and int(pre.max()) >= nV1:
        raise ValueError(f"{path}: presynaptic id out of range")
    post = np.repeat(np.arange(nV1), cnt)
    k = (pre % blockSize >= nE).astype(np.int64) * 2 + (post % blockSize >= nE).astype(np.int64)
    w = (w * sr[k]).astype(np.float32)
    return off, pre, w, dist


def write_list_of_lists(path, lists, dtype):
    with open(path, "wb") as f:
        f.write(struct.pack("<I", len(lists)))
        for l in lists:
            a = np.asarray(l, dtype)
            f.write(struct.pack("<I", a.size))
            f.write(a.tobytes())


def read_list_of_lists(path, dtype):
    out = []
    with open(path, "rb") as f:
        n = _u32(f, 1)
        for _ in range(n):
            k = _u32(f, 1)
            out.append(np.fromfile(f, dtype, k) if k else np.zeros(0, dtype))
    return out


def write_lgn_v1(id_path, s_path, ids_lists, s_lists):
    write_list_of_lists(id_path, ids_lists, "<u4")
    maxList = max((len(l) for l in s_lists), default=0)
    with open(s_path, "wb") as f:
        f.write(struct.pack("<2I", len(s_lists), maxList))
        for l in s_lists:
            a = np.asarray(l, "<f4")
            f.write(struct.pack("<I", a.size))
            f.write(a.tobytes())


def read_lgn_v1(id_path, s_path, sRatioLGN, typeAcc):
    """-> (nLGN_of [nV1], sLGN [maxList][nV1], idLGN [maxList][nV1]): the
    transposed layout of read_LGN (src/patch.h:30-77), strengths scaled by
    sRatioLGN of the neuron's type."""
    ids = read_list_of_lists(id_path, "<u4")
    with open(s_path, "rb") as f:
        nList, maxList = struct.unpack("<2I", f.read(8))
        s_lists = []
        for _ in range(nList):
            k = _u32(f, 1)
            s_lists.append(np.fromfile(f, "<f4", k) if k else np.zeros(0, "<f4"))
    if len(ids) != nList:
        raise ValueError("LGN_V1_idList and LGN_V1_sList list counts differ")
    bs = int(typeAcc[-1])
    nV1 = nList
    mx = max(1, maxList)
    sLGN = np.zeros((mx, nV1), np.float32)
    idLGN = np.zeros((mx, nV1), np.uint32)
    nof = np.zeros(nV1, np.uint32)
    for i in range(nV1):
        k = len(s_lists[i])
        if len(ids[i]) != k:
            raise ValueError(f"neuron {i}: id and strength lists differ in length")
        t = int(np.searchsorted(typeAcc, i % bs, side="right"))
        sLGN[:k, i] = s_lists[i] * np.float32(sRatioLGN[t])
        idLGN[:k, i] = ids[i]
        nof[i] = k
    return nof, sLGN, idLGN


# ------------------------------------------------------------------ fLGN
_FLGN_2N = ("polar", "ecc", "rw", "rh", "orient", "k", "ratio", "tauR", "tauD", "nR", "nD", "delay")


def write_flgn(path, lgn: dict, lgn_type=None):
    nL = lgn["polar"].shape[1]
    with open(path, "wb") as f:
        f.write(struct.pack("<I", nL))
        t = np.zeros(nL, "<u4") if lgn_type is None else np.asarray(lgn_type, "<u4")
        f.write(t.tobytes())
        for k in _FLGN_2N:
            f.write(np.ascontiguousarray(lgn[k], "<f4").tobytes())
        for k in ("spont", "c50", "sharpness"):
            f.write(np.ascontiguousarray(lgn[k], "<f4").tobytes())
        f.write(np.ascontiguousarray(lgn["coneType"], "<u4").tobytes())
        f.write(np.ascontiguousarray(lgn["covariant"], "<f4").tobytes())


def read_flgn(path):
    with open(path, "rb") as f:
        nL = _u32(f, 1)
        out = {"type": np.fromfile(f, "<u4", nL)}
        for k in _FLGN_2N:
            out[k] = _f32(f, 2 * nL).reshape(2, nL)
        for k in ("spont", "c50", "sharpness"):
            out[k] = _f32(f, nL)
        out["coneType"] = np.fromfile(f, "<u4", 2 * nL).reshape(2, nL)
        out["covariant"] = _f32(f, nL)
    return out


# ------------------------------------------------------------------ outputs
class RawDataWriter:
    """rawData<suffix>.bin: header (dt f4, nt u4, nV1 u4, iModel i4, mI u4,
    hWrite u4, ngFF u4, ngE u4, ngI u4), then per step
    sInfo[nV1], dep[nV1], w[nV1] (iModel 1), v[nV1], gFF[ngFF][nV1],
    hFF[ngFF][nV1] (hWrite), gE[ngE][nV1], gI[ngI][nV1], hE, hI (hWrite),
    gap[mI]."""

    def __init__(self, path, dt, nt, nV1, iModel, mI, hWrite, ngFF, ngE, ngI):
        self.f = open(path, "wb")
        self.f.write(struct.pack("<fIIiIIIII", dt, nt, nV1, iModel, mI, int(hWrite), ngFF, ngE, ngI))
        self.nV1, self.iModel, self.mI, self.hWrite = nV1, iModel, mI, hWrite

    def write_step(self, sInfo, dep, w, v, gFF, hFF, gE, gI, hE, hI, gap=None):
        parts = [sInfo, dep]
        if self.iModel == 1:
            parts.append(w)
        parts += [v, gFF]
        if self.hWrite:
            parts.append(hFF)
        parts += [gE, gI]
        if self.hWrite:
            parts += [hE, hI]
        if self.mI:
            parts.append(np.zeros(self.mI, np.float32) if gap is None else gap)
        for a in parts:
            self.f.write(np.ascontiguousarray(a, "<f4").tobytes())

    def close(self):
        self.f.close()


def read_raw_data(path):
    """-> dict(header..., sInfo [nt, nV1], and the per-step traces)."""
    with open(path, "rb") as f:
        dt, nt, nV1, iModel, mI, hW, ngFF, ngE, ngI = struct.unpack("<fIIiIIIII", f.read(36))
        per = nV1 * (1 + (3 if iModel == 1 else 2) + (ngE + ngI + ngFF) * (1 + hW)) + mI
        data = np.fromfile(f, "<f4", per * nt)
    if data.size != per * nt:
        raise ValueError(f"{path}: truncated")
    data = data.reshape(nt, per)
    out = dict(dt=dt, nt=nt, nV1=nV1, iModel=iModel, mI=mI, hWrite=hW, ngFF=ngFF, ngE=ngE, ngI=ngI)
    o = 0

    def take(k):
        nonlocal o
        a = data[:, o:o + k * nV1].reshape(nt, k, nV1) if k else None
        o += k * nV1
        return a
    out["sInfo"] = take(1)[:, 0]
    out["dep"] = take(1)[:, 0]
    out["w"] = take(1)[:, 0] if iModel == 1 else None
    out["v"] = take(1)[:, 0]
    out["gFF"] = take(ngFF)
    out["hFF"] = take(ngFF) if hW else None
    out["gE"] = take(ngE)
    out["gI"] = take(ngI)
    out["hE"] = take(ngE) if hW else None
    out["hI"] = take(ngI) if hW else None
    out["gap"] = data[:, o:o + mI]
    return out


def spikes_from_sinfo(sInfo: np.ndarray, dt: float):
    """Spike times per neuron from a [nt, nV1] sInfo array, the decoding of
    readSpike (src/readPatchOutput.py:314-340): sInfo >= 1 is nsp + tsp/dt;
    several spikes in one step are spread around the mean time."""
    nt, nV1 = sInfo.shape
    out = [[] for _ in range(nV1)]
    its, js = np.nonzero(sInfo >= 1)
    for it, j in zip(its, js):
        s = float(sInfo[it, j])
        nsp = int(np.floor(s))
        tsp = s - nsp
        if nsp > 1:
            dtsp = tsp / nsp if 1 - tsp > 0.5 else (1 - tsp) / nsp
            t0 = tsp - (nsp // 2) * dtsp
            out[j].extend(((it + t0 + k * dtsp) * dt) for k in range(nsp))
        else:
            out[j].append((it + tsp) * dt)
    return [np.asarray(x) for x in out]


class SLGNWriter:
    """sLGN<suffix>.bin, the strengths of a plastic run (learnData_FF = 1,
    src/patch.cu:4767-4777, 6196-6200, 6936-6938): header nt u4,
    sampleInterval_LGN_V1 u4, dt f4, nV1 u4, max_LGNperV1 u4, sRatioLGN[2]
    f4, nLearnTypeFF_E u4, nLearnTypeFF_I u4, gmaxLGN[nLearnTypeFF] f4 (as
    given in the cfg), FF_InfRatio f4; then sLGN [max_LGNperV1][nV1] f4 after
    every sampleInterval steps (steps S, 2S, ... <= nt; the initial strengths
    are not written)."""

    def __init__(self, path, sp, maxLGN: int):
        self.f = open(path, "wb")
        self.f.write(struct.pack("<IIfII", sp.nt, sp.sampleInterval_LGN_V1, sp.dt, sp.nV1, maxLGN))
        self.f.write(struct.pack("<2f", *sp.sRatioLGN[:2]))
        self.f.write(struct.pack("<II", 1, 0))               # learning 3: FF_E only
        self.f.write(struct.pack("<2f", sp.gmaxLGN, sp.FF_InfRatio))

    def write(self, sLGN: np.ndarray):
        self.f.write(np.ascontiguousarray(sLGN, "<f4").tobytes())

    def close(self):
        self.f.close()


def read_slgn(path) -> dict:
    """The header and the records [n][max_LGNperV1][nV1] of an sLGN file."""
    b = Path(path).read_bytes()
    nt, S = struct.unpack_from("<II", b, 0)
    dt, = struct.unpack_from("<f", b, 8)
    nV1, maxLGN = struct.unpack_from("<II", b, 12)
    sR = struct.unpack_from("<2f", b, 20)
    nE, nI = struct.unpack_from("<II", b, 28)
    o = 36
    gmax = struct.unpack_from(f"<{nE + nI}f", b, o)
    o += 4 * (nE + nI)
    inf, = struct.unpack_from("<f", b, o)
    o += 4
    rec = np.frombuffer(b, "<f4", offset=o).reshape(-1, maxLGN, nV1)
    return dict(nt=nt, sampleInterval=S, dt=dt, nV1=nV1, maxLGN=maxLGN, sRatioLGN=sR,
                nLearnTypeFF_E=nE, nLearnTypeFF_I=nI, gmaxLGN=gmax, FF_InfRatio=inf, sLGN=rec)


class StepWriter:
    """LGN_fr (header nt, nLGN) or LGN_sp (no header) per-step float rows."""

    def __init__(self, path, nt=None, n=None):
        self.f = open(path, "wb")
        if nt is not None:
            self.f.write(struct.pack("<2I", nt, n))

    def write(self, rows: np.ndarray):
        self.f.write(np.ascontiguousarray(rows, "<f4").tobytes())

    def close(self):
        self.f.close()


# ------------------------------------------------------------ LGN positions
def write_lgn_vpos(path, nLGN_I, nLGN_C, max_ecc, x0, xspan, y0, yspan, x, y, types, polar,
                   ecc_deg):
    """LGN_vpos: u32 nLGN_I, nLGN_C; f32 max_ecc; f32 x0, xspan, y0, yspan;
    f32 x[n], y[n]; u32 type[n]; f32 polar[n] (rad), ecc[n] (deg)."""
    n = nLGN_I + nLGN_C
    with open(path, "wb") as f:
        f.write(struct.pack("<2I", nLGN_I, nLGN_C))
        f.write(struct.pack("<5f", max_ecc, x0, xspan, y0, yspan))
        for a, dt in ((x, "<f4"), (y, "<f4"), (types, "<u4"), (polar, "<f4"), (ecc_deg, "<f4")):
            a = np.ascontiguousarray(a, dtype=dt)
            if a.size != n:
                raise ValueError("LGN_vpos arrays must have nLGN_I + nLGN_C entries")
            f.write(a.tobytes())


def read_lgn_vpos(path) -> dict:
    with open(path, "rb") as f:
        nI, nC = _u32(f), _u32(f)
        max_ecc, x0, xspan, y0, yspan = (float(v) for v in _f32(f, 5))
        n = nI + nC
        x, y = _f32(f, n), _f32(f, n)
        types = np.fromfile(f, dtype="<u4", count=n)
        polar, ecc = _f32(f, n), _f32(f, n)
        if types.size != n or ecc.size != n:
            raise ValueError(f"{path}: truncated")
    return dict(nLGN_I=nI, nLGN_C=nC, nLGN=n, max_ecc=max_ecc, x0=x0, xspan=xspan, y0=y0,
                yspan=yspan, x=x, y=y, type=types, polar=polar, ecc_deg=ecc)


def write_lgn_surface_id(path, sx, sy):
    """LGN_surfaceID: u32 max_xid, max_yid; i32 x[nLGN], y[nLGN]."""
    sx = np.ascontiguousarray(sx, dtype="<i4")
    sy = np.ascontiguousarray(sy, dtype="<i4")
    with open(path, "wb") as f:
        f.write(struct.pack("<2I", int(sx.max()), int(sy.max())))
        f.write(sx.tobytes())
        f.write(sy.tobytes())


def read_lgn_surface_id(path, nLGN: int):
    """(nsx, nsy, x[nLGN], y[nLGN]): the surface is (max_xid + 1) x (max_yid +
    1) (src/patch.cu:2582-2594); every id must lie on it, and no two cells may
    share a position (each cell's spike record has its own texel)."""
    with open(path, "rb") as f:
        nsx, nsy = _u32(f) + 1, _u32(f) + 1
        xy = np.fromfile(f, dtype="<i4", count=2 * nLGN)
    if xy.size != 2 * nLGN:
        raise ValueError(f"{path}: truncated")
    x, y = xy[:nLGN], xy[nLGN:]
    if nLGN and (x.min() < 0 or y.min() < 0 or x.max() >= nsx or y.max() >= nsy):
        raise ValueError(f"{path}: a surface id lies outside the {nsx} x {nsy} surface")
    if np.unique(y.astype(np.int64) * nsx + x).size != nLGN:
        raise ValueError(f"{path}: two LGN cells share a surface position")
    return nsx, nsy, x, y


# ------------------------------------------------------- connection statistics
def write_constats(path, nType: int, ffRatio):
    """conStats: u32 nType, nV1; f32 ffRatio[nV1] (the part the simulator
    reads; genCon writes further statistics after it)."""
    ff = np.ascontiguousarray(ffRatio, dtype="<f4")
    with open(path, "wb") as f:
        f.write(struct.pack("<2I", nType, ff.size))
        f.write(ff.tobytes())


def read_constats(path, nType: int, nV1: int) -> np.ndarray:
    with open(path, "rb") as f:
        nt, n = _u32(f), _u32(f)
        if nt != nType or n != nV1:
            raise ValueError(f"{path}: nType {nt}, nV1 {n}; expected {nType}, {nV1}")
        ff = _f32(f, n)
    if np.isnan(ff).any():
        raise ValueError(f"{path}: ffRatio holds NaN")
    return ff


def write_v1_feature(path, features):
    """V1_feature: u32 nFeature; f32 [nFeature][nV1] (feature 0 ocular
    dominance, feature 1 orientation preference in [0, 1))."""
    a = np.ascontiguousarray(features, dtype="<f4")
    with open(path, "wb") as f:
        f.write(struct.pack("<I", a.shape[0]))
        f.write(a.tobytes())


def read_v1_feature(path, nV1: int) -> np.ndarray:
    with open(path, "rb") as f:
        nF = _u32(f)
        a = _f32(f, nF * nV1)
    return a.reshape(nF, nV1)


# ------------------------------------------------------------- patchV1_cfg
def write_patchv1_cfg(path, sp, nLGN: int, stimulus_filename: str, view: tuple,
                      sRatioV1=(1.0, 1.0, 1.0, 1.0), noisyDep=(0.0, 0.0), virtual_LGN=False,
                      learning: int = 0, synFail=(0.0, 0.0, 0.0, 0.0), synFailFF=(0.0, 0.0)):
    """The run's parameters in the reference's patchV1_cfg layout
    (src/patch.cu:6953-7028), Float = f32: two neuron types (E, I), one
    level of each in the type hierarchy.  view = (normViewDistance, L_x0,
    L_y0, R_x0, R_y0).  The initial-condition distributions v0/w0/gFF0/gE0/
    gI0 (the reference's [mean, std] per type) and spE0/spI0 are written as
    zeros: this simulator draws its initial state itself."""
    f4 = lambda v: np.asarray(v, dtype="<f4").tobytes()
    u4 = lambda v: np.asarray(v, dtype="<u4").tobytes()
    ngFF, ngE, ngI = len(sp.riseTimeFF), len(sp.riseTimeE), len(sp.riseTimeI)
    fn = stimulus_filename.encode()
    flat = lambda rows: [x for r in rows for x in r]
    parts = [
        u4([4]), f4([-65.0, 0.0, -75.0]), u4([2, 1, 1]),
        f4(sp.vR), f4(sp.vThres), f4(sp.gL), f4(sp.vT), u4([1, 1]),
        u4([sp.nE, sp.blockSize]), f4(sp.sRatioLGN), f4(sRatioV1),
        f4([sp.frRatioLGN, sp.convolRatio]), u4([sp.frameRate]), u4([len(fn)]), fn,
        u4([nLGN, sp.nV1, sp.nt]), f4([sp.dt]), f4(view),
        f4(sp.tonicDep), f4(noisyDep), np.asarray([int(virtual_LGN)], "<i4").tobytes(),
        np.asarray([sp.seed], "<u8").tobytes(), np.asarray([sp.iModel, learning], "<i4").tobytes(),
        u4([ngFF, ngE, ngI]),
        f4(sp.riseTimeFF), f4(sp.decayTimeFF), f4(sp.riseTimeE), f4(sp.decayTimeE),
        f4(sp.riseTimeI), f4(sp.decayTimeI),
        f4(flat(sp.pFF)), f4(flat(sp.pE)), f4(flat(sp.pI)),
        f4(sp.C), f4(sp.tRef), f4(sp.a), f4(sp.b), f4(sp.tau_w), f4(sp.deltaT),
        f4([0.0, 0.0]), f4([0.0, 0.0]),                       # spE0, spI0
        f4([0.0] * 4) * 5,                                    # v0, w0, gFF0, gE0, gI0
        f4(synFail), f4(synFailFF),
    ]
    with open(path, "wb") as f:
        for b in parts:
            f.write(b)
