"""Snapshots of a running simulation in the reference's snapshot layout, and
resuming from them (src/patch.cu:6793-6912 writes, :5489-5684 reads).

File layout, in order (little endian; u32 unless noted):

    it0 (the next step), frameCycle, iFrameHead, oldFrameHead, currentFrame,
    iFramePhaseHead, iFramePhaseTail, mFramePhaseHead, mFramePhaseTail,
    iStatus, currentTimeSlot
    iModel (i32), learning (i32)
    nV1, nLGN, ngTypeFF, ngTypeE, ngTypeI, nType, trainDepth
    LGN  leftTimeRate f32[nLGN], lastNegLogRand f32[nLGN], generator[nLGN]
    plastic runs (learning != 0; src/patch.cu:6845-6851, 5556-5563):
         sLGN f32[maxLGN][nV1]; the post variables f32 vLTD_FF_E[nblock][nE],
         vTrip_FF_E[nblock][nE], vAvgE[nblock][nE][2], vAvgI[nblock][nI]
         (learning 3: no FF_I rule); the LGN trace f32[nLGN]
    V1   rGenCond generator[nV1], rNoisy generator[nV1], tBack f32[nV1]
    spikeTrain f32[trainDepth][nV1]
    state block f32: depC[nV1], (w[nV1] if iModel == 1), v[nV1],
         gFF[ngFF][nV1], hFF[ngFF][nV1], gE[ngE][nV1], gI[ngI][nV1],
         hE[ngE][nV1], hI[ngI][nV1], gap[nblock*nI]

A generator slot is 48 bytes, the size of the reference's MRG32k3a state.
The random numbers here come from counter-based Philox streams (the stream
id, the entity and the step are the counter), so a slot holds that stream's
identity instead of a state: u32 'PHLX' tag, u32 entity, u32 key low, u32 key
high, u32 stream id (0xffffffff: none), zeros -- except that the rNoisy slot
of a neuron (the voltage-noise stream) holds in its sixth word the float32
bits of the neuron's Ornstein-Uhlenbeck noise value, the only state that
stream carries from step to step (the reference keeps it in last_noise and
does not write it).  The frame-ring fields are
this simulator's: currentFrame frames have entered, iFrameHead =
currentFrame mod maxFrame, the parvo phase fields and iStatus are 0.  With no
axonal delays trainDepth is 1 and the spike train is the last step's output;
with delays (or delayed far connections) it is the whole ring and
currentTimeSlot the slot of step it0
(a ring of another length is cut or padded keeping the newest steps, as
src/patch.cu:5580-5660 does).

A snapshot is taken at a frame-period boundary (the LGN stage runs a frame
period ahead of V1).  Before writing, the last step's pending recurrent
increments are added to the state, as the reference's end-of-step recal_G
has done when it writes; the continuing run is unchanged by this, and a run
restored from the file continues bit for bit.
"""
from __future__ import annotations

import ctypes as C
import struct

import numpy as np
import torch

from . import abi
from .abi import check, ptr

GEN_BYTES = 48
TAG = 0x584C4850                  # 'PHLX'
NO_STREAM = 0xFFFFFFFF
STREAM_LGN_SPIKE = 1              # the Philox stream ids of the kernels
STREAM_V1_BG = 3
STREAM_V1_NOISE = 7
HEADER_U32 = 11


def generator_slots(n: int, key_lo: int, key_hi: int, stream: int, carried=None) -> bytes:
    """n 48-byte generator slots naming Philox stream `stream` of entities
    0..n-1; `carried` (float32[n]) is the value the stream carries between
    steps, in the sixth word."""
    a = np.zeros((n, GEN_BYTES // 4), np.uint32)
    a[:, 0] = TAG
    a[:, 1] = np.arange(n, dtype=np.uint32)
    a[:, 2] = key_lo
    a[:, 3] = key_hi
    a[:, 4] = stream
    if carried is not None:
        a[:, 5] = np.ascontiguousarray(carried, dtype=np.float32).view(np.uint32)
    return a.tobytes()


def carried_values(slots: bytes, n: int) -> np.ndarray:
    """The float32 values the generator slots carry (zeros for slots that are
    not this format's, such as the reference's MRG32k3a states)."""
    a = np.frombuffer(slots, np.uint32, n * (GEN_BYTES // 4)).reshape(n, GEN_BYTES // 4)
    out = a[:, 5].copy().view(np.float32)
    out[a[:, 0] != TAG] = 0.0
    return out


def state_block_floats(nV1: int, iModel: int, ngFF: int, ngE: int, ngI: int, mI: int) -> int:
    return nV1 * (2 + (1 if iModel == 1 else 0) + 2 * ngFF + 2 * (ngE + ngI)) + mI


def learn_floats(nV1: int, nblock: int, nE: int, maxLGN: int, nLGN: int) -> tuple:
    """Float counts of the plastic sections (learning 3): sLGN, the post
    variables (learnVarFFsize, src/patch.cu:4366-4368) and the LGN trace
    (LGN_learnPreSize, :2620)."""
    return maxLGN * nV1, 2 * nE * nblock + nV1 + nE * nblock, nLGN


def file_size(nV1, nLGN, iModel, ngFF, ngE, ngI, mI, trainDepth=1, learn=None) -> int:
    """Bytes of a snapshot (the reference's section sizes, src/patch.cu:5479, 3902-3907);
    learn = (nblock, nE, maxLGN) for a plastic run."""
    extra = 4 * sum(learn_floats(nV1, learn[0], learn[1], learn[2], nLGN)) if learn else 0
    return (4 * HEADER_U32 + 4 * 2 + 4 * 7 + nLGN * (8 + GEN_BYTES) + nV1 * (2 * GEN_BYTES + 4)
            + 4 * trainDepth * nV1 + 4 * state_block_floats(nV1, iModel, ngFF, ngE, ngI, mI)
            + extra)


def post_to_file(lvar: dict, nblock: int, bs: int, nE: int) -> np.ndarray:
    """The post variables [nV1] (excitatory entries) in the reference's
    learnVarFF layout (vAvgE's second slot and vAvgI are 0 with learning 3)."""
    nV1 = nblock * bs
    eid = np.arange(nV1).reshape(nblock, bs)[:, :nE].reshape(-1)
    avg = np.zeros((nblock * nE, 2), np.float32)
    avg[:, 0] = lvar["vAvg"][eid]
    return np.concatenate([lvar["vLTD"][eid], lvar["vTrip"][eid], avg.reshape(-1),
                           np.zeros(nblock * (bs - nE), np.float32)]).astype(np.float32)


def post_from_file(post: np.ndarray, nblock: int, bs: int, nE: int) -> dict:
    """post_to_file's inverse: [nV1] arrays, inhibitory entries 0."""
    nV1, k = nblock * bs, nblock * nE
    eid = np.arange(nV1).reshape(nblock, bs)[:, :nE].reshape(-1)
    out = {}
    for name, a in (("vLTD", post[:k]), ("vTrip", post[k:2 * k]), ("vAvg", post[2 * k:4 * k:2])):
        v = np.zeros(nV1, np.float32)
        v[eid] = a
        out[name] = v
    return out


def write_file(path, h: dict, lgn: dict, v1: dict, spike_train: np.ndarray, block: dict,
               keys=(0, 0), learn: dict | None = None):
    """Write a snapshot from host arrays.  h: the header fields; lgn:
    leftTimeRate, lastNegLogRand; v1: tBack, optionally lastNoise; block:
    depC, w (iModel 1), v, gFF, hFF, gE, gI, hE, hI, gap; learn (h learning
    != 0): sLGN, post (the file layout), pre (the LGN trace)."""
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32).tobytes()
    nV1, nLGN = h["nV1"], h["nLGN"]
    with open(path, "wb") as f:
        f.write(struct.pack("<11I", h["it0"], h["frameCycle"], h["iFrameHead"], h["oldFrameHead"],
                            h["currentFrame"], 0, 0, 0, 0, h.get("iStatus", 0),
                            h.get("currentTimeSlot", 0)))
        f.write(struct.pack("<2i", h["iModel"], h.get("learning", 0)))
        f.write(struct.pack("<7I", nV1, nLGN, h["ngTypeFF"], h["ngTypeE"], h["ngTypeI"],
                            h["nType"], h.get("trainDepth", 1)))
        f.write(f32(lgn["leftTimeRate"]))
        f.write(f32(lgn["lastNegLogRand"]))
        f.write(generator_slots(nLGN, keys[0], keys[1], STREAM_LGN_SPIKE))
        if h.get("learning", 0):
            for k in ("sLGN", "post", "pre"):
                f.write(f32(learn[k]))
        f.write(generator_slots(nV1, keys[0], keys[1], STREAM_V1_BG))
        f.write(generator_slots(nV1, keys[0], keys[1], STREAM_V1_NOISE,
                                carried=v1.get("lastNoise", np.zeros(nV1, np.float32))))
        f.write(f32(v1["tBack"]))
        f.write(f32(spike_train))
        f.write(f32(block["depC"]))
        if h["iModel"] == 1:
            f.write(f32(block["w"]))
        for k in ("v", "gFF", "hFF", "gE", "gI", "hE", "hI", "gap"):
            f.write(f32(block[k]))


def read_file(path, learn=None) -> dict:
    """Read a snapshot written by write_file (or by the reference, whose
    generator slots are then MRG32k3a states, returned as raw bytes).  A
    plastic run's file needs learn = (nblock, nE, maxLGN) for the section
    sizes (the header does not hold them)."""
    with open(path, "rb") as f:
        data = f.read()
    o = 0

    def take(fmt):
        nonlocal o
        v = struct.unpack_from(fmt, data, o)
        o += struct.calcsize(fmt)
        return v

    def floats(n):
        nonlocal o
        a = np.frombuffer(data, np.float32, n, o).copy()
        o += 4 * n
        return a

    def raw(n):
        nonlocal o
        b = data[o:o + n]
        o += n
        return b

    names = ("it0", "frameCycle", "iFrameHead", "oldFrameHead", "currentFrame",
             "iFramePhaseHead", "iFramePhaseTail", "mFramePhaseHead", "mFramePhaseTail",
             "iStatus", "currentTimeSlot")
    h = dict(zip(names, take("<11I")))
    h["iModel"], h["learning"] = take("<2i")
    if h["learning"] and learn is None:
        raise ValueError("a snapshot of a plastic run: pass learn = (nblock, nE, maxLGN)")
    h.update(zip(("nV1", "nLGN", "ngTypeFF", "ngTypeE", "ngTypeI", "nType", "trainDepth"),
                 take("<7I")))
    nV1, nLGN = h["nV1"], h["nLGN"]
    out = {"header": h}
    out["leftTimeRate"] = floats(nLGN)
    out["lastNegLogRand"] = floats(nLGN)
    out["lgn_generators"] = raw(nLGN * GEN_BYTES)
    if h["learning"]:
        nS, nP, nQ = learn_floats(nV1, learn[0], learn[1], learn[2], nLGN)
        out["sLGN"] = floats(nS).reshape(learn[2], nV1)
        out["post"] = floats(nP)
        out["pre"] = floats(nQ)
    out["rGenCond"] = raw(nV1 * GEN_BYTES)
    out["rNoisy"] = raw(nV1 * GEN_BYTES)
    out["lastNoise"] = carried_values(out["rNoisy"], nV1)
    out["tBack"] = floats(nV1)
    out["spikeTrain"] = floats(h["trainDepth"] * nV1).reshape(h["trainDepth"], nV1)
    ngFF, ngE, ngI = h["ngTypeFF"], h["ngTypeE"], h["ngTypeI"]
    rest = (len(data) - o) // 4
    mI = rest - state_block_floats(nV1, h["iModel"], ngFF, ngE, ngI, 0)
    if mI < 0 or (len(data) - o) % 4:
        raise ValueError(f"{path}: truncated snapshot ({len(data)} bytes)")
    out["depC"] = floats(nV1)
    if h["iModel"] == 1:
        out["w"] = floats(nV1)
    out["v"] = floats(nV1)
    for k, ng in (("gFF", ngFF), ("hFF", ngFF), ("gE", ngE), ("gI", ngI), ("hE", ngE),
                  ("hI", ngI)):
        out[k] = floats(ng * nV1).reshape(ng, nV1)
    out["gap"] = floats(mI)
    assert o == len(data)
    return out


def save(sim, path) -> dict:
    """Snapshot of `sim` (a GpuSim at a frame-period boundary) to `path`;
    returns the header."""
    net, sp, tm = sim.net, sim.net.spec, sim.tm
    npf, maxF = tm["ntPerFrame"], tm["maxFrame"]
    if getattr(net, "batch", None) is not None:
        raise ValueError("snapshots of batched trials are not supported")
    if sim.it % npf != 0:
        raise ValueError(f"snapshot at step {sim.it}: snapshots are taken at frame-period "
                         f"boundaries ({npf} steps)")
    K = sim.it // npf
    sim.finalize()                          # pending increments into the state
    torch.cuda.synchronize()
    learn = getattr(sim, "learn", False)
    if sim.lgn_to <= sim.it:
        # nothing of period K issued yet: the live LGN state is its start
        lgn = torch.stack([sim.lTR, sim.lNL])
        xpre = sim.xVar if learn else None
    else:
        lgn = sim.lgn_at[K % 2]             # the LGN state at the start of period K
        xpre = sim.xStart2[K % 2][0] if learn else None   # the trace at its first step
    dep = torch.empty(sp.nV1, dtype=torch.float32, device=sim.dev)
    check(sim.lib.vc_v1_dep_current(None, C.byref(sim.p), C.byref(sim.s), ptr(dep)),
          "vc_v1_dep_current")
    sim.check()
    st = {k: v.cpu().numpy() for k, v in sim.state.items()}
    ngFF, ngE, ngI = int(sim.p.ngFF), int(sim.p.ngE), int(sim.p.ngI)
    nFr = K * tm["denorm"]                  # frames entered (denorm per period)
    h = dict(it0=sim.it, frameCycle=nFr // maxF, iFrameHead=nFr % maxF,
             oldFrameHead=(nFr - 1) % maxF, currentFrame=nFr, iModel=sp.iModel,
             learning=int(sp.learning) if learn else 0,
             nV1=sp.nV1, nLGN=sp.nLGN, ngTypeFF=ngFF, ngTypeE=ngE, ngTypeI=ngI,
             nType=int(sim.p.nType), trainDepth=1)
    if getattr(sim, "delay", False) or getattr(sim, "far_only", False):
        # the whole ring (axonal delays, or far connections); the next step's slot
        h.update(trainDepth=sim.depth, currentTimeSlot=(sim.it + sim.slot_off) % sim.depth)
        train = sim.train.cpu().numpy()
    else:
        train = sim.sInfo_last.cpu().numpy()[None]
    lg = lgn.cpu().numpy()
    block = dict(depC=dep.cpu().numpy(), w=st["w"], v=st["v"], gFF=st["gFF"][:ngFF],
                 hFF=st["hFF"][:ngFF], gE=st["gE"][:ngE], gI=st["gI"][:ngI],
                 hE=st["hE"][:ngE], hI=st["hI"][:ngI],
                 gap=sim.gap.cpu().numpy() if getattr(sim, "gap", None) is not None
                 else np.zeros(sp.nblock * sp.nI, np.float32))
    lrn = None
    if learn:
        lv = {k: v.cpu().numpy() for k, v in sim.lvar.items()}
        lrn = dict(sLGN=sim.sLGN.cpu().numpy(), post=post_to_file(lv, sp.nblock, sp.blockSize,
                                                                   sp.nE),
                   pre=xpre.cpu().numpy())
    write_file(path, h, dict(leftTimeRate=lg[0], lastNegLogRand=lg[1]),
               dict(tBack=st["tBack"], lastNoise=sim.lastNoise.cpu().numpy()), train, block,
               keys=(int(sim.p.seed_lo), int(sim.p.seed_hi)), learn=lrn)
    return h


def restore(sim, path, as_init: bool = False) -> dict:
    """Load the snapshot at `path` into `sim` (a GpuSim of the same network).
    as_init (the reference's asInit): take the state as the initial
    condition and start at step 0; otherwise resume at the snapshot's step."""
    net, sp = sim.net, sim.net.spec
    learn = getattr(sim, "learn", False)
    d = read_file(path, learn=(sp.nblock, sp.nE, int(sim.sLGN.shape[0])) if learn else None)
    h = d["header"]
    if bool(h["learning"]) != learn:
        raise ValueError(f"snapshot {path} has learning = {h['learning']}, this run "
                         f"{sp.learning}")
    ngFF, ngE, ngI = int(sim.p.ngFF), int(sim.p.ngE), int(sim.p.ngI)
    mine = (sp.iModel, sp.nV1, sp.nLGN, ngFF, ngE, ngI)
    theirs = (h["iModel"], h["nV1"], h["nLGN"], h["ngTypeFF"], h["ngTypeE"], h["ngTypeI"])
    if mine != theirs:
        raise ValueError(f"snapshot {path} is of another network: (iModel, nV1, nLGN, ngFF, "
                         f"ngE, ngI) = {theirs}, this one {mine}")
    dep = torch.empty(sp.nV1, dtype=torch.float32, device=sim.dev)
    check(sim.lib.vc_v1_dep_current(None, C.byref(sim.p), C.byref(sim.s), ptr(dep)),
          "vc_v1_dep_current")
    if not np.array_equal(dep.cpu().numpy().view(np.uint32), d["depC"].view(np.uint32)):
        raise ValueError(f"snapshot {path}: its depC differs from this network's tonic drive")
    sim.seek(0 if as_init else h["it0"])
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=sim.dev)
    st = sim.state
    st["v"].copy_(T(d["v"]))
    st["tBack"].copy_(T(d["tBack"]))
    if h["iModel"] == 1:
        st["w"].copy_(T(d["w"]))
    for k, ng in (("gFF", ngFF), ("hFF", ngFF), ("gE", ngE), ("gI", ngI), ("hE", ngE),
                  ("hI", ngI)):
        st[k][:ng].copy_(T(d[k]))
    if getattr(sim, "gap", None) is not None:       # the gap sums after the last step
        sim.gap.copy_(T(d["gap"]))
    sim.lastNoise.copy_(T(d["lastNoise"]))
    sim.lTR.copy_(T(d["leftTimeRate"]))
    sim.lNL.copy_(T(d["lastNegLogRand"]))
    if learn:
        # the strengths and the post variables (src/patch.cu:5556-5563); the
        # homeostatic totals stay those of the network's initial strengths,
        # which the reference computes before it reads the snapshot
        # (:4336-4349); the LGN trace is the file's (the reference's resumed
        # run restarts it from 0, its surface being cleared at start-up)
        sim.sLGN.copy_(T(d["sLGN"]))
        for k, v in post_from_file(d["post"], sp.nblock, sp.blockSize, sp.nE).items():
            sim.lvar[k].copy_(T(v))
        sim.xVar.copy_(T(d["pre"]))
        if sim.totalFF is not None:
            from .params import learn_ff_totals
            tot, inf = learn_ff_totals(net)
            sim.totalFF.copy_(T(tot))
            sim.totalFFinf.copy_(T(inf))
    ring = d["spikeTrain"]
    if getattr(sim, "delay", False) or getattr(sim, "far_only", False):
        it0 = 0 if as_init else h["it0"]
        Df, D = int(h["trainDepth"]), sim.depth
        cts = int(h.get("currentTimeSlot", 0)) % Df   # the file's slot of step it0
        out = np.zeros((D, sp.nV1), np.float32)
        if Df == D:
            sim.slot_off = (cts - it0) % D
            out[:] = ring
        else:                                  # another ring length: keep the newest steps
            sim.slot_off = 0
            for a in range(1, min(Df, D) + 1):
                out[(it0 - a) % D] = ring[(cts - a) % Df]
        bits = np.packbits((out >= 1.0).reshape(D, sp.nV1 // 64, 64), axis=-1,
                           bitorder="little")
        sim.train.copy_(T(out))
        sim.trainMask.copy_(T(bits.reshape(D, -1).view(np.int64)))
        last = out[(it0 + sim.slot_off - 1) % D]
    else:
        last = ring[-1]
    sim.sInfo_step.copy_(T(last))
    sim.sInfo_last = sim.sInfo_step
    torch.cuda.synchronize()
    return h
