"""Snapshot file layout (CPU): the sections, their order and sizes follow the
reference's writer (src/patch.cu:6793-6912) and reader (:5489-5684); a
written file reads back unchanged."""
import struct

import numpy as np
import pytest


def _reference_size(nV1, nLGN, iModel, ngFF, ngE, ngI, mI, trainDepth):
    # the reference's own section sizes: 11 PosInt + 2 int + 7 Size header,
    # LGN_convolSize = 2*nLGN*sizeof(Float) + nLGN*sizeof(curandStateMRG32k3a)
    # (src/patch.cu:5479), two MRG32k3a states and tBack per neuron, the
    # spike train, and totalSize = vSize + depSize + ghSize + ffSize + gapSize
    # (+ wSize for AdEx) (:3902-3907, :4001-4004)
    mrg = 48
    header = 11 * 4 + 2 * 4 + 7 * 4
    lgn_convol = 2 * nLGN * 4 + nLGN * mrg
    v1 = 2 * nV1 * mrg + nV1 * 4
    train = trainDepth * nV1 * 4
    total = nV1 * 4 + nV1 * 4 + 2 * nV1 * (ngE + ngI) * 4 + 2 * nV1 * ngFF * 4 + mI * 4
    if iModel == 1:
        total += nV1 * 4
    return header + lgn_convol + v1 + train + total


@pytest.mark.parametrize("iModel", [0, 1])
def test_round_trip_and_layout(pkg, tmp_path, iModel):
    from visual_cortex_amd import snapshot as S
    rng = np.random.default_rng(3 + iModel)
    nV1, nLGN, ngFF, ngE, ngI, mI = 256, 96, 2, 2, 1, 64
    f = lambda *s: rng.standard_normal(s).astype(np.float32)
    h = dict(it0=640, frameCycle=1, iFrameHead=3, oldFrameHead=2, currentFrame=10,
             iModel=iModel, nV1=nV1, nLGN=nLGN, ngTypeFF=ngFF, ngTypeE=ngE, ngTypeI=ngI,
             nType=2, trainDepth=1)
    lgn = dict(leftTimeRate=f(nLGN), lastNegLogRand=f(nLGN))
    block = dict(depC=f(nV1), w=f(nV1), v=f(nV1), gFF=f(ngFF, nV1), hFF=f(ngFF, nV1),
                 gE=f(ngE, nV1), gI=f(ngI, nV1), hE=f(ngE, nV1), hI=f(ngI, nV1), gap=f(mI))
    tBack, train = f(nV1), f(1, nV1)
    p = tmp_path / "snap.bin"
    S.write_file(p, h, lgn, dict(tBack=tBack), train, block, keys=(11, 22))
    raw = p.read_bytes()
    assert len(raw) == _reference_size(nV1, nLGN, iModel, ngFF, ngE, ngI, mI, 1)
    assert len(raw) == S.file_size(nV1, nLGN, iModel, ngFF, ngE, ngI, mI)
    # fixed offsets of the header fields
    assert struct.unpack_from("<I", raw, 0)[0] == 640
    assert struct.unpack_from("<I", raw, 16)[0] == 10             # currentFrame
    assert struct.unpack_from("<2i", raw, 44) == (iModel, 0)
    assert struct.unpack_from("<7I", raw, 52) == (nV1, nLGN, ngFF, ngE, ngI, 2, 1)
    # leftTimeRate right after the header, the state block at the end
    o = 80
    assert np.array_equal(np.frombuffer(raw, np.float32, nLGN, o), lgn["leftTimeRate"])
    end = np.frombuffer(raw[-4 * mI:], np.float32)
    assert np.array_equal(end, block["gap"])
    # generator slots name the Philox streams
    slots = np.frombuffer(raw, np.uint32, nLGN * 12, o + 8 * nLGN).reshape(nLGN, 12)
    assert (slots[:, 0] == S.TAG).all() and np.array_equal(slots[:, 1], np.arange(nLGN))
    assert (slots[:, 2] == 11).all() and (slots[:, 3] == 22).all() and (slots[:, 4] == 1).all()
    d = S.read_file(p)
    for k, v in h.items():
        assert d["header"][k] == v, k
    for k in ("leftTimeRate", "lastNegLogRand"):
        assert np.array_equal(d[k], lgn[k])
    assert np.array_equal(d["tBack"], tBack) and np.array_equal(d["spikeTrain"], train)
    for k, v in block.items():
        if k == "w" and iModel == 0:
            assert "w" not in d
            continue
        assert np.array_equal(d[k], v), k


def test_truncated_file_is_rejected(pkg, tmp_path):
    from visual_cortex_amd import snapshot as S
    nV1, nLGN = 64, 16
    z = lambda *s: np.zeros(s, np.float32)
    h = dict(it0=0, frameCycle=0, iFrameHead=0, oldFrameHead=0, currentFrame=0, iModel=0,
             nV1=nV1, nLGN=nLGN, ngTypeFF=1, ngTypeE=1, ngTypeI=1, nType=2, trainDepth=1)
    p = tmp_path / "snap.bin"
    S.write_file(p, h, dict(leftTimeRate=z(nLGN), lastNegLogRand=z(nLGN)), dict(tBack=z(nV1)),
                 z(1, nV1), dict(depC=z(nV1), v=z(nV1), gFF=z(1, nV1), hFF=z(1, nV1),
                                 gE=z(1, nV1), gI=z(1, nV1), hE=z(1, nV1), hI=z(1, nV1),
                                 gap=z(16)))
    raw = p.read_bytes()
    q = tmp_path / "short.bin"
    q.write_bytes(raw[:len(raw) - 4 * nV1 * 3])
    with pytest.raises(ValueError):
        S.read_file(q)


def test_plastic_sections_layout(pkg, tmp_path):
    """A plastic run's sections sit between the LGN generators and the V1
    ones (src/patch.cu:6845-6851): sLGN [maxLGN][nV1], the post variables in
    the reference's learnVarFF layout (vLTD_FF_E, vTrip_FF_E over
    [nblock][nE], vAvgE [nblock][nE][2], vAvgI [nblock][nI]; 2 nE nblock +
    nV1 + nE nblock floats), the LGN trace [nLGN]."""
    from visual_cortex_amd import snapshot as S
    rng = np.random.default_rng(9)
    nblock, bs, nE, maxLGN, nLGN, ngFF, ngE, ngI = 3, 128, 96, 5, 40, 1, 1, 1
    nV1, mI = nblock * bs, nblock * (bs - nE)
    f = lambda *s: rng.standard_normal(s).astype(np.float32)
    h = dict(it0=64, frameCycle=0, iFrameHead=1, oldFrameHead=0, currentFrame=1, iModel=0,
             learning=3, nV1=nV1, nLGN=nLGN, ngTypeFF=ngFF, ngTypeE=ngE, ngTypeI=ngI, nType=2,
             trainDepth=1)
    isE = (np.arange(nV1) % bs) < nE
    lv = {k: np.where(isE, f(nV1), 0).astype(np.float32) for k in ("vLTD", "vTrip", "vAvg")}
    post = S.post_to_file(lv, nblock, bs, nE)
    assert post.size == 2 * nE * nblock + nV1 + nE * nblock
    k = nE * nblock
    assert np.array_equal(post[:nE], lv["vLTD"][:nE])              # block 0, eid = local id
    assert np.array_equal(post[k + nE:k + 2 * nE], lv["vTrip"][bs:bs + nE])   # block 1
    assert np.array_equal(post[2 * k:2 * k + 4], [lv["vAvg"][0], 0, lv["vAvg"][1], 0])
    learn = dict(sLGN=f(maxLGN, nV1), post=post, pre=f(nLGN))
    block = dict(depC=f(nV1), v=f(nV1), gFF=f(ngFF, nV1), hFF=f(ngFF, nV1), gE=f(ngE, nV1),
                 gI=f(ngI, nV1), hE=f(ngE, nV1), hI=f(ngI, nV1), gap=f(mI))
    p = tmp_path / "snap_learn.bin"
    S.write_file(p, h, dict(leftTimeRate=f(nLGN), lastNegLogRand=f(nLGN)), dict(tBack=f(nV1)),
                 f(1, nV1), block, learn=learn)
    raw = p.read_bytes()
    assert len(raw) == S.file_size(nV1, nLGN, 0, ngFF, ngE, ngI, mI,
                                   learn=(nblock, nE, maxLGN))
    assert struct.unpack_from("<2i", raw, 44) == (0, 3)
    o = 80 + nLGN * (8 + S.GEN_BYTES)
    assert np.array_equal(np.frombuffer(raw, np.float32, maxLGN * nV1, o),
                          learn["sLGN"].reshape(-1))
    with pytest.raises(ValueError, match="plastic"):
        S.read_file(p)
    d = S.read_file(p, learn=(nblock, nE, maxLGN))
    assert np.array_equal(d["sLGN"], learn["sLGN"]) and np.array_equal(d["pre"], learn["pre"])
    back = S.post_from_file(d["post"], nblock, bs, nE)
    for name in lv:
        assert np.array_equal(back[name], lv[name]), name
    assert np.array_equal(d["gap"], block["gap"])
