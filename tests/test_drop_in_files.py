"""The reference's remaining input files and its patchV1_cfg output:
LGN_vpos, LGN_surfaceID, conStats, V1_feature round trips and checks; the
per-neuron tonic drive the reference derives from tonicDep, CmoreDep,
minTonicRatio and the orientation boost (src/patch.cu:2843-2931, 4118-4222),
checked at the points where its formula has a closed form; patchV1_cfg read
back field by field with the layout src/readPatchOutput.py:32-88 parses."""
import math

import numpy as np
import pytest


def test_lgn_vpos_and_surface_round_trip(pkg, tmp_path):
    from visual_cortex_amd import formats as F
    rng = np.random.default_rng(1)
    n = 37
    x, y = rng.random(n).astype(np.float32), rng.random(n).astype(np.float32)
    t = rng.integers(0, 6, n).astype(np.uint32)
    pol, ecc = rng.random(n).astype(np.float32), rng.random(n).astype(np.float32)
    F.write_lgn_vpos(tmp_path / "v.bin", 30, 7, 1.5, 0.1, 2.0, -0.3, 1.0, x, y, t, pol, ecc)
    d = F.read_lgn_vpos(tmp_path / "v.bin")
    assert (d["nLGN_I"], d["nLGN_C"], d["nLGN"]) == (30, 7, 37)
    assert d["max_ecc"] == np.float32(1.5) and d["yspan"] == np.float32(1.0)
    for k, a in (("x", x), ("y", y), ("type", t), ("polar", pol), ("ecc_deg", ecc)):
        assert np.array_equal(d[k], a), k
    sx, sy = np.arange(n) % 7, np.arange(n) // 7
    F.write_lgn_surface_id(tmp_path / "s.bin", sx, sy)
    nsx, nsy, bx, by = F.read_lgn_surface_id(tmp_path / "s.bin", n)
    assert (nsx, nsy) == (7, 6) and np.array_equal(bx, sx) and np.array_equal(by, sy)
    sy2 = sy.copy()
    sy2[3] = sy2[10]
    sx2 = sx.copy()
    sx2[3] = sx2[10]
    F.write_lgn_surface_id(tmp_path / "dup.bin", sx2, sy2)
    with pytest.raises(ValueError, match="share"):
        F.read_lgn_surface_id(tmp_path / "dup.bin", n)


def test_constats_and_feature_round_trip(pkg, tmp_path):
    from visual_cortex_amd import formats as F
    ff = np.linspace(0, 1, 64, dtype=np.float32)
    F.write_constats(tmp_path / "c.bin", 2, ff)
    assert np.array_equal(F.read_constats(tmp_path / "c.bin", 2, 64), ff)
    with pytest.raises(ValueError):
        F.read_constats(tmp_path / "c.bin", 2, 65)
    feat = np.stack([np.sign(np.sin(np.arange(64))), np.arange(64) / 64.0]).astype(np.float32)
    F.write_v1_feature(tmp_path / "f.bin", feat)
    assert np.array_equal(F.read_v1_feature(tmp_path / "f.bin", 64), feat)


def _cfg(text):
    from visual_cortex_amd import config
    return config.Config(text)


def _tonic(text, ff, feature=None, nblock=2, bs=64, typeAcc=(48, 64), tonicDep=(0.8, 0.6)):
    from visual_cortex_amd import loader
    return loader.tonic_drive(_cfg(text), tonicDep, np.array(typeAcc), nblock, bs, ff, feature)


def test_tonic_drive_closed_forms(pkg):
    rng = np.random.default_rng(2)
    nblock, bs = 2, 64
    ff = rng.uniform(0.2, 0.7, nblock * bs).astype(np.float32)
    isE = (np.arange(nblock * bs) % bs) < 48
    # without CmoreDep and boost: the type's tonicDep
    t = _tonic("CmoreDep = false\nminTonicRatio = 0.5\nreadFeature = false", ff)
    assert np.array_equal(t, np.where(isE, np.float32(0.8), np.float32(0.6)))
    # CmoreDep: the type's largest ffRatio gets tonicDep*minTonicRatio, the
    # smallest tonicDep, and the drive is affine in ffRatio in between
    t = _tonic("CmoreDep = true\nminTonicRatio = 0.25, 0.5\nreadFeature = false", ff)
    for typ, td, m in ((isE, 0.8, 0.25), (~isE, 0.6, 0.5)):
        f, v = ff[typ], t[typ]
        assert v[np.argmax(f)] == pytest.approx(td * m, rel=1e-6)
        assert v[np.argmin(f)] == pytest.approx(td, rel=1e-6)
        lin = td * (m + (1 - m) * (f.max() - f) / (f.max() - f.min()))
        np.testing.assert_allclose(v, lin, rtol=1e-5)
    # the default of CmoreDep is true (src/patch.cu:275)
    assert np.array_equal(_tonic("minTonicRatio = 0.25, 0.5\nreadFeature = false", ff), t)


def test_orientation_boost_closed_forms(pkg):
    """von Mises boost around iOri/nOri: 1 at the preferred orientation, the
    baseline at the orthogonal one; kappa = 0 gives the baseline everywhere."""
    nblock, bs, nOri, iOri = 1, 64, 6, 2
    fOri = (iOri % nOri) / nOri + 0.5                  # 0.8333..
    op = np.full(nblock * bs, fOri, np.float32)
    op[1::2] = fOri - 0.5                              # the orthogonal orientation
    feat = np.stack([np.ones_like(op), op])
    ff = np.full(nblock * bs, 0.5, np.float32)         # every ffRatio equal: no CmoreDep scaling
    base, kap = 0.3, 2.0
    text = (f"nOri = {nOri}\niOri = {iOri}\nboostOri = {base}, {kap}\nboostOri = {base}, {kap}\n"
            "minTonicRatio = 0.5\nCmoreDep = true")
    t = _tonic(text, ff, feat, nblock=nblock, bs=bs)
    td = np.where((np.arange(bs) % bs) < 48, 0.8, 0.6)
    np.testing.assert_allclose(t[0::2], td[0::2], rtol=1e-6)
    np.testing.assert_allclose(t[1::2], td[1::2] * base, rtol=1e-5)
    # in between: the von Mises shape base + (1 - base) (exp(k(cos 2 pi d - 1)) - e^-2k)/(1 - e^-2k)
    d = 0.125
    op2 = np.full(nblock * bs, fOri + d, np.float32)
    t2 = _tonic(text, ff, np.stack([np.ones_like(op2), op2]), nblock=nblock, bs=bs)
    e = math.exp(-2 * kap)
    b = base + (1 - base) * (math.exp(kap * (math.cos(2 * math.pi * d) - 1)) - e) / (1 - e)
    np.testing.assert_allclose(t2, td * b, rtol=1e-5)
    # kappa 0: the baseline
    t0 = _tonic(text.replace(f"{base}, {kap}", f"{base}, 0.0"), ff, feat, nblock=nblock, bs=bs)
    np.testing.assert_allclose(t0, td * base, rtol=1e-6)


def test_feature_options_refused_like_the_reference(pkg):
    from visual_cortex_amd import loader
    with pytest.raises(ValueError, match="readFeature"):
        loader.feature_needed(_cfg("minTonicRatio = 0.5"))            # readFeature true, nOri 0
    with pytest.raises(ValueError, match="iOri"):
        loader.feature_needed(_cfg("nOri = 4\nboostOri = 0.5, 1.0"))
    assert not loader.feature_needed(_cfg("readFeature = false"))


def test_saved_network_with_tonic_files_loads_the_drive(pkg, tmp_path):
    from visual_cortex_amd import config, loader, network
    net = network.build_network(network.spec_tiny(nt=64, tonicDep=(0.5, 0.4)))
    sp = net.spec
    rng = np.random.default_rng(3)
    ff = rng.uniform(0.1, 0.9, sp.nV1).astype(np.float32)
    feat = np.stack([np.ones(sp.nV1), rng.random(sp.nV1)]).astype(np.float32)
    tonic = dict(CmoreDep=True, minTonicRatio=[0.3, 0.6], nOri=8, iOri=[1],
                 boostOri=[0.4, 1.5, 0.5, 0.5], readFeature=True, ffRatio=ff, feature=feat)
    cfgp = loader.save_network(net, tmp_path / "in", tonic=tonic)
    cfg = config.Config.from_file(cfgp)
    back = loader.load_network(cfg)
    ref = loader.tonic_drive(cfg, sp.tonicDep, np.array([sp.nE, sp.blockSize]), sp.nblock,
                             sp.blockSize, ff, feat)
    assert np.array_equal(back.tonicDep, ref)
    assert back.tonicDep.std() > 0


def read_cfg_like_the_reference(path):
    """The field order of src/readPatchOutput.py:32-88 (f4 precision)."""
    with open(path, "rb") as f:
        u4 = lambda n=1: np.fromfile(f, "<u4", n)
        f4 = lambda n=1: np.fromfile(f, "<f4", n)
        out = dict(prec=int(u4()[0]))
        out["vL"], out["vE"], out["vI"] = f4(3)
        nType, nTypeE, nTypeI = (int(x) for x in u4(3))
        out.update(nType=nType, nTypeE=nTypeE, nTypeI=nTypeI)
        for k in ("vR", "vThres", "gL", "vT"):
            out[k] = f4(nType)
        out["nTypeHierarchy"] = u4(2)
        out["typeAcc"] = u4(nType)
        out["sRatioLGN"] = f4(nType)
        out["sRatioV1"] = f4(nType * nType)
        out["frRatioLGN"], out["convolRatio"] = f4(2)
        out["frameRate"] = int(u4()[0])
        ln = int(u4()[0])
        out["inputFn"] = np.fromfile(f, f"S{ln}", 1)[0].decode("UTF-8")
        out["nLGN"], out["nV1"], out["nt"] = (int(x) for x in u4(3))
        out["dt"] = f4()[0]
        out["view"] = f4(5)
        out["tonicDep"] = f4(nType)
        out["noisyDep"] = f4(nType)
        out["virtual_LGN"] = int(np.fromfile(f, "<i4", 1)[0])
        out["seed"] = int(np.fromfile(f, "<u8", 1)[0])
        out["iModel"], out["learning"] = (int(x) for x in np.fromfile(f, "<i4", 2))
        ng = [int(x) for x in u4(3)]
        for k, n in zip(("grFF", "gdFF", "grE", "gdE", "grI", "gdI"),
                        (ng[0], ng[0], ng[1], ng[1], ng[2], ng[2])):
            out[k] = f4(n)
        out["pFF"], out["pE"], out["pI"] = f4(ng[0] * nType), f4(ng[1] * nType), f4(ng[2] * nType)
        for k in ("C", "tRef", "a", "b", "tau_w", "deltaT"):
            out[k] = f4(nType)
        out["rest"] = np.fromfile(f, "<f4")
    return out


def test_patchv1_cfg_layout(pkg, tmp_path):
    from visual_cortex_amd import formats as F, network
    sp = network.spec_c1(seed=77)
    F.write_patchv1_cfg(tmp_path / "p.bin", sp, 1024, "/data/stimulus.bin",
                        (0.5, 0.5, 0.5, 1.0, 1.0), sRatioV1=(1.0, 0.5, 2.0, 1.0))
    d = read_cfg_like_the_reference(tmp_path / "p.bin")
    assert d["prec"] == 4 and (d["vL"], d["vE"], d["vI"]) == (-65.0, 0.0, -75.0)
    assert (d["nType"], d["nTypeE"], d["nTypeI"]) == (2, 1, 1)
    assert np.array_equal(d["vR"], np.float32(sp.vR)) and np.array_equal(d["tRef"], np.float32(sp.tRef))
    assert np.array_equal(d["typeAcc"], [768, 1024])
    nE = d["typeAcc"][d["nTypeHierarchy"][0] - 1]          # as read_cfg derives nE
    assert nE == 768
    assert np.array_equal(d["sRatioV1"], np.float32([1.0, 0.5, 2.0, 1.0]))
    assert d["frameRate"] == 125 and d["inputFn"] == "/data/stimulus.bin"
    assert (d["nLGN"], d["nV1"], d["nt"]) == (1024, 4096, 8000) and d["dt"] == np.float32(0.125)
    assert d["seed"] == 77 and d["iModel"] == 0 and d["learning"] == 0
    assert np.array_equal(d["gdFF"], np.float32(sp.decayTimeFF))
    assert np.array_equal(d["pE"], np.float32([x for r in sp.pE for x in r]))
    # spE0, spI0, v0, w0, gFF0, gE0, gI0, synFail, synFailFF: 2+2+5*4+4+2 floats
    assert d["rest"].size == 30
