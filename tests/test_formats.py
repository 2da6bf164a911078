"""The reference's file formats and cfg semantics: round trips, the cfg
parser's Boost program_options rules, and a synthetic network saved in the
reference's input files and loaded back through a cfg."""
import numpy as np
import pytest


def test_cfg_semantics(pkg):
    from visual_cortex_amd import config
    text = """
    # comment
    dt = 0.125 # in ms
    vThres = -20.0, -50.0
    pFF = 0.9, 0.1
    pFF = 0.8, 0.2
    inputFolder = /in/
    conV1_suffix = abc
    fV1_conMat = !/abs/conmat.bin
    [sec]
    x = 3
    """
    c = config.Config(text, {"dt": 0.25})
    assert c.float("dt") == 0.25                                 # command line wins
    assert c.floats("vThres") == [-20.0, -50.0]
    assert c.floats("pFF") == [0.9, 0.1, 0.8, 0.2]                # repeated keys append
    assert c.int("sec.x") == 3
    assert c.path("fNeighborBlock") == "/in/neighborBlock-abc.bin"
    assert c.path("fV1_conMat") == "/abs/conmat.bin-abc.bin"       # '!' drops the folder only
    sp = config.spec_from_config(c)
    assert sp.pFF == ((0.9, 0.1), (0.8, 0.2)) and sp.dt == 0.25
    assert config.spec_from_config(config.Config("InhGap = true")).InhGap == 1
    # noise and failure keys as src/minimal5.cfg:75-120 writes them
    sp = config.spec_from_config(config.Config(
        "tau_noise = 2.5\nnoiseOnTonic = false\nnoisyDep = 0.025, 0.05\nsynFailFF = 0.1\n"
        "synFail = 0.2, 0.0\nsynFail = -0.5, 0.3\nsynPerConFF = 50, 60.0\n"
        "synPerCon = 5.0, 7.5\nsynPerCon = 4.26, 7.88"))
    assert sp.tau_noise == 2.5 and sp.noiseOnTonic == 0 and sp.noisyDep == (0.025, 0.05)
    assert sp.synFailFF == (0.1, 0.1) and sp.synPerConFF == (50.0, 60.0)
    assert sp.synFail == (0.2, 0.0, 0.0, 0.3)                     # negative rates become 0
    assert sp.synPerCon == (5.0, 7.5, 4.26, 7.88)
    assert config.spec_from_config(config.Config("")).noiseOnTonic == 1   # the default
    with pytest.raises(ValueError):
        config.spec_from_config(config.Config("synFail = 0.1, 0.1"))     # needs nType x nType
    sp = config.spec_from_config(config.Config("spE0 = 0.1, 0\nspI0 = 2, 3\nSCsplit = 4"))
    assert sp.spE0 == (0.1, 0.0) and sp.spI0 == (2.0, 3.0) and sp.SCsplit == 4
    with pytest.raises(ValueError):
        config.spec_from_config(config.Config("spE0 = 0.1"))             # [type][simple, complex]
    sp = config.spec_from_config(config.Config(
        "gFF0 = 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8\n"
        "gFF0 = 1.1, 1.2, 1.3, 1.4, 1.5, 1.6, 1.7, 1.8\n"
        "gE0 = 2, 0.1, 3, 0.2, 2, 0.1, 3, 0.2\ngE0 = 4, 0.1, 5, 0.2, 4, 0.1, 5, 0.2\n"
        "gI0 = 6, 0.1, 7, 0.2\ngI0 = 8, 0.1, 9, 0.2\nw0 = 0.5, 0.1, 0.2, 0.05\nuse_v0 = true"))
    assert sp.init_gFF0[8] == 1.1 and sp.init_gI0 == (6, 0.1, 7, 0.2, 8, 0.1, 9, 0.2)
    assert sp.init_w0 == (0.5, 0.1, 0.2, 0.05) and sp.use_v0 == 1
    with pytest.raises(ValueError):              # the reference requires all three with w0
        config.spec_from_config(config.Config("gE0 = 2, 0.1, 3, 0.2, 2, 0.1, 3, 0.2"))
    sp = config.spec_from_config(config.Config("noDelay = 0\nnoFarDelay = 0\nspeedOfThought = 0.5"))
    assert sp.noDelay == 0 and sp.speedOfThought == 0.5 and sp.noFarDelay == 0
    sp = config.spec_from_config(config.Config(""))
    assert sp.noDelay == 1 and sp.noFarDelay == 1                       # the reference's defaults
    with pytest.raises(ValueError):              # noFarDelay defaults to 1: refused with delays
        config.spec_from_config(config.Config("noDelay = 0"))
    sp = config.spec_from_config(config.Config("noDelay = 1\nnoFarDelay = 0"))
    assert sp.noDelay == 1 and sp.noFarDelay == 0                     # far delays only


def test_stimulus_and_lists_round_trip(pkg, tmp_path):
    from visual_cortex_amd import formats as F
    fr = np.random.default_rng(0).random((3, 3, 5, 7)).astype(np.float32)
    F.write_stimulus(tmp_path / "s.bin", fr, (0.1, 0.2, 0.3), 0.5, 2.0)
    st = F.read_stimulus(tmp_path / "s.bin")
    assert np.array_equal(st["frames"], fr) and st["range_deg"] == 2.0 and st["buffer_deg"] == 0.5
    lists = [np.array([1, 2, 3], np.uint32), np.zeros(0, np.uint32), np.array([7], np.uint32)]
    F.write_list_of_lists(tmp_path / "l.bin", lists, "<u4")
    back = F.read_list_of_lists(tmp_path / "l.bin", "<u4")
    assert all(np.array_equal(a, b) for a, b in zip(lists, back))


def test_raw_data_round_trip_and_spike_decoding(pkg, tmp_path):
    from visual_cortex_amd import formats as F
    nV1, nt = 8, 3
    w = F.RawDataWriter(tmp_path / "r.bin", 0.125, nt, nV1, 1, 2, True, 2, 2, 1)
    rng = np.random.default_rng(1)
    recs = []
    for it in range(nt):
        sI = np.full(nV1, -60.0, np.float32)
        sI[it] = 1.25                                       # a spike at 1/4 of the step
        parts = [sI] + [rng.random(nV1).astype(np.float32) for _ in range(3)] + \
                [rng.random((2, nV1)).astype(np.float32) for _ in range(2)] + \
                [rng.random((2, nV1)).astype(np.float32), rng.random((1, nV1)).astype(np.float32),
                 rng.random((2, nV1)).astype(np.float32), rng.random((1, nV1)).astype(np.float32)]
        w.write_step(*parts)
        recs.append(parts)
    w.close()
    d = F.read_raw_data(tmp_path / "r.bin")
    assert d["nt"] == nt and d["mI"] == 2
    for it in range(nt):
        assert np.array_equal(d["sInfo"][it], recs[it][0])
        assert np.array_equal(d["v"][it], recs[it][3])
        assert np.array_equal(d["gI"][it], recs[it][7])
        assert np.array_equal(d["hI"][it], recs[it][9])
    sp = F.spikes_from_sinfo(d["sInfo"], 0.125)
    assert np.allclose(sp[1], [(1 + 0.25) * 0.125])
    # two spikes in one step are spread around the mean time (readSpike)
    two = F.spikes_from_sinfo(np.array([[2.5]], np.float32), 1.0)[0]
    assert len(two) == 2


def test_network_files_round_trip(pkg, tmp_path):
    """A synthetic network written in the reference's input files and read
    back through a cfg gives identical arrays (masked CSR, LGN->V1, LGN
    parameters, stimulus)."""
    from visual_cortex_amd import config, loader, network
    net = network.build_network(network.spec_tiny(nt=128))
    cfg_path = loader.save_network(net, tmp_path / "in")
    back = loader.load_network(config.Config.from_file(cfg_path))
    a, b = net.masked_csr(), back.masked_csr()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert np.array_equal(back.nbr, net.nbr) and np.array_equal(back.nNear, net.nNear)
    np.testing.assert_array_equal(back.idLGN, net.idLGN)
    np.testing.assert_allclose(back.sLGN, net.sLGN, rtol=1e-6)
    for k in ("polar", "ecc", "rw", "rh", "k", "tauR", "nD", "coneType"):
        assert np.array_equal(back.lgn[k], net.lgn[k]), k
    assert np.array_equal(back.lgn["logA"], net.lgn["logA"])
    assert np.array_equal(back.lgn["logB"], net.lgn["logB"])
    assert np.array_equal(back.frames, net.frames)
    assert back.spec.nLGN == net.spec.nLGN and back.spec.nE == net.spec.nE
    assert back.spec.vThres == net.spec.vThres and back.spec.pE == net.spec.pE


def test_delaymat_round_trip(pkg, tmp_path):
    """With noDelay = 0 the V1_delayMat file is written with every pair's
    distance and read back for the connected pairs: the same float32 values
    the positions give, in the masked CSR's order."""
    from visual_cortex_amd import config, loader, network
    net = network.build_network(network.spec_tiny(nt=64, noDelay=0, speedOfThought=0.5),
                                with_frames=True)
    cfg_path = loader.save_network(net, tmp_path / "in")
    back = loader.load_network(config.Config.from_file(cfg_path))
    assert back.spec.noDelay == 0 and back.spec.speedOfThought == 0.5
    assert back.rec_dist is not None
    n = net.n_rec_synapses_csr()
    a, b = net.rec_distances()[:n], back.rec_distances()[:n]
    assert a.size == n and np.array_equal(a.view(np.uint32), b.view(np.uint32))
    sp = net.spec
    raw = np.fromfile(tmp_path / "in" / "V1_delayMat-vc.bin", "<f4", offset=4)
    assert raw.size == sp.nblock * net.maxNear * sp.blockSize ** 2
    D = raw.reshape(sp.nblock, net.maxNear, sp.blockSize, sp.blockSize)
    assert np.all(D[:, :, np.arange(sp.blockSize), np.arange(sp.blockSize)][:, 0] == 0)  # self


def test_sratio_v1_scales_strengths_by_type_pair(pkg, tmp_path):
    """sRatioV1 = [EE, EI, IE, II] multiplies each V1 strength by the
    factor of its (pre type, post type) pair in float32 on reading, and a
    zero factor removes those synapses."""
    from visual_cortex_amd import config, loader, network
    net = network.build_network(network.spec_tiny(nt=64))
    cfg_path = loader.save_network(net, tmp_path / "in")
    text = open(cfg_path).read()
    assert "sRatioV1 = 1.0\n" in text                # repeated keys append: replace the line
    with open(cfg_path, "w") as f:
        f.write(text.replace("sRatioV1 = 1.0\n", "sRatioV1 = 0.5, 0.25\nsRatioV1 = 2.0, 0.0\n"))
    back = loader.load_network(config.Config.from_file(cfg_path))
    nE = net.spec.nE
    sr = np.array([0.5, 0.25, 2.0, 0.0], np.float32)
    for (p0, q0, w0), (p1, q1, w1) in zip(net.rec_pairs, back.rec_pairs):
        k = (p0 >= nE).astype(int) * 2 + (q0 >= nE).astype(int)
        keep = k != 3
        assert np.array_equal(p1, p0[keep]) and np.array_equal(q1, q0[keep])
        assert np.array_equal(w1.view(np.uint32), (w0 * sr[k])[keep].astype(np.float32).view(np.uint32))


def test_v1_vec_round_trip(pkg, tmp_path):
    """V1_vec written and read back (both connectLongRange forms); sRatioV1
    scales the strengths by pre and post type on reading; a saved delayed
    network with far lists loads them back."""
    from visual_cortex_amd import config, formats, loader, network
    net = network.build_network(network.spec_tiny(nt=64, nblock=4, noDelay=0,
                                                  speedOfThought=0.5, nFar=5))
    sp = net.spec
    off, pre, w, d = net.far
    for clr in (0, 1):
        path = tmp_path / f"V1_vec{clr}.bin"
        formats.write_v1_vec(path, sp.nV1, net.far, connect_long_range=clr)
        assert path.stat().st_size == 4 + 4 * sp.nV1 * (1 + clr) + 12 * pre.size
        o2, p2, w2, d2 = formats.read_v1_vec(path, sp.nV1, sp.blockSize, sp.nE)
        for a, b in ((off, o2), (pre, p2), (w, w2), (d, d2)):
            assert np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))
    sr = [0.5, 2.0, 3.0, 0.25]
    _, _, w3, _ = formats.read_v1_vec(tmp_path / "V1_vec0.bin", sp.nV1, sp.blockSize, sp.nE, sr)
    post = np.repeat(np.arange(sp.nV1), np.diff(off.astype(np.int64)))
    k = (pre % sp.blockSize >= sp.nE) * 2 + (post % sp.blockSize >= sp.nE)
    assert np.array_equal(w3, (w * np.float32(sr)[k]).astype(np.float32))
    assert len(set(k.tolist())) == 4
    cfgp = loader.save_network(net, tmp_path / "in")
    back = loader.load_network(config.Config.from_file(cfgp))
    assert back.spec.noFarDelay == 0 and back.far_delayed()
    for a, b in zip(net.far, back.far):
        assert np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))
    # no far connections: an all-zero nVec file, read back as none
    net0 = network.build_network(network.spec_tiny(nt=64))
    cfg0 = loader.save_network(net0, tmp_path / "in0")
    assert (tmp_path / "in0" / "V1_vec-vc.bin").stat().st_size == 4 + 4 * net0.spec.nV1
    assert loader.load_network(config.Config.from_file(cfg0)).far is None
