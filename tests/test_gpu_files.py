"""The file-level drop-in on the GPU: a network saved in the reference's
input files and run through the cfg-driven CLI produces the same spikes as
the same network run directly, and its rawData output decodes with the
reference reader's semantics."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_cli_from_files_matches_direct_run(pkg, tmp_path, monkeypatch):
    from visual_cortex_amd import cli, formats, loader, network
    from visual_cortex_amd.sim import GpuSim
    print("test_cli_from_files_matches_direct_run: tiny, 192 steps", flush=True)
    monkeypatch.chdir(tmp_path)                 # the last-step snapshot goes to the cwd
    net = network.build_network(network.spec_tiny(nt=192, bgStrength=(0.4, 0.4)))
    cfg = loader.save_network(net, tmp_path / "run")
    with open(cfg, "a") as f:
        f.write("rawData = true\nsaveLGN_fr = true\ngetLGN_sp = true\nhWrite = true\n")
    assert cli.main(["-c", str(cfg)]) == 0
    raw = formats.read_raw_data(tmp_path / "run" / "rawData-vc.bin")
    assert raw["nt"] == 192 and raw["nV1"] == net.spec.nV1
    g = GpuSim(net)
    g.run(192)
    s_direct = g.sInfo()
    assert np.array_equal(raw["sInfo"].view(np.uint32), s_direct.view(np.uint32))
    g.finalize()
    st = g.state_numpy()
    assert np.array_equal(raw["v"][-1], st["v"])
    assert np.array_equal(raw["gE"][-1], st["gE"]) and np.array_equal(raw["hI"][-1], st["hI"])
    lsp = np.fromfile(tmp_path / "run" / "LGN_sp-vc.bin", "<f4").reshape(192, -1)
    assert np.array_equal(lsp.view(np.uint32), g.lgn_sinfo()[:192].view(np.uint32))
    spikes = formats.spikes_from_sinfo(raw["sInfo"], net.spec.dt)
    assert sum(len(x) for x in spikes) == int((s_direct >= 1).sum()) > 0


def test_cli_delayed_run_from_files(pkg, tmp_path, monkeypatch):
    """noDelay = 0 from the files (distances from V1_delayMat) gives the same
    spikes as the direct delayed run (distances computed on the device)."""
    from visual_cortex_amd import cli, formats, loader, network
    from visual_cortex_amd.sim import GpuSim
    print("test_cli_delayed_run_from_files: tiny, 192 steps, speed 0.5", flush=True)
    monkeypatch.chdir(tmp_path)
    net = network.build_network(network.spec_tiny(nt=192, bgStrength=(0.4, 0.4), noDelay=0,
                                                  speedOfThought=0.5))
    cfg = loader.save_network(net, tmp_path / "run")
    with open(cfg, "a") as f:
        f.write("rawData = true\nhWrite = true\n")
    assert cli.main(["-c", str(cfg)]) == 0
    raw = formats.read_raw_data(tmp_path / "run" / "rawData-vc.bin")
    g = GpuSim(net)
    assert g.delay
    g.run(192)
    s_direct = g.sInfo()
    assert int((s_direct >= 1).sum()) > 100
    assert np.array_equal(raw["sInfo"].view(np.uint32), s_direct.view(np.uint32))
    st = g.state_numpy()
    assert np.array_equal(raw["gE"][-1], st["gE"]) and np.array_equal(raw["hI"][-1], st["hI"])


def test_cli_plastic_run_writes_slgn(pkg, tmp_path, monkeypatch):
    """learning = 3 from the cfg: the CLI's outputs and its sLGN file (one
    record every sampleInterval_LGN_V1 steps) equal the direct plastic run's
    spikes and strengths."""
    from visual_cortex_amd import cli, formats, loader, network
    from visual_cortex_amd.sim import GpuSim
    print("test_cli_plastic_run_writes_slgn: tiny, 192 steps, learning 3", flush=True)
    monkeypatch.chdir(tmp_path)
    keys = dict(tauLTP=16.8, tauLTD=33.7, tauTrip=114.0, r_LTD=1.0, A_LGN=0.05,
                tauAvg=(1000.0, 1000.0), targetFR=(5.0, 5.0), gmaxLGN=5.0, gminLGN=0.0)
    net = network.build_network(network.spec_tiny(nt=192, bgStrength=(0.4, 0.4), learning=3,
                                                  sampleInterval_LGN_V1=64, learnData_FF=1,
                                                  **keys))
    cfg = loader.save_network(net, tmp_path / "run")
    with open(cfg, "a") as f:
        f.write("rawData = true\nhWrite = true\nlearning = 3\nsampleInterval_LGN_V1 = 64\n")
        for k, v in keys.items():
            f.write(f"{k} = {', '.join(str(x) for x in v) if isinstance(v, tuple) else v}\n")
    assert cli.main(["-c", str(cfg)]) == 0
    raw = formats.read_raw_data(tmp_path / "run" / "rawData-vc.bin")
    paths = sorted((tmp_path / "run").glob("sLGN*"))
    assert len(paths) == 1, paths
    d = formats.read_slgn(paths[0])
    assert (d["nt"], d["sampleInterval"], d["nLearnTypeFF_E"], d["nLearnTypeFF_I"]) == (192, 64, 1, 0)
    assert d["sLGN"].shape == (3, net.sLGN.shape[0], net.spec.nV1)
    g = GpuSim(net)
    assert g.learn
    for k in range(3):
        g.run(64)
        want = g.learn_state_numpy()["sLGN"]
        assert np.array_equal(d["sLGN"][k].view(np.uint32), want.view(np.uint32)), k
    assert (d["sLGN"][-1] != net.sLGN).sum() > 50
    assert np.array_equal(raw["sInfo"].view(np.uint32), g.sInfo().view(np.uint32))
