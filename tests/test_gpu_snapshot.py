"""Snapshot / resume on the GPU: a run restored from a snapshot file
continues bit for bit like the run that wrote it (V1 output of every step and
the final state), for both V1 launch kinds and both neuron models; the
reference's continuation contract (src/patch.cu:5489-5684, 6793-6912)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _spec(case):
    from visual_cortex_amd import network
    s = network.spec_tiny(nt=512, tonicDep=(0.9, 0.9))
    return network.lif(s) if case == "lif" else s


CASES = ("adex", "lif")


@pytest.mark.parametrize("persistent", [True, False])
@pytest.mark.parametrize("case", CASES)
def test_resume_is_bitwise(pkg, tmp_path, case, persistent):
    import torch
    from visual_cortex_amd import network, snapshot
    from visual_cortex_amd.sim import GpuSim
    print(f"test_resume_is_bitwise: {case} persistent={persistent}", flush=True)
    net = network.build_network(_spec(case))
    npf = GpuSim(net, record=False).tm["ntPerFrame"]
    first, then = 2 * npf, 3 * npf
    a = GpuSim(net, record=True, persistent=persistent)
    a.run(first)
    path = tmp_path / "snapShot_test.bin"
    h = snapshot.save(a, path)
    assert h["it0"] == first and h["currentFrame"] == first // npf
    a.run(then)
    ref = a.sInfo()[first:]
    ref_state = a.state_numpy()
    assert (ref >= 1.0).sum() > 0, "the case must spike after the snapshot"

    b = GpuSim(net, record=True, persistent=persistent)
    hb = snapshot.restore(b, path)
    assert hb["it0"] == first and b.it == first
    b.run(then)
    got = b.sInfo()
    got_state = b.state_numpy()
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    for k in ref_state:
        assert np.array_equal(got_state[k].view(np.uint32), ref_state[k].view(np.uint32)), k
    torch.cuda.synchronize()


def test_snapshot_only_at_frame_boundary(pkg, tmp_path):
    from visual_cortex_amd import network, snapshot
    from visual_cortex_amd.sim import GpuSim
    print("test_snapshot_only_at_frame_boundary", flush=True)
    net = network.build_network(network.spec_tiny(nt=256))
    g = GpuSim(net, record=False)
    g.run(10)
    with pytest.raises(ValueError):
        snapshot.save(g, tmp_path / "x.bin")


def test_restore_as_initial_condition(pkg, tmp_path):
    from visual_cortex_amd import network, snapshot
    from visual_cortex_amd.sim import GpuSim
    print("test_restore_as_initial_condition", flush=True)
    net = network.build_network(network.spec_tiny(nt=256, tonicDep=(0.9, 0.9)))
    a = GpuSim(net, record=False)
    npf = a.tm["ntPerFrame"]
    a.run(npf)
    path = tmp_path / "s.bin"
    snapshot.save(a, path)
    st = a.state_numpy()
    b = GpuSim(net, record=True)
    snapshot.restore(b, path, as_init=True)
    assert b.it == 0
    sb = b.state_numpy()
    for k in ("v", "tBack", "gE", "hE", "gI", "hI", "gFF", "hFF"):
        assert np.array_equal(sb[k].view(np.uint32), st[k].view(np.uint32)), k
    b.run(npf)
    assert np.isfinite(b.sInfo()).all()


def test_cli_snapshot_and_resume(pkg, tmp_path, monkeypatch):
    """patch_fast's snapshot options: a run of 128 steps writes a snapshot at
    its last step; a second run restored from it (asInit = false) for 64
    more steps reproduces steps 128-191 of one uninterrupted run."""
    import shutil
    from visual_cortex_amd import cli, formats, loader, network
    from visual_cortex_amd.sim import GpuSim
    print("test_cli_snapshot_and_resume: tiny, 128 + 64 steps", flush=True)
    monkeypatch.chdir(tmp_path)                 # snapshots go to the working directory
    net = network.build_network(network.spec_tiny(nt=192, bgStrength=(0.4, 0.4)))
    cfg = loader.save_network(net, tmp_path / "run")
    base = cfg.read_text()
    cfg.write_text(base + "rawData = true\nhWrite = true\nnt = 128\nsnapshotInterval = 64\n")
    assert cli.main(["-c", str(cfg)]) == 0
    snaps = sorted(tmp_path.glob("snapShot_*"))
    assert len(snaps) == 1                      # delPrevSnapshot: only the last is kept
    shutil.copy(snaps[0], tmp_path / "resume_vc.bin")
    cfg.write_text(base + "rawData = true\nhWrite = true\nnt = 64\nsnapshotInterval = 100000\n"
                   f"fSnapshot = !{tmp_path}/resume\nsnapshot_suffix = vc\nasInit = false\n")
    assert cli.main(["-c", str(cfg)]) == 0
    raw = formats.read_raw_data(tmp_path / "run" / "rawData-vc.bin")
    g = GpuSim(net)
    g.run(192)
    direct = g.sInfo()[128:]
    assert raw["nt"] == 64
    assert np.array_equal(raw["sInfo"].view(np.uint32), direct.view(np.uint32))
