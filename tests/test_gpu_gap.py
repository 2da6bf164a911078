"""GPU parity of the gap junctions between inhibitory neurons (the
reference's InhGap: the gap sums of recal_G_mat_nd / recal_G_mat,
src/coredynamics.cu:2524-2540, 2237-2260, and the gapS / gap terms of the
update, src/coredynamics.cuh:107-117): vc_v1_gap against the oracle's
orc_v1_gap on random step outputs, and whole simulations with gap junctions
against the oracle's -- undelayed and delayed, with InhGap off (gapS alone,
as the reference's fast update kernel applies it), at the small plain
configuration's size -- all bit for bit; the per-step gap sums of rawData,
snapshot resume, a network read back from its files, and the persistent
launch's refusal."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _tables(g):
    return dict(TW=g.TW.cpu().numpy(), SW=g.SW.cpu().numpy(), SCx=g.SCx.cpu().numpy(),
                SCy=g.SCy.cpu().numpy(), maxConvol=g.maxConvol.cpu().numpy())


def _net(**kw):
    from visual_cortex_amd import network
    base = dict(nGap=6, InhGap=1, bgStrength=(0.4, 0.4), nblock=4)
    base.update(kw)
    return network.build_network(network.spec_tiny(**base))


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("delayed", [False, True])
def test_gap_kernel_bitwise(pkg, delayed):
    import oracle
    from visual_cortex_amd.abi import check, ptr
    from visual_cortex_amd.sim import GpuSim
    _flush(f"test_gap_kernel_bitwise: 4 blocks, nGap 6, delayed {delayed}")
    net = _net(noDelay=0 if delayed else 1, speedOfThought=0.25)
    sp = net.spec
    g = GpuSim(net)
    o = oracle.OracleSim(net, pkg, tables=_tables(g))
    assert g.gap_on and not g.persistent
    rng = np.random.default_rng(11)
    D = g.depth if delayed else 1
    for cur in ((0, D // 2, D - 1) if delayed else (0,)):
        train = rng.uniform(-70.0, -52.0, (D, sp.nV1)).astype(np.float32)
        spk = rng.random((D, sp.nV1)) < 0.15
        train[spk] = (1.0 + rng.random(int(spk.sum()))).astype(np.float32)
        o._gap(train, cur)
        tr = torch.as_tensor(train, device="cuda")
        lag = ptr(g.gap_lag) if delayed else None
        check(g.lib.vc_v1_gap(None, C.byref(g.p), C.byref(g.s), C.byref(g.rec), ptr(tr), D, cur,
                              lag), "vc_v1_gap")
        torch.cuda.synchronize()
        got = g.gap.cpu().numpy()
        assert np.abs(o.gap).max() > 10.0
        assert np.array_equal(_bits(got), _bits(o.gap)), cur


CASES = {
    # name: (spec kwargs, steps)
    "undelayed": (dict(), 256),
    "gapS_only": (dict(InhGap=0), 256),
    "delayed": (dict(noDelay=0, speedOfThought=0.5), 200),
    "small_plain": (None, 640),
}


@pytest.mark.parametrize("case", list(CASES))
def test_gap_simulation_bitwise(pkg, case):
    import oracle
    from visual_cortex_amd import network
    from visual_cortex_amd.sim import GpuSim
    kw, nsteps = CASES[case]
    _flush(f"test_gap_simulation_bitwise[{case}]: {kw} {nsteps} steps")
    if kw is None:                       # 4 x 1024 conductance-LIF neurons, 1024 LGN cells
        net = network.build_network(network.spec_c1(nGap=8, InhGap=1))
    else:
        net = _net(**kw)
    sp = net.spec
    g = GpuSim(net)
    assert g.gapS is not None and not g.persistent
    assert g.gap_on == bool(sp.InhGap)
    o = oracle.OracleSim(net, pkg, tables=_tables(g))
    if g.gap_on:
        assert np.array_equal(_bits(g.gap.cpu().numpy()), _bits(o.gap))   # pre-loop sums
    g.run(nsteps // 2)
    g.run(nsteps - nsteps // 2)
    o.run(nsteps)
    sg, so = g.sInfo(), o.sInfo()
    isI = (np.arange(sp.nV1) % sp.blockSize) >= sp.nE
    nI_sp = int((so[:, isI] >= 1).sum())
    _flush(f"  V1 spikes gpu={int((sg >= 1).sum())} oracle={int((so >= 1).sum())} "
           f"(inhibitory {nI_sp})")
    assert nI_sp > 20
    diff = np.argwhere(_bits(sg) != _bits(so))
    assert diff.size == 0, f"first sInfo difference at (step, neuron) {diff[0]}"
    g.finalize()                         # the last step's pending recurrent increments
    sgs = g.state_numpy()
    for k in ("v", "w", "tBack", "gFF", "hFF", "gE", "hE", "gI", "hI"):
        assert np.array_equal(_bits(sgs[k]), _bits(o.st[k])), k
    if g.gap_on:
        assert np.array_equal(_bits(g.gap.cpu().numpy()), _bits(o.gap))


def test_gap_changes_the_dynamics(pkg):
    """The same network with InhGap on and off, and without junctions: three
    different runs (gapS alone already changes the inhibitory leak)."""
    from visual_cortex_amd.sim import GpuSim
    _flush("test_gap_changes_the_dynamics")
    runs = []
    for kw in (dict(), dict(InhGap=0), dict(nGap=0, InhGap=0)):
        g = GpuSim(_net(**kw))
        g.run(192)
        runs.append(g.sInfo())
    assert not np.array_equal(runs[0], runs[1])
    assert not np.array_equal(runs[1], runs[2])


def test_gap_outputs_snapshot_and_files(pkg, tmp_path):
    """run_with_outputs writes each step's gap sums to rawData (the oracle's,
    bit for bit); a snapshot taken mid-run resumes bit for bit; the network
    saved in the reference's files (V1_gapMat, V1_gapVec) and read back gives
    the same spikes."""
    import oracle
    from visual_cortex_amd import config, formats as F, loader, snapshot
    from visual_cortex_amd.sim import GpuSim
    _flush("test_gap_outputs_snapshot_and_files")
    net = _net()
    sp = net.spec
    g = GpuSim(net, record=False)
    npf = g.tm["ntPerFrame"]
    n = 2 * npf
    paths = dict(rawData=tmp_path / "raw.bin", lgn_fr=tmp_path / "fr.bin",
                 lgn_sp=tmp_path / "sp.bin")
    g.run_with_outputs(n, paths, dict(rawData=True, lgn_fr=False, lgn_sp=False))
    raw = F.read_raw_data(paths["rawData"])
    o = oracle.OracleSim(net, pkg, tables=_tables(g))
    gaps = []
    for _ in range(n):
        o.run(1)
        gaps.append(o.gap.copy())
    assert np.array_equal(_bits(raw["gap"]), _bits(np.stack(gaps)))
    assert np.array_equal(_bits(raw["sInfo"]), _bits(o.sInfo()))
    # snapshot after two periods, resume for two more
    a = GpuSim(net)
    a.run(2 * npf)
    path = tmp_path / "snap.bin"
    snapshot.save(a, path)
    a.run(2 * npf)
    b = GpuSim(net)
    snapshot.restore(b, path)
    assert np.array_equal(_bits(snapshot.read_file(path)["gap"]), _bits(gaps[-1]))
    b.run(2 * npf)
    assert np.array_equal(_bits(b.sInfo()), _bits(a.sInfo()[2 * npf:]))
    # files
    cfgp = loader.save_network(net, tmp_path / "in")
    back = loader.load_network(config.Config.from_file(cfgp))
    c = GpuSim(back)
    assert c.gap_on
    c.run(4 * npf)
    assert np.array_equal(_bits(c.sInfo()), _bits(a.sInfo()))


def test_gap_persistent_launch_refuses(pkg):
    from visual_cortex_amd import abi
    from visual_cortex_amd.abi import ptr
    from visual_cortex_amd.sim import GpuSim
    _flush("test_gap_persistent_launch_refuses")
    net = _net()
    g = GpuSim(net)
    sp = net.spec
    out = torch.empty(4, sp.nV1, device="cuda")
    args = (None, C.byref(g.p), C.byref(g.s), C.byref(g.ff), C.byref(g.rec), g.spk2, 0,
            ptr(g.lgn_sInfo2[0]), sp.nLGN, None, sp.nV1, 0, 4, ptr(out), sp.nV1)
    rc = g.lib.vc_v1_run_persistent(*args, ptr(g.scratch), 1, ptr(g.status))
    assert rc == abi.VC_E_UNSUPPORTED
    rc = g.lib.vc_v1_run_persistent_noalloc(*args, ptr(g.status))
    assert rc == abi.VC_E_UNSUPPORTED
    assert "gap" in g.lib.vc_last_error().decode()
