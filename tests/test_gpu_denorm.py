"""GPU parity when a stimulus frame is not a whole number of steps (the
reference's denorm > 1, src/util/util.h:250-278, src/patch.cu:4999-5030,
5969): 60 Hz frames at 8000 steps/s (denorm 3, a frame every 133 1/3 steps)
and 90 Hz (denorm 9); whole runs against the oracle bit for bit through
both V1 launch kinds, and a snapshot resumed at a period boundary."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _tables(g):
    return dict(TW=g.TW.cpu().numpy(), SW=g.SW.cpu().numpy(), SCx=g.SCx.cpu().numpy(),
                SCy=g.SCy.cpu().numpy(), maxConvol=g.maxConvol.cpu().numpy())


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


CASES = {
    "60Hz": dict(frameRate=60, nt=1000),
    "90Hz": dict(frameRate=90, tau=30.0, nKernelSample=30, nt=1700),
}


def _net(**kw):
    from visual_cortex_amd import network
    base = dict(bgStrength=(0.4, 0.4), nblock=2)
    base.update(kw)
    return network.build_network(network.spec_tiny(**base))


@pytest.mark.parametrize("persistent", [False, True])
@pytest.mark.parametrize("case", list(CASES))
def test_denorm_run_bitwise(pkg, case, persistent):
    import oracle
    from visual_cortex_amd.sim import GpuSim
    net = _net(**CASES[case])
    g = GpuSim(net, persistent=persistent)
    tm = g.tm
    assert tm["denorm"] > 1
    nsteps = tm["ntPerFrame"] + tm["ntPerFrame"] // 2 + 37
    _flush(f"test_denorm_run_bitwise[{case}, persistent {persistent}]: denorm {tm['denorm']}, "
           f"{tm['ntPerFrame']} steps per period, {nsteps} steps")
    o = oracle.OracleSim(net, pkg, tables=_tables(g))
    g.run(nsteps // 3)
    g.run(nsteps - nsteps // 3)
    o.run(nsteps)
    sg, so = g.sInfo(), o.sInfo()
    lg, lo = g.lgn_sinfo(), o.lgn_sinfo()
    assert (so >= 1).sum() > 100 and (lo >= 1).sum() > 20
    n = min(len(lg), len(lo))
    assert np.array_equal(_bits(lg[:n]), _bits(lo[:n]))
    diff = np.argwhere(_bits(sg) != _bits(so))
    assert diff.size == 0, f"first sInfo difference at (step, neuron) {diff[0]}"


def test_denorm_snapshot_resume_bitwise(pkg, tmp_path):
    from visual_cortex_amd import snapshot
    from visual_cortex_amd.sim import GpuSim
    _flush("test_denorm_snapshot_resume_bitwise: 60 Hz frames")
    net = _net(frameRate=60, nt=1300, tonicDep=(0.9, 0.9))
    a = GpuSim(net)
    npf, den = a.tm["ntPerFrame"], a.tm["denorm"]
    a.run(npf)
    path = tmp_path / "snapShot_denorm.bin"
    h = snapshot.save(a, path)
    assert h["currentFrame"] == den
    a.run(npf // 2 + 11)
    ref = a.sInfo()[npf:]
    b = GpuSim(net)
    snapshot.restore(b, path)
    b.run(npf // 2 + 11)
    assert (ref >= 1).sum() > 0
    assert np.array_equal(_bits(b.sInfo()), _bits(ref))
