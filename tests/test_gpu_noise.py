"""GPU parity of the voltage noise (noisyDep: white, Ornstein-Uhlenbeck with
tau_noise, noiseOnTonic) and the synaptic failure (synFailFF on LGN
synapses, synFail on V1 ones) of the reference's update and recal_G
kernels (src/coredynamics.cu:1424-1445, 1476-1480, 2766-2769): whole
simulations against the oracle bit for bit, through the one-step and the
persistent launches, batched trials against the trials run alone, snapshot
resume of the noise state, and the refusals of the paths without the
draws."""
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


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


NOISE = dict(noisyDep=(0.04, 0.08), noiseOnTonic=0)
OU = dict(noisyDep=(0.04, 0.08), noiseOnTonic=0, tau_noise=3.0)
TONIC = dict(noisyDep=(0.1, 0.2), noiseOnTonic=1, tonicDep=(0.7, 0.8))
FAILFF = dict(synFailFF=(0.3, 0.2), synPerConFF=(20.0, 30.0))
FAIL = dict(synFail=(0.2, 0.3, 0.4, 0.1), synPerCon=(5.0, 7.5, 4.26, 7.88))
CASES = {
    "white": NOISE,
    "ou": OU,
    "on_tonic": TONIC,
    "fail_ff": FAILFF,
    "fail_rec": FAIL,
    "all": {**OU, **FAILFF, **FAIL},
}


def _net(**kw):
    from visual_cortex_amd import network
    base = dict(bgStrength=(0.4, 0.4), nblock=4)
    base.update(kw)
    return network.build_network(network.spec_tiny(**base))


@pytest.mark.parametrize("persistent", [False, True])
@pytest.mark.parametrize("case", list(CASES))
def test_noise_and_failure_bitwise(pkg, case, persistent):
    import oracle
    from visual_cortex_amd.sim import GpuSim
    nsteps = 256
    _flush(f"test_noise_and_failure_bitwise[{case}, persistent {persistent}]: {nsteps} steps")
    net = _net(**CASES[case])
    g = GpuSim(net, persistent=persistent)
    assert g.persistent == persistent
    o = oracle.OracleSim(net, pkg, tables=_tables(g))
    g.run(nsteps // 2 + 3)
    g.run(nsteps - nsteps // 2 - 3)
    o.run(nsteps)
    sg, so = g.sInfo(), o.sInfo()
    _flush(f"  V1 spikes gpu={int((sg >= 1).sum())} oracle={int((so >= 1).sum())}")
    assert (so >= 1).sum() > 100
    diff = np.argwhere(_bits(sg) != _bits(so))
    assert diff.size == 0, f"first sInfo difference at (step, neuron) {diff[0]}"
    g.finalize()
    sgs = g.state_numpy()
    for k in ("v", "w", "tBack", "gFF", "hFF", "gE", "hE", "gI", "hI"):
        assert np.array_equal(_bits(sgs[k]), _bits(o.st[k])), k
    assert np.array_equal(_bits(g.lastNoise.cpu().numpy()), _bits(o.lastNoise))
    if case in ("ou", "all"):
        assert np.abs(o.lastNoise).max() > 0.01


def test_noise_changes_the_dynamics(pkg):
    """Each ingredient changes the spikes of the same network and seed."""
    from visual_cortex_amd.sim import GpuSim
    _flush("test_noise_changes_the_dynamics")
    runs = {}
    for name, kw in [("none", {}), *CASES.items()]:
        g = GpuSim(_net(**kw))
        g.run(160)
        runs[name] = g.sInfo()
    for name in CASES:
        assert not np.array_equal(runs[name], runs["none"]), name


def test_noise_small_plain_configuration(pkg):
    """4 x 1024 neurons, 1024 LGN cells, everything on, 640 steps, through
    the persistent launch."""
    import oracle
    from visual_cortex_amd import network
    from visual_cortex_amd.sim import GpuSim
    _flush("test_noise_small_plain_configuration: spec_c1 + OU noise + failures, 640 steps")
    net = network.build_network(network.spec_c1(**OU, **FAILFF, **FAIL))
    g = GpuSim(net)
    assert g.persistent
    o = oracle.OracleSim(net, pkg, tables=_tables(g))
    g.run(640)
    o.run(640)
    sg, so = g.sInfo(), o.sInfo()
    _flush(f"  V1 spikes gpu={int((sg >= 1).sum())} oracle={int((so >= 1).sum())}")
    assert (so >= 1).sum() > 150
    diff = np.argwhere(_bits(sg) != _bits(so))
    assert diff.size == 0, f"first sInfo difference at (step, neuron) {diff[0]}"


def test_noise_batched_trials(pkg):
    """Trials with noise and failure batched into one launch give each
    trial's own run bit for bit (the draws are keyed by the trial's seed and
    its own neuron ids)."""
    from visual_cortex_amd import batch, network
    from visual_cortex_amd.sim import GpuSim
    _flush("test_noise_batched_trials: 3 trials, 192 steps")
    specs = [network.spec_tiny(nblock=nb, seed=300 + i, ori_deg=30.0 * i, bgStrength=(0.4, 0.4),
                               **OU, **FAILFF, **FAIL)
             for i, nb in enumerate((2, 1, 3))]
    nets = [batch.build_network_fast(s, device="cuda") for s in specs]
    g = GpuSim(batch.batch_networks(nets))
    g.run(192)
    sb = g.sInfo()
    off = 0
    for t, n in enumerate(nets):
        gs = GpuSim(n)
        gs.run(192)
        assert np.array_equal(_bits(sb[:, off:off + n.spec.nV1]), _bits(gs.sInfo())), t
        off += n.spec.nV1


def test_noise_snapshot_resume(pkg, tmp_path):
    """The Ornstein-Uhlenbeck value travels in the snapshot (the rNoisy
    generator slots): a restored run continues bit for bit."""
    from visual_cortex_amd import snapshot
    from visual_cortex_amd.sim import GpuSim
    _flush("test_noise_snapshot_resume")
    net = _net(**OU, **FAIL)
    npf = GpuSim(net).tm["ntPerFrame"]
    a = GpuSim(net)
    a.run(2 * npf)
    path = tmp_path / "snap.bin"
    snapshot.save(a, path)
    ln = snapshot.read_file(path)["lastNoise"]
    assert np.array_equal(_bits(ln), _bits(a.lastNoise.cpu().numpy())) and np.abs(ln).max() > 0
    a.run(2 * npf)
    b = GpuSim(net)
    snapshot.restore(b, path)
    b.run(2 * npf)
    assert np.array_equal(_bits(b.sInfo()), _bits(a.sInfo()[2 * npf:]))


def test_refusals(pkg):
    """Paths without the draws refuse: V1 synaptic failure with conduction
    delays, the snapshot step of the adjoint, vc_v1_rec_apply with failure
    (vc_v1_rec_apply_at names the step), the OU noise without its state."""
    from visual_cortex_amd import abi
    from visual_cortex_amd.abi import ptr
    from visual_cortex_amd.sim import GpuSim
    _flush("test_refusals")
    with pytest.raises(NotImplementedError):
        GpuSim(_net(noDelay=0, speedOfThought=0.5, **FAIL))
    g = GpuSim(_net(**FAIL), persistent=False)
    rc = g.lib.vc_v1_rec_apply(None, C.byref(g.p), C.byref(g.s), C.byref(g.rec),
                               C.byref(g.spk[0]))
    assert rc == abi.VC_E_UNSUPPORTED and "vc_v1_rec_apply_at" in g.lib.vc_last_error().decode()
    row = torch.zeros(g.net.spec.nV1, device="cuda")
    snap = torch.zeros(64, g.net.spec.nV1, device="cuda")
    rc = g.lib.vc_v1_step_snapshot(None, C.byref(g.p), C.byref(g.s), C.byref(g.ff),
                                   C.byref(g.rec), C.byref(g.spk[0]), C.byref(g.spk[1]),
                                   ptr(g.lgn_sInfo2[0]), None, 0, ptr(row), ptr(snap))
    assert rc == abi.VC_E_UNSUPPORTED
    h = GpuSim(_net(**OU), persistent=False)
    h.s.lastNoise = None
    rc = h.lib.vc_v1_step(None, C.byref(h.p), C.byref(h.s), C.byref(h.ff), C.byref(h.rec),
                          C.byref(h.spk[0]), C.byref(h.spk[1]), ptr(h.lgn_sInfo2[0]), None, 0,
                          ptr(row), None, None)
    assert rc == abi.VC_E_INVALID and "lastNoise" in h.lib.vc_last_error().decode()
    torch.cuda.synchronize()
