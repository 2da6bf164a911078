"""GPU parity of the multi-spike update (compute_V_collect_spike_learnFF,
src/coredynamics.cu:427-1233, used when dt > min tRef): whole simulations
with refractory periods shorter than dt, so that neurons fire several
times in a step, against the oracle bit for bit through the one-step and the
persistent launches, with voltage noise and LGN synaptic failure (the
restarts re-read the LGN inputs and redraw their failure factors), and with
axonal delays (the delayed gather reads spike counts above 1)."""
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


BASE = dict(dt=0.5, tRef=(0.1, 0.08), tonicDep=(3.0, 12.0), bgStrength=(0.4, 0.4), nblock=4)
CASES = {
    "plain": {},
    "white_failff": dict(noisyDep=(0.04, 0.08), noiseOnTonic=0, synFailFF=(0.3, 0.2),
                         synPerConFF=(20.0, 30.0)),
    "ou": dict(noisyDep=(0.04, 0.08), noiseOnTonic=0, tau_noise=3.0),
    "one_type": dict(tRef=(1.0, 0.08)),
}


def _net(**kw):
    from visual_cortex_amd import network
    base = dict(BASE)
    base.update(kw)
    return network.build_network(network.spec_tiny(**base))


@pytest.mark.parametrize("persistent", [False, True])
@pytest.mark.parametrize("case", list(CASES))
def test_multi_spike_bitwise(pkg, case, persistent):
    import oracle
    from visual_cortex_amd.sim import GpuSim
    nsteps = 160
    _flush(f"test_multi_spike_bitwise[{case}, persistent {persistent}]: {nsteps} steps")
    net = _net(**CASES[case])
    g = GpuSim(net, persistent=persistent)
    assert g.persistent == persistent
    o = oracle.OracleSim(net, pkg, tables=_tables(g))
    g.run(nsteps // 2 + 5)
    g.run(nsteps - nsteps // 2 - 5)
    o.run(nsteps)
    sg, so = g.sInfo(), o.sInfo()
    nmulti = int((so >= 2).sum())
    _flush(f"  V1 spike steps gpu={int((sg >= 1).sum())} oracle={int((so >= 1).sum())}, "
           f"with 2+ spikes {nmulti}")
    assert nmulti > 100
    diff = np.argwhere(_bits(sg) != _bits(so))
    assert diff.size == 0, f"first sInfo difference at (step, neuron) {diff[0]}"
    g.finalize()
    sgs = g.state_numpy()
    for k in ("v", "w", "tBack", "gFF", "hFF", "gE", "hE", "gI", "hI"):
        assert np.array_equal(_bits(sgs[k]), _bits(o.st[k])), k


def test_multi_spike_with_delays_bitwise(pkg):
    import oracle
    from visual_cortex_amd.sim import GpuSim
    nsteps = 120
    _flush(f"test_multi_spike_with_delays_bitwise: {nsteps} steps")
    net = _net(noDelay=0, speedOfThought=0.5)
    g = GpuSim(net)
    o = oracle.OracleSim(net, pkg, tables=_tables(g))
    g.run(nsteps)
    o.run(nsteps)
    sg, so = g.sInfo(), o.sInfo()
    assert int((so >= 2).sum()) > 100
    diff = np.argwhere(_bits(sg) != _bits(so))
    assert diff.size == 0, f"first sInfo difference at (step, neuron) {diff[0]}"
