"""Batched trials on the GPU: independent trials of uneven size, orientation
and seed concatenated into one network give, trial by trial, bit-identical
results to each trial run alone, and match the oracle on the batch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _trials(pkg):
    from visual_cortex_amd import batch, network
    specs = [network.spec_tiny(nblock=nb, seed=100 + i, ori_deg=22.5 * i, bgStrength=(0.4, 0.4))
             for i, nb in enumerate((1, 3, 2))]
    return [batch.build_network_fast(s, device="cuda") for s in specs]


def test_batched_equals_serial_and_oracle(pkg):
    import oracle
    from visual_cortex_amd import batch
    from visual_cortex_amd.sim import GpuSim
    print("test_batched_equals_serial_and_oracle: 3 trials, 192 steps", flush=True)
    nets = _trials(pkg)
    bnet = batch.batch_networks(nets)
    g = GpuSim(bnet)
    g.run(192)
    sb = g.sInfo()
    assert (sb >= 1).sum() > 0
    off = 0
    for t, n in enumerate(nets):
        gs = GpuSim(n)
        gs.run(192)
        st = gs.sInfo()
        assert np.array_equal(sb[:, off:off + n.spec.nV1].view(np.uint32), st.view(np.uint32)), t
        off += n.spec.nV1
    tables = dict(TW=g.TW.cpu().numpy(), SW=g.SW.cpu().numpy(), SCx=g.SCx.cpu().numpy(),
                  SCy=g.SCy.cpu().numpy(), maxConvol=g.maxConvol.cpu().numpy())
    o = oracle.OracleSim(bnet, pkg, tables=tables)
    o.run(192)
    assert np.array_equal(o.sInfo().view(np.uint32), sb.view(np.uint32))
    assert np.array_equal(o.lgn_sinfo().view(np.uint32), g.lgn_sinfo().view(np.uint32))


def test_autotuned_shapes_are_bitwise_equal(pkg):
    """Every one-step launch shape the autotuner may pick gives the same
    results bit for bit (workgroup size and gather tile change only the
    schedule), and the autotuner returns one of them."""
    import numpy as np
    from visual_cortex_amd import network, tune
    from visual_cortex_amd.sim import GpuSim
    print("test_autotuned_shapes_are_bitwise_equal: tiny, 192 steps", flush=True)
    net = network.build_network(network.spec_tiny(nt=256, bgStrength=(0.4, 0.4)))
    ref = GpuSim(net, persistent=False)
    ref.run(192)
    s_ref = ref.sInfo()
    assert (s_ref >= 1).sum() > 0
    for wg, t in tune.candidates(net.spec.blockSize):
        g = GpuSim(net, persistent=False)
        g.p.stepWG, g.p.gatherTile = wg, t
        g.run(192)
        assert np.array_equal(g.sInfo().view(np.uint32), s_ref.view(np.uint32)), (wg, t)
    tuned = GpuSim(net, persistent=False, tune=True)
    assert tuned.shape in tune.candidates(net.spec.blockSize)
    tuned.run(192)
    assert np.array_equal(tuned.sInfo().view(np.uint32), s_ref.view(np.uint32))
