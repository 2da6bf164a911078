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
