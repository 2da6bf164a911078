"""GPU parity of the far (long-range) connections with delays (the
reference's recal_G_vec + sum_G, src/coredynamics.cu:181-372, 2314-2349):
vc_v1_rec_far against the oracle's orc_v1_rec_far on random spike-train
rings, and whole delayed simulations with far lists against the oracle's,
all bit for bit; its status flags; a snapshot of such a run resumes bit for
bit; a network read back from its files runs the same; far connections with
undelayed near ones (noDelay 1, noFarDelay 0) equal the oracle too."""
import ctypes as C
import pathlib
import tempfile

import numpy as np
import pytest
import torch

from test_delay_oracle import random_ring, zero_state
from test_far_oracle import far_call

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _tables(g):
    return dict(TW=g.TW.cpu().numpy(), SW=g.SW.cpu().numpy(), SCx=g.SCx.cpu().numpy(),
                SCy=g.SCy.cpu().numpy(), maxConvol=g.maxConvol.cpu().numpy())


def _net(**kw):
    from visual_cortex_amd import network
    base = dict(noDelay=0, speedOfThought=0.25, bgStrength=(0.4, 0.4), nblock=4, nFar=12)
    base.update(kw)
    return network.build_network(network.spec_tiny(**base))


def ring_masks(train):
    D, n = train.shape
    return np.packbits((train >= 1.0).reshape(D, n // 64, 64), axis=-1,
                       bitorder="little").reshape(D, -1).view(np.int64)


def _state_struct(g, dev):
    from visual_cortex_amd import abi
    from visual_cortex_amd.abi import ptr
    return abi.V1State(*(ptr(dev[k]) for k in ("v", "w", "tBack", "gFF", "hFF", "gE", "hE",
                                                "gI", "hI")), ptr(g.tonic))


@pytest.mark.parametrize("speed", [0.25, 1.0])
def test_rec_far_kernel_bitwise(pkg, speed):
    import oracle
    from visual_cortex_amd.abi import check, ptr
    from visual_cortex_amd.sim import GpuSim
    _flush(f"test_rec_far_kernel_bitwise: 4 blocks, nFar 12, speed {speed}")
    net = _net(speedOfThought=speed)
    sp = net.spec
    g = GpuSim(net)
    assert g.far_on and g.far.maxFar == 12
    o = oracle.OracleSim(net, pkg, tables=_tables(g))
    assert o.far_on and o.depth == g.depth
    D = g.depth
    rng = np.random.default_rng(int(speed * 100) + 1)
    for cur in (0, D // 2, D - 1):
        train = random_ring(rng, D, sp.nV1, p_spike=0.15)
        st = zero_state(net)
        for k in ("gE", "hE", "gI", "hI"):
            st[k][:] = rng.random(st[k].shape).astype(np.float32)
        ref = {k: v.copy() for k, v in st.items()}
        assert far_call(o, ref, net.far, train, D, cur, speed) == 0
        dev = {k: torch.as_tensor(v, device="cuda") for k, v in st.items()}
        s = _state_struct(g, dev)
        tr = torch.as_tensor(train, device="cuda")
        tm = torch.as_tensor(ring_masks(train), device="cuda")
        status = torch.zeros(1, dtype=torch.int32, device="cuda")
        check(g.lib.vc_v1_rec_far(None, C.byref(g.p), C.byref(s), C.byref(g.far), ptr(tr),
                                  ptr(tm), D, cur, speed, ptr(status)), "vc_v1_rec_far")
        torch.cuda.synchronize()
        assert int(status.item()) == 0
        assert int(g.far_t[-1].abs().max()) == 0             # work bits left clear
        hits = 0
        for k in ("gE", "hE", "gI", "hI"):
            a = dev[k].cpu().numpy()
            assert np.array_equal(a.view(np.uint32), ref[k].view(np.uint32)), (cur, k)
            hits += int((a != st[k]).sum())
        assert hits > 50


def test_rec_far_status_flags(pkg):
    """A ring shorter than the delays of spiking far presynaptic neurons sets
    bit 0, an index entry out of range bit 1; the connections concerned add
    nothing, and the work bits are left clear."""
    from visual_cortex_amd import abi
    from visual_cortex_amd.abi import check, ptr
    from visual_cortex_amd.sim import GpuSim
    _flush("test_rec_far_status_flags")
    net = _net()
    sp = net.spec
    g = GpuSim(net)
    st = zero_state(net)
    dev = {k: torch.as_tensor(v, device="cuda") for k, v in st.items()}
    s = _state_struct(g, dev)
    tr = torch.full((1, sp.nV1), 1.5, device="cuda")
    tm = torch.full((1, sp.nV1 // 64), -1, dtype=torch.int64, device="cuda")
    status = torch.zeros(1, dtype=torch.int32, device="cuda")
    check(g.lib.vc_v1_rec_far(None, C.byref(g.p), C.byref(s), C.byref(g.far), ptr(tr), ptr(tm),
                              1, 0, 0.25, ptr(status)), "vc_v1_rec_far")
    torch.cuda.synchronize()
    assert int(status.item()) == 1
    assert float(dev["gE"].abs().max()) == 0.0 and float(dev["gI"].abs().max()) == 0.0
    assert int(g.far_t[-1].abs().max()) == 0
    # presynaptic neuron 0 with one index entry pointing past the last neuron
    i32 = dict(dtype=torch.int32, device="cuda")
    cnt = torch.zeros(sp.nV1, **i32)
    pre = torch.zeros(1, sp.nV1, **i32)
    w = torch.ones(1, sp.nV1, device="cuda")
    dist = torch.zeros(1, sp.nV1, device="cuda")
    pmask = torch.zeros(sp.nV1 // 64, dtype=torch.int64, device="cuda")
    pmask[0] = 1
    outOff = torch.ones(sp.nV1 + 1, **i32)
    outOff[0] = 0
    outPost = torch.tensor([sp.nV1 + 5], **i32)
    outJ = torch.zeros(1, **i32)
    hit = torch.zeros(sp.nV1, dtype=torch.int64, device="cuda")
    far = abi.FarConn(1, ptr(cnt), ptr(pre), ptr(w), ptr(dist), ptr(pmask), ptr(outOff),
                      ptr(outPost), ptr(outJ), ptr(hit))
    status.zero_()
    check(g.lib.vc_v1_rec_far(None, C.byref(g.p), C.byref(s), C.byref(far), ptr(tr), ptr(tm),
                              1, 0, 0.25, ptr(status)), "vc_v1_rec_far")
    torch.cuda.synchronize()
    assert int(status.item()) == 2
    assert float(dev["gE"].abs().max()) == 0.0 and int(hit.abs().max()) == 0


def test_rec_far_gate(pkg):
    """No far presynaptic neuron above 1.0 in the current slot (outputs of
    exactly 1.0, or spikes of other neurons only): nothing is added, though
    earlier spikes arrive; one far presynaptic spike opens it again -- as
    the oracle."""
    import oracle
    from visual_cortex_amd.abi import check, ptr
    from visual_cortex_amd.sim import GpuSim
    _flush("test_rec_far_gate")
    net = _net(nFar=4, nblock=9)
    sp = net.spec
    g = GpuSim(net)
    o = oracle.OracleSim(net, pkg, tables=_tables(g))
    D = g.depth
    isfar = np.unpackbits(net.far_pre_mask().view(np.uint8), bitorder="little").astype(bool)
    assert 0 < isfar.sum() < sp.nV1
    rng = np.random.default_rng(77)
    cur = D // 2
    base = random_ring(rng, D, sp.nV1, p_spike=0.2)
    cases = []
    t1 = base.copy()
    t1[cur] = np.where(isfar, 1.0, t1[cur]).astype(np.float32)      # exactly 1.0: closed
    t1[cur][~isfar] = 1.7                                           # non-far spikes only
    cases.append((t1, False))
    t2 = t1.copy()
    t2[cur][np.nonzero(isfar)[0][3]] = 1.25                         # one far spike: open
    cases.append((t2, True))
    for train, open_ in cases:
        st = zero_state(net)
        ref = {k: v.copy() for k, v in st.items()}
        assert far_call(o, ref, net.far, train, D, cur, sp.speedOfThought) == 0
        dev = {k: torch.as_tensor(v, device="cuda") for k, v in st.items()}
        s = _state_struct(g, dev)
        tr = torch.as_tensor(train, device="cuda")
        tm = torch.as_tensor(ring_masks(train), device="cuda")
        check(g.lib.vc_v1_rec_far(None, C.byref(g.p), C.byref(s), C.byref(g.far), ptr(tr),
                                  ptr(tm), D, cur, sp.speedOfThought, None), "vc_v1_rec_far")
        torch.cuda.synchronize()
        for k in ("gE", "hE", "gI", "hI"):
            a = dev[k].cpu().numpy()
            assert np.array_equal(a.view(np.uint32), ref[k].view(np.uint32)), (open_, k)
        assert (float(np.abs(ref["gE"]).max()) > 0) == open_


FCASES = {
    # name: (spec kwargs, steps)
    "slow": (dict(speedOfThought=0.25), 256),
    "fast": (dict(speedOfThought=2.0), 200),
    "many": (dict(speedOfThought=0.5, nFar=40, nblock=9), 160),
}


@pytest.mark.parametrize("case,events", [(c, True) for c in FCASES] + [("slow", False)])
def test_far_simulation_bitwise(pkg, case, events):
    import oracle
    from visual_cortex_amd.sim import GpuSim
    kw, nsteps = FCASES[case]
    _flush(f"test_far_simulation_bitwise[{case}, events={events}]: {kw} {nsteps} steps")
    net = _net(**kw)
    g = GpuSim(net, ff_events=events)
    assert g.delay and g.far_on
    o = oracle.OracleSim(net, pkg, tables=_tables(g))
    assert o.far_on and o.depth == g.depth
    g.run(nsteps // 2)
    g.run(nsteps - nsteps // 2)
    o.run(nsteps)
    sg, so = g.sInfo(), o.sInfo()
    nsp = int((so >= 1).sum())
    _flush(f"  depth {g.depth}, V1 spikes gpu={int((sg >= 1).sum())} oracle={nsp}")
    assert nsp > 200
    diff = np.argwhere(sg.view(np.uint32) != so.view(np.uint32))
    assert diff.size == 0, f"first sInfo difference at (step, neuron) {diff[0]}"
    sgs = g.state_numpy()
    for k in ("v", "w", "tBack", "gFF", "hFF", "gE", "hE", "gI", "hI"):
        assert np.array_equal(sgs[k].view(np.uint32), o.st[k].view(np.uint32)), k


def test_far_connections_change_the_dynamics(pkg):
    """Without its far lists, or with noFarDelay 1 (no far input, as the
    reference's distance-independent branch adds nothing), the network gives
    another spike train; the two ways of switching far input off agree."""
    from dataclasses import replace
    from visual_cortex_amd.sim import GpuSim
    _flush("test_far_connections_change_the_dynamics")
    net = _net(sFar=0.5)
    a = GpuSim(net)
    a.run(256)
    far = net.far
    net.far = None
    b = GpuSim(net)
    assert not b.far_on
    b.run(256)
    net.far = far
    net.spec = replace(net.spec, noFarDelay=1)
    c = GpuSim(net)
    assert not c.far_on
    c.run(256)
    sa, sb, sc = a.sInfo(), b.sInfo(), c.sInfo()
    assert not np.array_equal(sa, sb)
    assert np.array_equal(sb.view(np.uint32), sc.view(np.uint32))


def test_far_snapshot_resume_bitwise(pkg, tmp_path):
    from visual_cortex_amd import snapshot
    from visual_cortex_amd.sim import GpuSim
    _flush("test_far_snapshot_resume_bitwise: 2+2 frame periods")
    net = _net()
    npf = GpuSim(net).tm["ntPerFrame"]
    full = GpuSim(net)
    full.run(4 * npf)
    ref = full.sInfo()
    a = GpuSim(net)
    a.run(2 * npf)
    path = tmp_path / "snap.bin"
    snapshot.save(a, path)
    b = GpuSim(net)
    snapshot.restore(b, path)
    b.run(2 * npf)
    assert np.array_equal(b.sInfo().view(np.uint32), ref[2 * npf:].view(np.uint32))


def test_far_network_from_files(pkg, tmp_path):
    """Saved in the reference's formats (V1_vec included) and read back, the
    network with far lists gives the same spikes."""
    from visual_cortex_amd import config, loader
    from visual_cortex_amd.sim import GpuSim
    _flush("test_far_network_from_files")
    net = _net()
    cfgp = loader.save_network(net, tmp_path / "in")
    back = loader.load_network(config.Config.from_file(cfgp))
    assert back.n_far() == net.n_far()
    a, b = GpuSim(net), GpuSim(back)
    assert b.far_on and a.depth == b.depth
    a.run(192)
    b.run(192)
    sa = a.sInfo()
    assert int((sa >= 1).sum()) > 100
    assert np.array_equal(sa.view(np.uint32), b.sInfo().view(np.uint32))


@pytest.mark.parametrize("speed", [0.25, 2.0])
def test_far_with_undelayed_near_bitwise(pkg, speed):
    """noDelay 1 with noFarDelay 0: undelayed near synapses (applied in the
    step, recal_G_mat_nd) and delayed far connections after them; the whole
    run equals the oracle's bit for bit, and a snapshot resumes bit for bit."""
    import oracle
    from visual_cortex_amd import snapshot
    from visual_cortex_amd.sim import GpuSim
    _flush(f"test_far_with_undelayed_near_bitwise: speed {speed}")
    net = _net(noDelay=1, noFarDelay=0, speedOfThought=speed)
    g = GpuSim(net)
    assert g.far_only and not g.delay and not g.persistent
    o = oracle.OracleSim(net, pkg, tables=_tables(g))
    assert o.far_only and o.depth == g.depth
    npf = g.tm["ntPerFrame"]
    g.run(2 * npf)
    o.run(4 * npf)
    with tempfile.TemporaryDirectory() as td:
        path = pathlib.Path(td) / "snap.bin"
        snapshot.save(g, path)
        g.run(2 * npf)
        b = GpuSim(net)                  # resumed from the snapshot
        snapshot.restore(b, path)
        b.run(2 * npf)
    sg, so = g.sInfo(), o.sInfo()
    nsp = int((so >= 1).sum())
    _flush(f"  depth {g.depth}, V1 spikes gpu={int((sg >= 1).sum())} oracle={nsp}")
    assert nsp > 200
    diff = np.argwhere(sg.view(np.uint32) != so.view(np.uint32))
    assert diff.size == 0, f"first sInfo difference at (step, neuron) {diff[0]}"
    sgs = g.state_numpy()
    for k in ("v", "w", "tBack", "gFF", "hFF", "gE", "hE", "gI", "hI"):
        assert np.array_equal(sgs[k].view(np.uint32), o.st[k].view(np.uint32)), k
    assert np.array_equal(b.sInfo().view(np.uint32), sg[2 * npf:].view(np.uint32))


@pytest.mark.parametrize("kind", ["delayed", "far_only"])
def test_batched_far_trials_equal_serial(pkg, kind):
    """Trials with far lists batched into one network, with delayed near
    connections and with undelayed ones: each trial's output equals its own
    run bit for bit (per-trial far gate), and there are steps in which one
    trial's gate is open while another's is closed with far arrivals due."""
    from test_far_oracle import FAR_BATCH, gate_splits
    from visual_cortex_amd import batch, network
    from visual_cortex_amd.sim import GpuSim
    _flush(f"test_batched_far_trials_equal_serial[{kind}]: 3 trials, speed 0.5, 160 steps")
    specs = [network.spec_tiny(nblock=nb, seed=310 + i, ori_deg=22.5 * i, bgStrength=(0.4, 0.4),
                               nFar=10, **FAR_BATCH[kind])
             for i, nb in enumerate((4, 2, 3))]
    nets = [network.build_network(s) for s in specs]
    bnet = batch.batch_networks(nets, device="cuda")
    g = GpuSim(bnet)
    assert g.far_on and g.delay == (kind == "delayed") and g.far_only == (kind == "far_only")
    g.run(160)
    sb = g.sInfo()
    off = 0
    serial = []
    for t, n in enumerate(nets):
        gs = GpuSim(n)
        assert gs.far_on and gs.depth <= g.depth
        gs.run(160)
        ss = gs.sInfo()
        serial.append(ss)
        assert int((ss >= 1).sum()) > 40, t
        assert np.array_equal(sb[:, off:off + n.spec.nV1].view(np.uint32), ss.view(np.uint32)), t
        off += n.spec.nV1
    assert gate_splits(nets, serial) > 0
