"""GPU parity of the axonal-delay path (noDelay = 0; the reference's
recal_G_mat, src/coredynamics.cu:2062-2309): the device distances against
the host's, vc_v1_rec_delay against the oracle's orc_v1_rec_delay on random
spike-train rings, and whole delayed simulations against the oracle's, all
bit for bit; snapshots of a delayed run resume bit for bit."""
import ctypes as C

import numpy as np
import pytest
import torch

from test_delay_oracle import delay_call, random_ring, zero_state

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _tables(g):
    return dict(TW=g.TW.cpu().numpy(), SW=g.SW.cpu().numpy(), SCx=g.SCx.cpu().numpy(),
                SCy=g.SCy.cpu().numpy(), maxConvol=g.maxConvol.cpu().numpy())


def _net(**kw):
    from visual_cortex_amd import network
    base = dict(noDelay=0, speedOfThought=0.5, bgStrength=(0.4, 0.4))
    base.update(kw)
    return network.build_network(network.spec_tiny(**base))


def test_device_distances_match_host(pkg):
    from visual_cortex_amd.sim import GpuSim
    _flush("test_device_distances_match_host: tiny spec, speed 0.5")
    net = _net()
    g = GpuSim(net)
    n = net.n_rec_synapses_csr()
    dg = g.dist[:n].cpu().numpy()
    dh = net.rec_distances()[:n]
    assert np.array_equal(dg.view(np.uint32), dh.view(np.uint32))
    assert g.depth == g.train.shape[0] >= 4
    # the per-synapse delay table: float32 arithmetic of src/coredynamics.cu:2173-2186
    f = np.float32
    t2p = dh / f(net.spec.speedOfThought)
    it2post = np.ceil(t2p / f(net.spec.dt)).astype(np.uint64)
    rem = it2post.astype(np.float32) * f(net.spec.dt) - t2p
    tab = g.syn[:n].cpu().numpy().view(np.uint32)
    w = net.masked_csr()[2][:n]
    assert np.array_equal(tab[:, 0], w.view(np.uint32))
    assert np.array_equal(tab[:, 1], rem.view(np.uint32))
    assert np.array_equal(tab[:, 2].astype(np.uint64), it2post)
    assert int(it2post.max()) == g.depth - 1


@pytest.mark.parametrize("speed", [0.125, 0.25, 1.0])   # rings longer and shorter than 32
def test_rec_delay_kernel_bitwise(pkg, speed):
    import oracle
    from visual_cortex_amd import abi
    from visual_cortex_amd.abi import check, ptr
    from visual_cortex_amd.sim import GpuSim
    _flush(f"test_rec_delay_kernel_bitwise: tiny spec, speed {speed}")
    net = _net(speedOfThought=speed)
    sp = net.spec
    g = GpuSim(net)
    o = oracle.OracleSim(net, pkg, tables=_tables(g))
    D = g.depth
    rng = np.random.default_rng(int(speed * 100))
    for cur in (0, D // 2, D - 1):
        train = random_ring(rng, D, sp.nV1)
        bits = np.packbits((train >= 1.0).reshape(D, sp.nV1 // 64, 64), axis=-1,
                           bitorder="little").reshape(D, -1).view(np.int64)
        st = zero_state(net)
        for k in ("gE", "hE", "gI", "hI"):           # nonzero start values
            st[k][:] = rng.random(st[k].shape).astype(np.float32)
        ref = {k: v.copy() for k, v in st.items()}
        assert delay_call(o, net, ref, train, D, cur, speed) == 0
        dev = {k: torch.as_tensor(v, device="cuda") for k, v in st.items()}
        s = abi.V1State(*(ptr(dev[k]) for k in ("v", "w", "tBack", "gFF", "hFF", "gE", "hE",
                                                 "gI", "hI")), ptr(g.tonic))
        tr = torch.as_tensor(train, device="cuda")
        tm = torch.as_tensor(bits, device="cuda")
        status = torch.zeros(1, dtype=torch.int32, device="cuda")
        check(g.lib.vc_v1_rec_delay(None, C.byref(g.p), C.byref(s), C.byref(g.rec),
                                    ptr(g.syn), ptr(tr), ptr(tm), D, cur, 0, ptr(status)),
              "vc_v1_rec_delay")
        torch.cuda.synchronize()
        assert int(status.item()) == 0
        hits = 0
        for k in ("gE", "hE", "gI", "hI"):
            a = dev[k].cpu().numpy()
            assert np.array_equal(a.view(np.uint32), ref[k].view(np.uint32)), (cur, k)
            hits += int((a != st[k]).sum())
        assert hits > 50
    # a ring shorter than the longest delay: flagged, nothing added
    st = zero_state(net)
    dev = {k: torch.as_tensor(v, device="cuda") for k, v in st.items()}
    s = abi.V1State(*(ptr(dev[k]) for k in ("v", "w", "tBack", "gFF", "hFF", "gE", "hE", "gI",
                                             "hI")), ptr(g.tonic))
    tr = torch.ones(1, sp.nV1, device="cuda") * 1.5
    tm = torch.full((1, sp.nV1 // 64), -1, dtype=torch.int64, device="cuda")
    status = torch.zeros(1, dtype=torch.int32, device="cuda")
    check(g.lib.vc_v1_rec_delay(None, C.byref(g.p), C.byref(s), C.byref(g.rec),
                                ptr(g.syn), ptr(tr), ptr(tm), 1, 0, 0, ptr(status)),
          "vc_v1_rec_delay")
    torch.cuda.synchronize()
    assert int(status.item()) == 1
    assert float(dev["gE"].abs().max()) == 0.0


DCASES = {
    # name: (spec kwargs, steps)
    "slow": (dict(speedOfThought=0.25), 256),
    "very_slow": (dict(speedOfThought=0.125), 256),
    "fast": (dict(speedOfThought=2.0), 256),
    "burst": (dict(speedOfThought=0.5, tonicDep=(2.0, 2.0), sE=1.5), 160),
    "mixed_tile": (dict(speedOfThought=0.5, nE=200, nblock=3), 200),
}


@pytest.mark.parametrize("case,events", [(c, True) for c in DCASES] + [("slow", False)])
def test_delayed_simulation_bitwise(pkg, case, events):
    import oracle
    from visual_cortex_amd import params
    from visual_cortex_amd.sim import GpuSim
    kw, nsteps = DCASES[case]
    _flush(f"test_delayed_simulation_bitwise[{case}, events={events}]: {kw} {nsteps} steps")
    net = _net(**kw)
    state = None
    if case == "burst":
        state = params.initial_state(net)
        state["v"][:] = -52.0
    g = GpuSim(net, state=state, ff_events=events)
    assert g.delay and not g.persistent
    o = oracle.OracleSim(net, pkg, tables=_tables(g), state=state)
    assert o.delay and o.depth == g.depth
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


def test_delays_change_the_dynamics(pkg):
    """The same network without delays gives another spike train (the delay
    path is really taken), with the same spike count order of magnitude."""
    from dataclasses import replace
    from visual_cortex_amd import network
    from visual_cortex_amd.sim import GpuSim
    _flush("test_delays_change_the_dynamics")
    net = _net(speedOfThought=0.25)
    a = GpuSim(net)
    a.run(256)
    net.spec = replace(net.spec, noDelay=1)
    b = GpuSim(net)
    assert not b.delay
    b.run(256)
    sa, sb = a.sInfo(), b.sInfo()
    assert not np.array_equal(sa, sb)
    na, nb = int((sa >= 1).sum()), int((sb >= 1).sum())
    assert na > 0 and nb > 0 and 0.2 < na / nb < 5


def test_delayed_snapshot_resume_bitwise(pkg, tmp_path):
    from visual_cortex_amd import snapshot
    from visual_cortex_amd.sim import GpuSim
    _flush("test_delayed_snapshot_resume_bitwise: tiny spec, speed 0.25, 2+2 frame periods")
    net = _net(speedOfThought=0.25)
    npf = GpuSim(net).tm["ntPerFrame"]
    full = GpuSim(net)
    full.run(4 * npf)
    ref = full.sInfo()
    a = GpuSim(net)
    a.run(2 * npf)
    path = tmp_path / "snap.bin"
    h = snapshot.save(a, path)
    assert h["trainDepth"] == a.depth
    b = GpuSim(net)
    snapshot.restore(b, path)
    b.run(2 * npf)
    tail = b.sInfo()
    assert np.array_equal(tail.view(np.uint32), ref[2 * npf:].view(np.uint32))


def test_batched_delayed_trials_equal_serial(pkg):
    """Trials of different sizes with axonal delays batched into one network:
    each trial's output equals its own delayed run bit for bit (the ring is
    as long as the longest trial's delays; a longer ring changes nothing)."""
    from visual_cortex_amd import batch, network
    from visual_cortex_amd.sim import GpuSim
    _flush("test_batched_delayed_trials_equal_serial: 3 trials, speed 0.5, 192 steps")
    specs = [network.spec_tiny(nblock=nb, seed=200 + i, ori_deg=22.5 * i, bgStrength=(0.4, 0.4),
                               noDelay=0, speedOfThought=0.5)
             for i, nb in enumerate((1, 3, 2))]
    nets = [batch.build_network_fast(s, device="cuda") for s in specs]
    bnet = batch.batch_networks(nets)
    assert bnet.spec.noDelay == 0
    g = GpuSim(bnet)
    assert g.delay
    g.run(192)
    sb = g.sInfo()
    assert int((sb >= 1).sum()) > 100
    off = 0
    for t, n in enumerate(nets):
        gs = GpuSim(n)
        assert gs.delay and gs.depth <= g.depth
        gs.run(192)
        assert np.array_equal(sb[:, off:off + n.spec.nV1].view(np.uint32),
                              gs.sInfo().view(np.uint32)), t
        off += n.spec.nV1


@pytest.mark.parametrize("extra", [5, 40])       # window staged in LDS, ring longer than 32
def test_rec_delay_window_of_a_longer_ring(pkg, extra):
    """A ring longer than the near delays need (as far connections make it),
    scanned over the near window only: the same sums as the oracle's dense
    scan of the whole ring; a window shorter than the longest delay is
    flagged."""
    import oracle
    from visual_cortex_amd import abi
    from visual_cortex_amd.abi import check, ptr
    from visual_cortex_amd.sim import GpuSim
    _flush(f"test_rec_delay_window_of_a_longer_ring: window + {extra} slots")
    speed = 0.5
    net = _net(speedOfThought=speed)
    sp = net.spec
    g = GpuSim(net)
    o = oracle.OracleSim(net, pkg, tables=_tables(g))
    W = g.window
    assert W == g.depth and W <= 32
    D = W + extra
    rng = np.random.default_rng(extra)
    for cur in (0, W - 1, D - 1):
        train = random_ring(rng, D, sp.nV1, p_spike=0.1)
        bits = np.packbits((train >= 1.0).reshape(D, sp.nV1 // 64, 64), axis=-1,
                           bitorder="little").reshape(D, -1).view(np.int64)
        st = zero_state(net)
        for k in ("gE", "hE", "gI", "hI"):
            st[k][:] = rng.random(st[k].shape).astype(np.float32)
        ref = {k: v.copy() for k, v in st.items()}
        assert delay_call(o, net, ref, train, D, cur, speed) == 0
        for win in (W, 0):                       # the window, and the whole ring
            dev = {k: torch.as_tensor(v, device="cuda") for k, v in st.items()}
            s = abi.V1State(*(ptr(dev[k]) for k in ("v", "w", "tBack", "gFF", "hFF", "gE",
                                                     "hE", "gI", "hI")), ptr(g.tonic))
            tr = torch.as_tensor(train, device="cuda")
            tm = torch.as_tensor(bits, device="cuda")
            status = torch.zeros(1, dtype=torch.int32, device="cuda")
            check(g.lib.vc_v1_rec_delay(None, C.byref(g.p), C.byref(s), C.byref(g.rec),
                                        ptr(g.syn), ptr(tr), ptr(tm), D, cur, win, ptr(status)),
                  "vc_v1_rec_delay")
            torch.cuda.synchronize()
            assert int(status.item()) == 0
            for k in ("gE", "hE", "gI", "hI"):
                a = dev[k].cpu().numpy()
                assert np.array_equal(a.view(np.uint32), ref[k].view(np.uint32)), (cur, win, k)
    status = torch.zeros(1, dtype=torch.int32, device="cuda")
    check(g.lib.vc_v1_rec_delay(None, C.byref(g.p), C.byref(s), C.byref(g.rec), ptr(g.syn),
                                ptr(tr), ptr(tm), D, 0, W - 1, ptr(status)), "vc_v1_rec_delay")
    torch.cuda.synchronize()
    assert int(status.item()) == 1
