"""Pins of the oracle's far-connection input (orc_v1_rec_far: the
reference's recal_G_vec with distance-dependent delays and sum_G,
src/coredynamics.cu:181-372, 2314-2349): every spike of a far presynaptic
neuron reaches its post neuron d/speed after it was fired and, if it lands in
the current step, adds nsp*w*p of the conductance shape over the rest of the
step -- summed here in float64 from absolute arrival times, not from the
two-slot arithmetic; the ring's length does not matter; far lists that hold
exactly the near synapses, in recal_G_mat's order, give recal_G_mat's sums
bit for bit; and nothing is added in a step in which no far presynaptic
neuron's output is above 1.0 (src/patch.cu:6597-6598).  Also the synthetic
lists."""
import ctypes as C

import numpy as np
import pytest

from test_delay_oracle import delay_call, random_ring, synapses, zero_state
from test_oracle import osim


def fp(a):
    return a.ctypes.data_as(C.c_void_p)


def far_call(sim, st, far, train, depth, cur, speed):
    off, pre, w, dist = (np.ascontiguousarray(x) for x in far)
    train = np.ascontiguousarray(train, dtype=np.float32)
    return sim.o.L.orc_v1_rec_far(C.byref(sim.p), fp(st["gE"]), fp(st["hE"]), fp(st["gI"]),
                                  fp(st["hI"]), fp(off), fp(pre), fp(w), fp(dist), speed,
                                  fp(train), depth, cur)


def far_event_sum(net, far, train, depth, cur, speed):
    sp = net.spec
    dt = sp.dt
    off, pre, w, d = far
    post = np.repeat(np.arange(sp.nV1), np.diff(off.astype(np.int64)))
    pre = pre.astype(np.int64)
    w = w.astype(np.float64)
    d = d.astype(np.float64)
    bs = sp.blockSize
    inh = (pre % bs) >= sp.nE
    tpost = ((post % bs) >= sp.nE).astype(int)
    out = {k: np.zeros((n, sp.nV1)) for k, n in (("gE", len(sp.riseTimeE)),
                                                  ("hE", len(sp.riseTimeE)),
                                                  ("gI", len(sp.riseTimeI)),
                                                  ("hI", len(sp.riseTimeI)))}
    for k in range(depth):
        sv = train[k, pre].astype(np.float64)
        nsp = np.floor(sv)
        arrival = (sv - nsp - (cur - k) % depth) * dt + d / speed
        ok = (sv >= 1.0) & (arrival >= 0.0) & (arrival < dt)
        r = dt - arrival
        for isI, rise, decay, p, gk, hk in ((False, sp.riseTimeE, sp.decayTimeE, sp.pE, "gE", "hE"),
                                            (True, sp.riseTimeI, sp.decayTimeI, sp.pI, "gI", "hI")):
            sel = ok & (inh == isI)
            for ig in range(len(rise)):
                A = nsp[sel] * w[sel] * np.array([p[0][ig], p[1][ig]])[tpost[sel]]
                rr = r[sel]
                dod = rise[ig] / (decay[ig] - rise[ig])
                np.add.at(out[gk][ig], post[sel], A * dod * (np.exp(-rr / decay[ig]) -
                                                             np.exp(-rr / rise[ig])))
                np.add.at(out[hk][ig], post[sel], A * np.exp(-rr / rise[ig]))
    return out


def _far_net(speed, nFar=12, nblock=4):
    from visual_cortex_amd import network
    return network.build_network(network.spec_tiny(noDelay=0, speedOfThought=speed,
                                                   nblock=nblock, nFar=nFar),
                                 with_frames=False)


def far_depth(net):
    from visual_cortex_amd import network
    sp = net.spec
    return network.train_depth(float(net.far[3].max()), sp.speedOfThought, sp.dt)


@pytest.mark.parametrize("speed,cur", [(0.25, 5), (0.5, 0), (2.0, 1)])
def test_far_input_equals_arrival_time_sum(pkg, speed, cur):
    net = _far_net(speed)
    sp = net.spec
    depth = far_depth(net)
    assert depth >= 3
    cur = cur % depth
    rng = np.random.default_rng(21 + cur)
    train = random_ring(rng, depth, sp.nV1, p_spike=0.15)
    sim = osim(pkg, net)
    st = zero_state(net)
    assert far_call(sim, st, net.far, train, depth, cur, speed) == 0
    ref = far_event_sum(net, net.far, train, depth, cur, speed)
    hits = 0
    for k in ("gE", "hE", "gI", "hI"):
        n = ref[k].shape[0]
        np.testing.assert_allclose(st[k][:n].astype(np.float64), ref[k], rtol=3e-5,
                                   atol=1e-7 * max(1.0, ref[k].max()))
        hits += int((ref[k] > 0).sum())
    assert hits > 100


def test_far_ring_length_and_short_ring(pkg):
    net = _far_net(0.5)
    sp = net.spec
    depth = far_depth(net)
    rng = np.random.default_rng(8)
    hist = random_ring(rng, depth, sp.nV1, p_spike=0.15)
    sim = osim(pkg, net)
    outs = []
    for D, cur in ((depth, depth - 1), (depth, 1 % depth), (depth + 7, 3)):
        train = np.full((D, sp.nV1), -60.0, np.float32)
        for a in range(depth):
            train[(cur - a) % D] = hist[a]
        st = zero_state(net)
        assert far_call(sim, st, net.far, train, D, cur, 0.5) == 0
        outs.append(st)
    for st in outs[1:]:
        for k in ("gE", "hE", "gI", "hI"):
            assert np.array_equal(st[k].view(np.uint32), outs[0][k].view(np.uint32))
    st = zero_state(net)
    assert far_call(sim, st, net.far, hist[:1], 1, 0, 0.5) != 0


def test_near_synapses_as_far_lists_give_recal_G_mat_sums(pkg):
    """The near synapses written as far lists (per post neuron: neighbour
    block order, then presynaptic id, as recal_G_mat visits them) reproduce
    orc_v1_rec_delay's sums bit for bit from nonzero start values."""
    from visual_cortex_amd import network
    speed = 0.5
    net = network.build_network(network.spec_tiny(noDelay=0, speedOfThought=speed),
                                with_frames=False)
    sp = net.spec
    bs = sp.blockSize
    post, pre, w, _ = synapses(net)
    b = post // bs
    ib = np.concatenate([np.full(net.rec_pairs[bb * net.maxNear + i][0].size, i)
                         for bb in range(sp.nblock) for i in range(int(net.nNear[bb]))])
    D = net.dense_delaymat()
    dist = D[b, ib, pre % bs, post % bs]
    keep = w > 0
    post, pre, w, dist, ib = post[keep], pre[keep], w[keep], dist[keep], ib[keep]
    order = np.lexsort((pre, ib, post))
    cnt = np.bincount(post, minlength=sp.nV1)
    off = np.zeros(sp.nV1 + 1, np.uint32)
    off[1:] = np.cumsum(cnt)
    far = (off, pre[order].astype(np.uint32), w[order].astype(np.float32),
           dist[order].astype(np.float32))
    nnz = net.n_rec_synapses_csr()
    depth = network.train_depth(float(net.rec_distances()[:nnz].max()), speed, sp.dt)
    rng = np.random.default_rng(4)
    sim = osim(pkg, net)
    for cur in (0, depth - 1):
        train = random_ring(rng, depth, sp.nV1, p_spike=0.15)
        a = zero_state(net)
        for k in ("gE", "hE", "gI", "hI"):
            a[k][:] = rng.random(a[k].shape).astype(np.float32)
        bst = {k: v.copy() for k, v in a.items()}
        start = {k: v.copy() for k, v in a.items()}
        assert delay_call(sim, net, a, train, depth, cur, speed) == 0
        assert far_call(sim, bst, far, train, depth, cur, speed) == 0
        for k in ("gE", "hE", "gI", "hI"):
            assert np.array_equal(a[k].view(np.uint32), bst[k].view(np.uint32)), (cur, k)
            assert int((a[k] != start[k]).sum()) > 50, k


def test_synthetic_far_lists():
    from visual_cortex_amd import network
    net = _far_net(0.5, nFar=10, nblock=9)
    sp = net.spec
    off, pre, w, d = net.far
    assert net.n_far() == 10 * sp.nV1 and np.all(np.diff(off.astype(np.int64)) == 10)
    post = np.repeat(np.arange(sp.nV1), 10)
    pb, qb = pre // sp.blockSize, post // sp.blockSize
    for b in range(sp.nblock):                     # never from a near block
        near = set(int(x) for x in net.nbr[b, :net.nNear[b]]) | {b}
        assert not (set(pb[qb == b].tolist()) & near) or len(near) == sp.nblock
    assert np.all(w > 0)
    ref = np.hypot(net.pos[pre, 0] - net.pos[post, 0], net.pos[pre, 1] - net.pos[post, 1])
    np.testing.assert_allclose(d, ref, rtol=1e-6)
    # the rest of the network does not depend on nFar
    net0 = network.build_network(network.spec_tiny(noDelay=0, speedOfThought=0.5, nblock=9),
                                 with_frames=False)
    assert net0.far is None
    assert np.array_equal(net0.pos, net.pos)
    assert all(np.array_equal(a[2], b[2]) for a, b in zip(net0.rec_pairs, net.rec_pairs))
    # noFarDelay: -1 follows noDelay; 1 turns the far input off (the reference's
    # distance-independent branch adds nothing)
    assert net.far_delayed()
    from dataclasses import replace
    net.spec = replace(sp, noFarDelay=1)
    assert not net.far_delayed()


def test_far_gate_on_current_far_spikes(pkg):
    """The reference runs recal_G_vec only when a far presynaptic neuron's
    output in the current step is above 1.0 (fill_fSpikeTrain,
    src/stats.cu:114-126): with every far presynaptic output of the step
    below or at 1.0 nothing is added, however many earlier spikes arrive;
    one such output above 1.0 gives the full arrival-time sum."""
    net = _far_net(0.5, nFar=4, nblock=9)
    sp = net.spec
    depth = far_depth(net)
    isfar = np.zeros(sp.nV1, bool)
    isfar[net.far[1].astype(np.int64)] = True
    assert 0 < isfar.sum() < sp.nV1
    rng = np.random.default_rng(9)
    cur = depth // 2
    train = random_ring(rng, depth, sp.nV1, p_spike=0.2)
    train[cur] = np.where(isfar, np.float32(1.0), np.float32(1.6))
    sim = osim(pkg, net)
    st = zero_state(net)
    assert far_call(sim, st, net.far, train, depth, cur, 0.5) == 0
    assert all(float(np.abs(st[k]).max()) == 0.0 for k in ("gE", "hE", "gI", "hI"))
    ref = far_event_sum(net, net.far, train, depth, cur, 0.5)
    assert ref["gE"].max() > 0                       # arrivals that the gate drops
    train[cur][np.nonzero(isfar)[0][0]] = 1.5
    st = zero_state(net)
    assert far_call(sim, st, net.far, train, depth, cur, 0.5) == 0
    ref = far_event_sum(net, net.far, train, depth, cur, 0.5)
    for k in ("gE", "hE", "gI", "hI"):
        n = ref[k].shape[0]
        np.testing.assert_allclose(st[k][:n].astype(np.float64), ref[k], rtol=3e-5,
                                   atol=1e-7 * max(1.0, ref[k].max()))


def gate_splits(nets, sinfos):
    """Steps in which the far gates of the trials differ (src/patch.cu
    :6597-6598: some far presynaptic neuron of the trial has output > 1.0)
    while a trial whose gate is closed has far arrivals due in that step
    (a spike of a far presynaptic neuron it2post or it2post - 1 steps
    earlier, it2post = ceil(d/speed/dt)), which the closed gate drops: the
    steps in which a network-wide gate would give other sums."""
    opened, due = [], []
    for n, s in zip(nets, sinfos):
        sp = n.spec
        off, pre, _, dist = n.far
        T = s.shape[0]
        pre = pre.astype(np.int64)
        opened.append((s[:, np.unique(pre)] > 1.0).any(1))
        it2 = np.ceil(dist.astype(np.float64) / sp.speedOfThought / sp.dt).astype(np.int64)
        fired = s[:, pre] >= 1.0                          # [T, connections]
        d = np.zeros(T, bool)
        for lag in (0, 1):
            L = it2 - lag
            for t in range(T):
                src = t - L
                ok = (src >= 0) & (L >= 1)
                if ok.any() and fired[src[ok], np.flatnonzero(ok)].any():
                    d[t] = True
        due.append(d)
    n = 0
    for i in range(len(nets)):
        for j in range(len(nets)):
            if i != j:
                n += int((opened[i] & ~opened[j] & due[j]).sum())
    return n


# far input through delayed near connections, and with undelayed ones
FAR_BATCH = {"delayed": dict(noDelay=0, speedOfThought=0.5),
             "far_only": dict(noDelay=1, noFarDelay=0, speedOfThought=0.5)}


@pytest.mark.parametrize("kind", list(FAR_BATCH))
def test_batched_far_trials_equal_serial(pkg, kind):
    """Trials with far lists batched into one network: each trial's output
    equals its own run bit for bit -- the far gate is per trial, as each
    trial is a separate run of the reference (the second trial fires far
    less, so in many steps only the first trial's gate is open while the
    second has far arrivals that its closed gate drops)."""
    import oracle
    from visual_cortex_amd import batch, network
    specs = [network.spec_tiny(nblock=nb, seed=300 + i, ori_deg=30.0 * i,
                               bgStrength=(0.4, 0.4),
                               nFar=10, **FAR_BATCH[kind])
             for i, nb in enumerate((4, 2))]
    nets = [network.build_network(s) for s in specs]
    b = batch.batch_networks(nets, device="cpu")
    assert b.n_far() == sum(n.n_far() for n in nets)
    ob = oracle.OracleSim(b, pkg)
    ob.run(96)
    sb = ob.sInfo()
    off = 0
    serial = []
    for n in nets:
        o = oracle.OracleSim(n, pkg)
        o.run(96)
        s = o.sInfo()
        serial.append(s)
        assert int((s >= 1).sum()) > 40
        assert np.array_equal(sb[:, off:off + n.spec.nV1].view(np.uint32), s.view(np.uint32))
        off += n.spec.nV1
    assert gate_splits(nets, serial) > 0


def test_batch_refuses_mixed_far_settings(pkg):
    """Trials batched together share one spec for the far gate: a trial whose
    far connections are delayed (noDelay 1, noFarDelay 0) cannot be batched
    with one whose far input is off (noFarDelay -1 -> noDelay 1)."""
    from visual_cortex_amd import batch, network
    a = network.build_network(network.spec_tiny(nFar=4, noDelay=1, noFarDelay=0, nt=64))
    b = network.build_network(network.spec_tiny(nFar=4, noDelay=1, noFarDelay=-1, nt=64, seed=7))
    assert a.far_delayed() and not b.far_delayed()
    for pair in ((a, b), (b, a)):
        with pytest.raises(ValueError, match="noFarDelay"):
            batch.batch_networks(list(pair), device="cpu")
