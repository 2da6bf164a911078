"""Gap junctions between inhibitory neurons (the reference's InhGap): the
oracle's gap sums and membrane terms, and the host-side data (files,
gapRatio, gapS), pinned by what the mathematics fixes.

  * orc_v1_gap against a float64 matrix product over the reference's
    V1_gapMat layout [b][ib][pre][post] (asymmetric strengths, so a
    transposed operand fails), with spiking partners replaced by their
    type's vThres;
  * the delayed form against the same product over each junction's lagged
    ring slot;
  * the steady state of a network with gap junctions and tonic drive only:
    at a fixed point of the step (Heun's method with the gap sum of the step
    before) every inhibitory voltage solves
        (gL + gapS_i) v_i - sum_j g_ij v_j = gL vL + depC_i,
    a linear system solved here in float64;
  * gapS = far strengths then near strengths in the reference's order, the
    files' strengths times gapRatio, and a save/load round trip."""
import numpy as np
import pytest


def _gap_net(pkg, **kw):
    from visual_cortex_amd import network as N
    spec = N.spec_tiny(nGap=6, InhGap=1, **kw)
    return N.build_network(spec)


def _vhat(train_row, p, nV1, bs):
    """Each neuron's voltage as a gap partner: vThres of its type when it spiked."""
    local = np.arange(nV1) % bs
    acc = [p.typeAcc[t] for t in range(p.nType)]
    typ = np.searchsorted(np.asarray(acc), local, side="right")
    vth = np.asarray([p.vThres[t] for t in range(p.nType)], np.float64)
    v = train_row.astype(np.float64)
    return np.where(v > 0, vth[typ], v)


def _oracle_gap(orc, net, p, train, cur, gapDist=None, speed=1.0):
    import ctypes as C
    import oracle as O
    sp = net.spec
    G = np.ascontiguousarray(net.dense_gapmat())
    out = np.zeros(sp.nblock * sp.nI, np.float32)
    nNear = np.ascontiguousarray(net.nNear, np.uint32)
    nbr = np.ascontiguousarray(net.nbr, np.uint32)
    train = np.ascontiguousarray(train, np.float32)
    rc = orc.L.orc_v1_gap(C.byref(p), net.maxNear, O._fp(nNear), O._fp(nbr), O._fp(G),
                          O._fp(gapDist), speed, O._fp(train), train.shape[0], cur, O._fp(out))
    assert rc == 0
    return out


def test_gap_sum_is_the_matrix_product(pkg, orc):
    from visual_cortex_amd import params as P
    net = _gap_net(pkg)
    sp = net.spec
    bs, nE, nI = sp.blockSize, sp.nE, sp.nI
    p = P.v1_params_struct(net)
    p.vThres[0], p.vThres[1] = -20.0, -47.5      # a partner takes its own type's threshold
    rng = np.random.default_rng(5)
    train = rng.uniform(-70.0, -52.0, (1, sp.nV1)).astype(np.float32)
    spk = rng.random(sp.nV1) < 0.2
    train[0, spk] = (1.0 + rng.random(int(spk.sum()))).astype(np.float32)
    got = _oracle_gap(orc, net, p, train, 0)
    G = net.dense_gapmat().astype(np.float64)            # [b, ib, pre l, post m]
    assert not np.allclose(G, G.transpose(0, 1, 3, 2))   # asymmetric: a transpose shows
    vh = _vhat(train[0], p, sp.nV1, bs)
    want = np.zeros((sp.nblock, nI))
    for b in range(sp.nblock):
        for ib in range(int(net.nNear[b])):
            bid = int(net.nbr[b, ib])
            want[b] += G[b, ib].T @ vh[bid * bs + nE: (bid + 1) * bs]
    scale = np.zeros((sp.nblock, nI))
    for b in range(sp.nblock):
        for ib in range(int(net.nNear[b])):
            bid = int(net.nbr[b, ib])
            scale[b] += np.abs(G[b, ib]).T @ np.abs(vh[bid * bs + nE: (bid + 1) * bs])
    err = np.abs(got.reshape(sp.nblock, nI) - want)
    assert (err <= 1e-6 * scale + 1e-6).all(), err.max()
    # the spike substitution matters here: with raw outputs the sums differ
    raw = np.zeros_like(want)
    for b in range(sp.nblock):
        for ib in range(int(net.nNear[b])):
            bid = int(net.nbr[b, ib])
            raw[b] += G[b, ib].T @ train[0, bid * bs + nE: (bid + 1) * bs].astype(np.float64)
    assert np.abs(raw - want).max() > 1.0


def test_gap_delayed_reads_each_junctions_lagged_slot(pkg, orc):
    from visual_cortex_amd import params as P
    net = _gap_net(pkg)
    sp = net.spec
    bs, nE, nI = sp.blockSize, sp.nE, sp.nI
    p = P.v1_params_struct(net)
    speed = 0.25                                          # mm/ms: lags of several steps
    D = np.ascontiguousarray(net.dense_gapdist())
    f = np.float32
    lag = np.ceil((D / f(speed)) / f(sp.dt)).astype(np.int64)
    depth = int(lag.max()) + 2
    rng = np.random.default_rng(9)
    train = rng.uniform(-70.0, -52.0, (depth, sp.nV1)).astype(np.float32)
    spk = rng.random((depth, sp.nV1)) < 0.15
    train[spk] = 1.5
    cur = 3
    got = _oracle_gap(orc, net, p, train, cur, gapDist=D, speed=speed)
    G = net.dense_gapmat().astype(np.float64)
    want = np.zeros((sp.nblock, nI))
    for b in range(sp.nblock):
        for ib in range(int(net.nNear[b])):
            bid = int(net.nbr[b, ib])
            for l, m in zip(*np.nonzero(G[b, ib])):
                slot = (cur - lag[b, ib, l, m]) % depth
                vh = _vhat(train[slot], p, sp.nV1, bs)[bid * bs + nE + l]
                want[b, m] += G[b, ib, l, m] * vh
    assert lag.max() >= 3
    np.testing.assert_allclose(got.reshape(sp.nblock, nI), want, rtol=1e-5, atol=1e-4)


def test_gap_steady_state_solves_the_linear_system(pkg, orc):
    """No chemical synapses, no LGN input, no background: only the tonic
    drive (below threshold) and the gap junctions act.  After 3000 steps
    (375 ms, ~70 membrane time constants) the oracle's voltages are the
    fixed point, which the float64 solution of the linear system gives."""
    import oracle as O
    from visual_cortex_amd import network as N
    spec = N.lif(N.spec_tiny(nGap=6, InhGap=1, KE=0, KI=0, m=0, bgStrength=(0.0, 0.0),
                             tonicDep=(0.5, 0.5)))
    net = N.build_network(spec)
    rng = np.random.default_rng(21)
    net.tonicDep = rng.uniform(0.2, 0.8, spec.nV1).astype(np.float32)
    sim = O.OracleSim(net, pkg)
    sim.run(3000, record=False)
    v = sim.st["v"].astype(np.float64)
    bs, nE, nI = spec.blockSize, spec.nE, spec.nI
    vL, vT, gL = -65.0, spec.vT[1], spec.gL[1]
    target = vL + (vT - vL) * net.tonicDep.astype(np.float64)
    depC = gL * (target - vL)                              # LIF tonic drive (a = 0)
    mI = spec.nblock * nI
    A = np.zeros((mI, mI))
    gapS = net.gap_strength_sums().astype(np.float64)
    Gd = net.dense_gapmat().astype(np.float64)
    for b in range(spec.nblock):
        for ib in range(int(net.nNear[b])):
            bid = int(net.nbr[b, ib])
            A[b * nI:(b + 1) * nI, bid * nI:(bid + 1) * nI] -= Gd[b, ib].T
    A[np.arange(mI), np.arange(mI)] += gL + gapS
    isI = (np.arange(spec.nV1) % bs) >= nE
    rhs = gL * vL + depC[isI]
    vI = np.linalg.solve(A, rhs)
    assert np.abs(gapS).max() > 0.5 and (v < -50.0).all()
    np.testing.assert_allclose(v[isI], vI, atol=2e-4)
    np.testing.assert_allclose(v[~isI], target[~isI], atol=2e-4)
    # the gap terms move the inhibitory fixed point away from the tonic target
    assert np.abs(vI - target[isI]).max() > 0.5


def test_gap_strength_sums_order_ratio_and_files(pkg, tmp_path):
    from visual_cortex_amd import config, formats as F, loader
    net = _gap_net(pkg)
    sp = net.spec
    mI = sp.nblock * sp.nI
    rng = np.random.default_rng(3)
    # far junctions: two per inhibitory neuron, from the other block
    cnt = np.full(mI, 2, np.int64)
    off = np.zeros(mI + 1, np.int64)
    np.cumsum(cnt, out=off[1:])
    post_b = np.repeat(np.arange(mI) // sp.nI, 2)
    pre = ((1 - post_b) * sp.blockSize + sp.nE + rng.integers(0, sp.nI, 2 * mI)).astype(np.uint32)
    g = rng.uniform(0.01, 0.3, 2 * mI).astype(np.float32)
    dist = rng.uniform(0.1, 0.5, 2 * mI).astype(np.float32)
    net.gap_far = (off.astype(np.uint32), pre, g, dist)
    # gapS: 0, + far strengths in file order, + near ones by (slot, partner), in float32
    gs = net.gap_strength_sums()
    want = np.zeros(mI, np.float32)
    for i in range(mI):
        a = np.float32(0.0)
        for j in range(int(off[i]), int(off[i + 1])):
            a = np.float32(a + g[j])
        b = i // sp.nI
        for ib in range(int(net.nNear[b])):
            P_, Q_, W_ = net.gap[b * net.maxNear + ib]
            sel = Q_ == (i % sp.nI)
            for w in W_[sel][np.argsort(P_[sel], kind="stable")]:
                a = np.float32(a + w)
        want[i] = a
    assert np.array_equal(gs.view(np.uint32), want.view(np.uint32))
    # files: strengths as they are; the loader multiplies by gapRatio
    cfg_path = loader.save_network(net, tmp_path, suffix="gp")
    text = cfg_path.read_text().replace("gapRatio = 1.0", "gapRatio = 0.5")
    cfg_path.write_text(text)
    net2 = loader.load_network(config.Config.from_file(str(cfg_path)))
    assert net2.spec.InhGap == 1
    for a, b in zip(net.gap, net2.gap):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        assert np.array_equal(b[2], (a[2] * np.float32(0.5)).astype(np.float32))
    assert np.array_equal(net2.gap_far[2], (g * np.float32(0.5)).astype(np.float32))
    assert np.array_equal(net2.gap_far[1], pre) and np.array_equal(net2.gap_far[3], dist)
    # a negative strength in the file is refused, as the reference asserts
    mpath = tmp_path / "V1_gapMat-gp.bin"
    raw = np.fromfile(mpath, np.uint8).copy()
    raw[4:8] = np.frombuffer(np.float32(-1.0).tobytes(), np.uint8)
    raw.tofile(mpath)
    with pytest.raises(ValueError):
        loader.load_network(config.Config.from_file(str(cfg_path)))


def test_gap_files_need_gap_ratio(pkg, tmp_path):
    from visual_cortex_amd import config, loader
    net = _gap_net(pkg)
    cfg_path = loader.save_network(net, tmp_path, suffix="gr")
    cfg_path.write_text(cfg_path.read_text().replace("gapRatio = 1.0\n", ""))
    with pytest.raises(ValueError, match="gapRatio"):
        loader.load_network(config.Config.from_file(str(cfg_path)))


def test_gap_ell_order_and_lags(pkg):
    """The ELL lists each inhibitory neuron's junctions by (slot, partner),
    with the partner's global id, and the lags are ceil((d / speed) / dt)."""
    net = _gap_net(pkg)
    sp = net.spec
    mg, cnt, Pm, Gm = net.gap_ell()
    G = net.dense_gapmat()
    for i in range(0, sp.nblock * sp.nI, 7):
        b, m = divmod(i, sp.nI)
        want = []
        for ib in range(int(net.nNear[b])):
            bid = int(net.nbr[b, ib])
            for l in np.nonzero(G[b, ib, :, m])[0]:
                want.append((bid * sp.blockSize + sp.nE + int(l), G[b, ib, l, m]))
        n = int(cnt[i])
        assert n == len(want)
        assert [int(x) for x in Pm[:n, i]] == [w[0] for w in want]
        assert np.array_equal(Gm[:n, i], np.asarray([w[1] for w in want], np.float32))
    lag = net.gap_lags(0.5, sp.dt)
    pos = net.pos
    for i in range(0, sp.nblock * sp.nI, 11):
        b = i // sp.nI
        post = b * sp.blockSize + sp.nE + i % sp.nI
        for j in range(int(cnt[i])):
            pre = int(Pm[j, i])
            x = np.float32(pos[pre, 0] - pos[post, 0])
            y = np.float32(pos[pre, 1] - pos[post, 1])
            d = np.float32(np.sqrt(np.float32(x * x + y * y)))
            assert lag[j, i] == int(np.ceil(np.float32(d / np.float32(0.5)) / np.float32(sp.dt)))
