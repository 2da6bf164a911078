"""Pins of the oracle's voltage noise and synaptic failure (the reference's
noisyDep / tau_noise / noiseOnTonic, src/coredynamics.cu:1424-1445, and
synFailFF / synFail, :1476-1480, 2766-2769) against what the mathematics
fixes:

* the normal draw: cos(2 pi x) and the Box-Muller transform against float64
  on the same 32-bit inputs, and the draws of the noise stream against the
  standard normal (Kolmogorov-Smirnov, moments, the largest value the 24-bit
  uniforms allow);
* white noise: on a linear membrane (LIF, no conductances) one Heun step
  moves v by sigma N (1 - gL dt / 2C) relative to the noiseless step, with
  sigma = sqrt(2 noisyDep (vT - vR) r dt), r = tonicDep or 1; the dep output
  is depC + sigma N;
* Ornstein-Uhlenbeck noise: one update is y' = y (1 - dt/tau + dt^2/2tau^2)
  + (sigma N / tau)(1 - dt/2tau) (Heun's method on dy/dt = -y/tau with the
  kick), v moves by dt/2 (y + y' - gL dt y / C), and over many neurons the
  stationary variance and lag-1 correlation are those of that recursion;
* failure: an input's conductance increment is scaled by 1 - p +
  sqrt(p (1 - p) / nSyn) N for the draw of (post neuron, step, stream,
  synapse), p = 1 removes it, p = 0 leaves every bit as it was."""
import ctypes as C

import numpy as np
import pytest

from test_oracle import flat_state, hand_network, osim

STREAM_FAILFF, STREAM_FAIL, STREAM_NOISE = 5, 6, 7


def fp(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def step_once(sim, lgn_row, it, do_rec=1):
    """One orc_v1_step of `sim` on its own state; returns (sInfo, dep)."""
    sp = sim.net.spec
    out = np.zeros(sp.nV1, np.float32)
    dep = np.zeros(sp.nV1, np.float32)
    st = sim.st
    rc = sim.o.L.orc_v1_step(C.byref(sim.p), fp(st["v"]), fp(st["w"]), fp(st["tBack"]),
                             fp(st["gFF"]), fp(st["hFF"]), fp(st["gE"]), fp(st["hE"]),
                             fp(st["gI"]), fp(st["hI"]), fp(sim.tonic), fp(sim.nLGN_of),
                             fp(sim.sLGN), fp(sim.idLGN),
                             fp(np.ascontiguousarray(lgn_row, np.float32)), sim.net.maxNear,
                             fp(sim.nNear), fp(sim.nbr), fp(sim.conMat), fp(sim.ipre),
                             fp(sim.npre), fp(sim.og), fp(sim.oh), it, fp(out), fp(dep), do_rec,
                             None, None, fp(sim.lastNoise))
    assert rc == 0
    return out, dep


def draws(o, p, ents, step, stream, packs):
    """The normals of (entity, step, stream, pack) under the run's key, float64."""
    return np.array([o.L.orc_draw_normal(p.seed_lo, p.seed_hi, int(e), step, stream, int(k),
                                         p.trial) for e, k in zip(ents, packs)], np.float64)


# ------------------------------------------------------------ the draw --
def test_cos2pi_against_float64(orc):
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.random(200000), np.arange(0, 1, 1 / 64),
                        np.array([0.125, 0.25, 0.375, 0.5, 0.625, 0.75, 0.875])])
    x = np.unique(x.astype(np.float32))
    x = x[x < 1.0]
    got = np.zeros_like(x)
    orc.L.orc_cos2pi_vec(x.size, fp(x), fp(got))
    ref = np.cos(2 * np.pi * x.astype(np.float64))
    err = np.abs(got.astype(np.float64) - ref)
    # a few float32 ulp of the result, and of 1 near the zeros of cos
    assert np.all(err <= 3 * np.spacing(np.abs(ref).astype(np.float32)) + 6e-8), err.max()
    assert got[x == 0][0] == 1.0 and got[x == np.float32(0.5)][0] == -1.0
    assert got[x == np.float32(0.25)][0] == 0.0


def test_box_muller_against_float64(orc):
    rng = np.random.default_rng(4)
    a = rng.integers(0, 2 ** 32, 100000, dtype=np.uint32)
    b = rng.integers(0, 2 ** 32, 100000, dtype=np.uint32)
    a[:4] = [0, 0xffffffff, 255, 256]                  # u1 = 2^-24, 1, 2^-24, 2 * 2^-24
    got = np.zeros(a.size, np.float32)
    orc.L.orc_normal_vec(a.size, fp(a), fp(b), fp(got))
    u1 = ((a >> 8).astype(np.float64) + 1.0) * 2.0 ** -24
    u2 = (b >> 8).astype(np.float64) * 2.0 ** -24
    ref = np.sqrt(-2.0 * np.log(u1)) * np.cos(2 * np.pi * u2)
    np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-6)
    assert got[1] == 0.0                                # u1 = 1: radius 0


def test_noise_stream_is_standard_normal(orc):
    from scipy import stats
    n = 60000
    z = np.array([orc.L.orc_draw_normal(0x1234, 0x5678, e, 77, STREAM_NOISE, 0, 0)
                  for e in range(n)])
    assert stats.kstest(z, "norm").pvalue > 1e-3
    assert abs(z.mean()) < 4.0 / np.sqrt(n)
    assert abs(z.var() - 1.0) < 4.0 * np.sqrt(2.0 / n)
    assert np.abs(z).max() <= np.sqrt(-2 * np.log(2.0 ** -24)) + 1e-5
    # another step, stream or pack is another draw
    z2 = np.array([orc.L.orc_draw_normal(0x1234, 0x5678, e, 78, STREAM_NOISE, 0, 0)
                   for e in range(2000)])
    z3 = np.array([orc.L.orc_draw_normal(0x1234, 0x5678, e, 77, STREAM_FAIL, 0, 0)
                   for e in range(2000)])
    z4 = np.array([orc.L.orc_draw_normal(0x1234, 0x5678, e, 77, STREAM_NOISE, 1, 0)
                   for e in range(2000)])
    for other in (z2, z3, z4):
        assert abs(np.corrcoef(z[:2000], other)[0, 1]) < 0.1
        assert not np.array_equal(z[:2000], other)


# ------------------------------------------------------- voltage noise --
def linear_net(pkg, **kw):
    base = dict(bgStrength=(0.0, 0.0), deltaT=(0.0, 0.0), a=(0.0, 0.0), b=(0.0, 0.0),
                vThres=(-20.0, -20.0))
    base.update(kw)
    return hand_network(pkg, nblock=2, bs=64, nE=48, **base)


@pytest.mark.parametrize("on_tonic", [0, 1])
def test_white_noise_one_step(pkg, on_tonic):
    net = linear_net(pkg, noisyDep=(0.05, 0.12), noiseOnTonic=on_tonic, tonicDep=(0.6, 0.9))
    clean = linear_net(pkg, noisyDep=(0.0, 0.0), noiseOnTonic=on_tonic, tonicDep=(0.6, 0.9))
    sp = net.spec
    rng = np.random.default_rng(5)
    st = flat_state(net, -60.0)
    st["v"][:] = rng.uniform(-64.0, -56.0, sp.nV1).astype(np.float32)
    a, b = osim(pkg, net, st), osim(pkg, clean, st)
    it = 37
    lgn = np.zeros(sp.nLGN, np.float32)
    _, dep_a = step_once(a, lgn, it)
    _, dep_b = step_once(b, lgn, it)
    t = (np.arange(sp.nV1) % sp.blockSize >= sp.nE).astype(int)
    nd = np.array(sp.noisyDep)[t]
    vT, vR = np.array(sp.vT)[t], np.array(sp.vR)[t]
    r = a.tonic.astype(np.float64) if on_tonic else 1.0
    N = draws(a.o, a.p, np.arange(sp.nV1), it, STREAM_NOISE, np.zeros(sp.nV1))
    sigma = np.sqrt(2 * nd * (vT - vR) * r * sp.dt) * N
    gL, Cm = np.array(sp.gL)[t], np.array(sp.C)[t]
    dv = a.st["v"].astype(np.float64) - b.st["v"].astype(np.float64)
    np.testing.assert_allclose(dv, sigma * (1 - gL * sp.dt / (2 * Cm)), rtol=0, atol=2e-5)
    np.testing.assert_allclose(dep_a.astype(np.float64) - dep_b, sigma, rtol=0, atol=2e-6)
    assert np.abs(sigma).max() > 0.1
    assert not a.lastNoise.any()                        # white noise keeps no state


def test_ou_noise_one_step(pkg):
    tau = 2.5
    net = linear_net(pkg, noisyDep=(0.05, 0.12), noiseOnTonic=0, tau_noise=tau)
    clean = linear_net(pkg, noisyDep=(0.0, 0.0), noiseOnTonic=0, tau_noise=tau)
    sp = net.spec
    rng = np.random.default_rng(6)
    st = flat_state(net, -60.0)
    st["v"][:] = rng.uniform(-64.0, -56.0, sp.nV1).astype(np.float32)
    a, b = osim(pkg, net, st), osim(pkg, clean, st)
    y0 = rng.normal(0.0, 0.2, sp.nV1).astype(np.float32)
    a.lastNoise[:] = y0
    b.lastNoise[:] = y0               # the clean run: no kick, y is not updated (noise1 = 0)
    it = 11
    lgn = np.zeros(sp.nLGN, np.float32)
    _, dep_a = step_once(a, lgn, it)
    _, dep_b = step_once(b, lgn, it)
    t = (np.arange(sp.nV1) % sp.blockSize >= sp.nE).astype(int)
    nd = np.array(sp.noisyDep)[t]
    vT, vR = np.array(sp.vT)[t], np.array(sp.vR)[t]
    N = draws(a.o, a.p, np.arange(sp.nV1), it, STREAM_NOISE, np.zeros(sp.nV1))
    x = np.sqrt(2 * nd * (vT - vR) * sp.dt) * N
    dt = sp.dt
    y = y0.astype(np.float64)
    y1 = y * (1 - dt / tau + dt * dt / (2 * tau * tau)) + (x / tau) * (1 - dt / (2 * tau))
    np.testing.assert_allclose(a.lastNoise, y1, rtol=1e-5, atol=1e-7)
    np.testing.assert_array_equal(b.lastNoise, y0)
    gL, Cm = np.array(sp.gL)[t], np.array(sp.C)[t]
    # both runs take y as the first slope's term; the noisy one adds y1 to the second
    dv = a.st["v"].astype(np.float64) - b.st["v"].astype(np.float64)
    np.testing.assert_allclose(dv, dt / 2 * y1, rtol=0, atol=2e-5)
    np.testing.assert_allclose(dep_a.astype(np.float64) - dep_b, y1, rtol=0, atol=2e-6)


def test_ou_noise_stationary_statistics(pkg):
    """Many steps of the recursion y' = A y + B N: variance B^2 / (1 - A^2),
    lag-1 correlation A, over 8 blocks x 64 neurons of one type."""
    tau, nd = 2.0, 0.08
    net = hand_network(pkg, nblock=8, bs=64, nE=64, bgStrength=(0.0, 0.0),
                       deltaT=(0.0, 0.0), a=(0.0, 0.0), b=(0.0, 0.0), vThres=(-20.0, -20.0),
                       noisyDep=(nd, nd), noiseOnTonic=0, tau_noise=tau)
    sp = net.spec
    sim = osim(pkg, net, flat_state(net, -62.0))
    lgn = np.zeros(sp.nLGN, np.float32)
    dt = sp.dt
    A = 1 - dt / tau + dt * dt / (2 * tau * tau)
    B = np.sqrt(2 * nd * (sp.vT[0] - sp.vR[0]) * dt) / tau * (1 - dt / (2 * tau))
    ys = []
    for it in range(400):
        step_once(sim, lgn, it, do_rec=0)
        ys.append(sim.lastNoise.astype(np.float64).copy())
    ys = np.array(ys[100:])                             # 6 tau of burn-in
    var = B * B / (1 - A * A)
    assert abs(ys.var() / var - 1) < 0.06, (ys.var(), var)
    rho = np.mean(ys[1:] * ys[:-1]) / np.mean(ys * ys)
    assert abs(rho - A) < 0.01, (rho, A)


# ---------------------------------------------------- synaptic failure --
def ff_net(pkg, **kw):
    from visual_cortex_amd import network
    return network.build_network(network.spec_tiny(bgStrength=(0.0, 0.0), **kw))


@pytest.mark.parametrize("pf", [0.3, 1.0])
def test_ff_failure_scales_each_input(pkg, pf):
    """One spiking LGN cell: every neuron that lists it once gets its FF
    conductance increment times the failure factor of (neuron, step,
    list slot); neurons without it are unchanged."""
    nsyn = (40.0, 25.0)
    net = ff_net(pkg, synFailFF=(pf, pf), synPerConFF=nsyn)
    clean = ff_net(pkg)
    sp = net.spec
    st = flat_state(net, -65.0)
    a, b = osim(pkg, net, st), osim(pkg, clean, st)
    cnt = np.bincount(a.idLGN[:, :].ravel()[
        (np.arange(a.idLGN.shape[0])[:, None] < a.nLGN_of[None, :]).ravel()], minlength=sp.nLGN)
    cell = int(np.argmax(cnt))
    lgn = np.zeros(sp.nLGN, np.float32)
    lgn[cell] = 1.4                                    # one spike at 0.4 dt
    it = 21
    step_once(a, lgn, it, do_rec=0)
    step_once(b, lgn, it, do_rec=0)
    slots = np.arange(a.idLGN.shape[0])[:, None]
    hit = (a.idLGN == cell) & (slots < a.nLGN_of[None, :])
    assert hit.sum(axis=0).max() == 1
    posts = np.nonzero(hit.any(axis=0))[0]
    slot = np.argmax(hit[:, posts], axis=0)
    t = (posts % sp.blockSize >= sp.nE).astype(int)
    N = draws(a.o, a.p, posts, it, STREAM_FAILFF, slot)
    fac = 1 - pf + np.sqrt(pf * (1 - pf) / np.array(nsyn)[t]) * N
    fac = np.maximum(fac, 0.0)                          # a negative strength adds nothing
    for k in ("gFF", "hFF"):
        ga, gb = a.st[k].astype(np.float64), b.st[k].astype(np.float64)
        np.testing.assert_allclose(ga[:, posts], gb[:, posts] * fac[None, :], rtol=2e-6,
                                   atol=1e-12)
        others = np.setdiff1d(np.arange(sp.nV1), posts)
        assert np.array_equal(ga[:, others], gb[:, others])
    assert (b.st["gFF"][:, posts] > 0).all()
    if pf == 1.0:
        assert not a.st["gFF"][:, posts].any()


def test_no_failure_changes_nothing(pkg):
    """p = 0 with any synapses per connection is the run without failure,
    bit for bit (the draws are skipped)."""
    a = osim(pkg, ff_net(pkg, synFailFF=(0.0, 0.0), synPerConFF=(3.0, 3.0),
                         synFail=(0.0,) * 4, synPerCon=(2.0,) * 4))
    b = osim(pkg, ff_net(pkg))
    a.run(64)
    b.run(64)
    assert np.array_equal(a.sInfo().view(np.uint32), b.sInfo().view(np.uint32))


@pytest.mark.parametrize("inh", [False, True])
def test_recurrent_failure_scales_each_synapse(pkg, inh):
    """One neuron spikes: every post neuron's recurrent increment is the
    synapse's strength times the failure factor of (post, step, pre), with
    the rate and synapses per connection of [pre type][post type]."""
    pfail = (0.2, 0.35, 0.5, 0.65)
    nsyn = (5.0, 7.5, 4.26, 7.88)
    net = ff_net(pkg, synFail=pfail, synPerCon=nsyn)
    clean = ff_net(pkg)
    sp = net.spec
    st = flat_state(net, -70.0)
    pre = sp.blockSize + (sp.nE + 5 if inh else 17)     # in block 1
    st["v"][pre] = 0.0 if not inh else -40.0            # above threshold: spikes this step
    a, b = osim(pkg, net, st), osim(pkg, clean, st)
    lgn = np.zeros(sp.nLGN, np.float32)
    it = 9
    out_a, _ = step_once(a, lgn, it)
    out_b, _ = step_once(b, lgn, it)
    assert np.array_equal(np.nonzero(out_a >= 1)[0], [pre])
    key = "gI" if inh else "gE"
    hk = "hI" if inh else "hE"
    col = b.conMat                                     # [nblock][maxNear][bs][bs]
    bs = sp.blockSize
    posts, w = [], []
    for bb in range(sp.nblock):
        for ib in range(int(net.nNear[bb])):
            if int(net.nbr[bb, ib]) == pre // bs:
                row = col[bb, ib, pre % bs]
                q = np.nonzero(row > 0)[0]
                posts.append(bb * bs + q)
                w.append(row[q])
    posts = np.concatenate(posts)
    assert posts.size > 10
    tpost = (posts % bs >= sp.nE).astype(int)
    tpre = 1 if inh else 0
    p_ = np.array(pfail)[tpre * 2 + tpost]
    ns = np.array(nsyn)[tpre * 2 + tpost]
    N = draws(a.o, a.p, posts, it, STREAM_FAIL, np.full(posts.size, pre))
    fac = np.maximum(1 - p_ + np.sqrt(p_ * (1 - p_) / ns) * N, 0.0)
    # the increments: the state after the step minus the state without the
    # spike's input (the clean run's increments are the strengths' alone)
    base = osim(pkg, clean, st)
    base_out, _ = step_once(base, lgn, it, do_rec=0)
    for k in (key, hk):
        inc_a = a.st[k].astype(np.float64) - base.st[k]
        inc_b = b.st[k].astype(np.float64) - base.st[k]
        np.testing.assert_allclose(inc_a[:, posts], inc_b[:, posts] * fac[None, :], rtol=1e-5,
                                   atol=1e-9)
        assert (inc_b[0, posts] > 0).all()
