"""The initial state of a run.

* vc_v1_initial_state (host code of the library) draws the reference's
  initial v, w and conductances (src/patch.cu:4405-4570) with the standard
  library's generator.  It is checked bit for bit against a from-scratch
  restatement, here, of the published algorithms it relies on: the minimal
  standard linear congruential engine (x <- 16807 x mod 2^31 - 1, the
  default_random_engine of libstdc++), generate_canonical<float, 24> (one
  engine value, (x - 1) / 2^31 in binary32), the uniform and the polar
  (Marsaglia) normal distributions with the second value of each pair kept
  for the next draw -- in the reference's fill order.  A draw order mistake
  (g and h per neuron) or a normal object per array (no carried value)
  fails it.
* rand_spInit's restatement (orc_v1_initial_spikes) is pinned by what the
  rules fix: the spike fraction of each chance class, uniform spike times,
  the state of a neuron that spiked, and the order of the chain (reference
  slot 0, then slots depthRef-1, ..., 1) through the refractory rule.
"""
import ctypes as C

import numpy as np
import pytest
from scipy import stats

from test_multi_oracle import _fp, _params

LIBM = C.CDLL("libm.so.6")
LIBM.logf.argtypes = [C.c_float]
LIBM.logf.restype = C.c_float
LIBM.sqrtf.argtypes = [C.c_float]
LIBM.sqrtf.restype = C.c_float
f32 = np.float32


class MinStd:
    """linear_congruential_engine<uint_fast32_t, 16807, 0, 2147483647>."""
    M = 2147483647

    def __init__(self, seed):
        x = seed % self.M
        self.x = 1 if x == 0 else x

    def __call__(self):
        self.x = (self.x * 16807) % self.M
        return self.x

    def canonical(self):
        # generate_canonical<float, 24>: one value, (x - min) * 1 / float(2^31 - 2 + 0 ...)
        s = f32(self() - 1)
        r = s / f32(2147483646.0)          # float(2147483646.0L) = 2^31
        if r >= f32(1.0):
            r = np.nextafter(f32(1.0), f32(0.0))
        return f32(r)


class Normal:
    """normal_distribution<float>(0, 1): Marsaglia's polar method."""

    def __init__(self):
        self.saved = None

    def __call__(self, g):
        if self.saved is not None:
            r, self.saved = self.saved, None
            return r
        while True:
            x = f32(f32(2.0) * g.canonical() - f32(1.0))
            y = f32(f32(2.0) * g.canonical() - f32(1.0))
            r2 = f32(f32(x * x) + f32(y * y))
            if not (r2 > 1.0 or r2 == 0.0):
                break
        mult = f32(LIBM.sqrtf(f32(f32(-2.0) * f32(LIBM.logf(r2))) / r2))
        self.saved = f32(x * mult)
        return f32(y * mult)


def restated_initial_state(seed, sp, tonic, nLGN, per_array_normal=False, interleaved=False):
    g = MinStd(seed)
    norm = Normal()
    n, bs = sp.nV1, sp.blockSize
    acc0 = [0, sp.nE, bs]
    ngFF, ngE, ngI = len(sp.riseTimeFF), len(sp.riseTimeE), len(sp.riseTimeI)
    out = dict(v=np.zeros(n, f32), w=np.zeros(n, f32), gFF=np.zeros((ngFF, n), f32),
               hFF=np.zeros((ngFF, n), f32), gE=np.zeros((ngE, n), f32), hE=np.zeros((ngE, n), f32),
               gI=np.zeros((ngI, n), f32), hI=np.zeros((ngI, n), f32))

    def draw(mean, sd, nonneg):
        nonlocal norm
        while True:
            r = f32(f32(norm(g) * f32(sd)) + f32(mean))
            if not nonneg or not r < 0:
                return r

    tmax = f32(np.max(tonic))
    for i in range(2):
        if sp.iModel == 1:
            m, s = sp.init_w0[2 * i], sp.init_w0[2 * i + 1]
            for b in range(sp.nblock):
                for j in range(acc0[i], acc0[i + 1]):
                    out["w"][b * bs + j] = draw(m, s, False)
        if sp.use_v0 or tmax > 1.0:
            a, c = f32(sp.vR[i]), f32(sp.vT[i])
            for b in range(sp.nblock):
                for j in range(acc0[i], acc0[i + 1]):
                    out["v"][b * bs + j] = f32(f32(g.canonical() * f32(c - a)) + a)
        else:
            for b in range(sp.nblock):
                for j in range(acc0[i], acc0[i + 1]):
                    out["v"][b * bs + j] = f32(f32(sp.vR[i]) + f32(f32(f32(sp.vT[i]) - f32(sp.vR[i]))
                                                               * tonic[b * bs + j]))

    def fill(name, dist, ng, i, k):
        d = [f32(x) for x in dist[(i * ng + k) * 4:(i * ng + k) * 4 + 4]]
        for b in range(sp.nblock):
            if interleaved:                   # g and h per neuron: not the reference's order
                for j in range(acc0[i], acc0[i + 1]):
                    out["g" + name][k, b * bs + j] = draw(d[0], f32(d[0] * d[1]), True)
                    out["h" + name][k, b * bs + j] = draw(d[2], f32(d[2] * d[3]), True)
            else:
                for j in range(acc0[i], acc0[i + 1]):
                    out["g" + name][k, b * bs + j] = draw(d[0], f32(d[0] * d[1]), True)
                for j in range(acc0[i], acc0[i + 1]):
                    out["h" + name][k, b * bs + j] = draw(d[2], f32(d[2] * d[3]), True)
            if name == "FF":
                for j in range(acc0[i], acc0[i + 1]):
                    if nLGN[b * bs + j] == 0:
                        out["gFF"][k, b * bs + j] = out["hFF"][k, b * bs + j] = 0.0

    for i in range(2):
        for k in range(ngFF):
            if per_array_normal:
                norm = Normal()
            fill("FF", sp.init_gFF0, ngFF, i, k)
    for i in range(2):
        for k in range(ngE):
            if per_array_normal:
                norm = Normal()
            fill("E", sp.init_gE0, ngE, i, k)
        for k in range(ngI):
            if per_array_normal:
                norm = Normal()
            fill("I", sp.init_gI0, ngI, i, k)
    return out


INIT = dict(init_w0=(0.3, 0.1, 0.0, 0.0), init_gFF0=(0.2, 0.5, 0.1, 1.0, 0.05, 0.3, 0.0, 0.0,
                                                      0.3, 0.4, 0.2, 0.8, 0.06, 0.2, 0.01, 2.0),
            init_gE0=(0.1, 0.5, 0.05, 0.5, 0.02, 1.0, 0.0, 0.0, 0.2, 0.3, 0.1, 1.5, 0.01, 0.5,
                      0.02, 0.5),
            init_gI0=(0.4, 0.3, 0.2, 0.6, 0.5, 1.2, 0.3, 0.4))


def _net(pkg, **kw):
    from visual_cortex_amd import network
    base = dict(nblock=2, blockSize=128, nE=96, KE=20, KI=5, m=4, **INIT)
    base.update(kw)
    return network.build_network(network.spec_tiny(**base))


@pytest.mark.parametrize("kw", [dict(), dict(use_v0=1), dict(iModel=0), dict(seed=0)])
def test_initial_state_is_the_reference_recipe(pkg, kw):
    """Bit for bit against the restated generator, with v from the tonic
    drive, v uniform (use_v0), LIF (no w draws) and seed 0 (the engine's
    zero-seed rule); neurons without LGN input get gFF = hFF = 0."""
    from visual_cortex_amd import params
    net = _net(pkg, **kw)
    sp = net.spec
    net.nLGN_of = net.nLGN_of.copy()
    net.nLGN_of[::7] = 0
    net.tonicDep = np.linspace(0.1, 0.9, sp.nV1).astype(np.float32)
    st = params.initial_state(net)
    seed = sp.seed + 1 + sp.nV1
    ref = restated_initial_state(seed, sp, net.tonicDep, net.nLGN_of)
    for k in ref:
        if k == "w" and sp.iModel == 0:
            assert not st["w"].any()
            continue
        assert np.array_equal(st[k].view(np.uint32), ref[k].view(np.uint32)), k
    assert (st["gFF"][:, ::7] == 0).all() and (st["hFF"] >= 0).all() and (st["gE"] >= 0).all()
    if not sp.use_v0:
        assert np.all(st["v"] < np.asarray(sp.vT)[(np.arange(sp.nV1) % sp.blockSize >= sp.nE)
                                                  .astype(int)])


@pytest.mark.parametrize("mistake", [dict(per_array_normal=True), dict(interleaved=True)])
def test_initial_state_pin_detects_mistakes(pkg, mistake):
    """A normal distribution object per array (no carried second value) or g
    and h drawn per neuron instead of per range disagree with the library."""
    from visual_cortex_amd import params
    net = _net(pkg)
    sp = net.spec
    st = params.initial_state(net)
    ref = restated_initial_state(sp.seed + 1 + sp.nV1, sp, net.tonicDep, net.nLGN_of, **mistake)
    assert any(not np.array_equal(st[k], ref[k]) for k in ref)


# ---------------------------------------------------------------- rand_spInit
def _spinit(pkg, orc, net, sp0, SCsplit, depthRef, depth, cur0, v0=-60.0):
    from visual_cortex_amd import abi
    p = _params(pkg, net)
    q = abi.V1SpInit()
    for i, x in enumerate(sp0):
        q.sp0[i] = x
    q.SCsplit, q.depthRef = SCsplit, depthRef
    n = net.spec.nV1
    v = np.full(n, v0, np.float32)
    w = np.linspace(0.0, 1.0, n).astype(np.float32)
    tb = np.zeros(n, np.float32)
    train = np.zeros((depth, n), np.float32)
    nl = np.ascontiguousarray(net.nLGN_of, np.uint32)
    rc = orc.L.orc_v1_initial_spikes(C.byref(p), _fp(nl), C.byref(q), _fp(v), _fp(w), _fp(tb),
                                     _fp(train), depth, cur0)
    assert rc == 0
    return p, v, w, tb, train


def test_initial_spike_fractions_and_times(pkg, orc):
    """Each (type, simple/complex) class spikes in reference slot 0 with its
    own chance (binomial, 5 sigma), at uniform times (KS); with tRef below
    the step every later slot has the same chance; a neuron that spiked is at
    vR with tBack = tRef - (1 - tsp) dt when positive."""
    net = _net(pkg, nblock=16, blockSize=256, nE=192, init_gFF0=None, init_gE0=None,
               init_gI0=None, init_w0=None, tRef=(0.05, 0.1), iModel=0)
    sp = net.spec
    net.nLGN_of = ((np.arange(sp.nV1) % 3) * 3 + 2).astype(np.uint32)   # 2, 5, 8 inputs
    nl = net.nLGN_of
    SC = 4
    chances = (0.3, 0.1, 0.5, 0.2)
    D = 5
    p, v, w, tb, train = _spinit(pkg, orc, net, chances, SC, D, D, 0)
    t = (np.arange(sp.nV1) % sp.blockSize >= sp.nE).astype(int)
    cls = t * 2 + (nl <= SC).astype(int)
    s0 = train[0]
    for c in range(4):
        sel = cls == c
        k, nn = int((s0[sel] >= 1).sum()), int(sel.sum())
        mu, sd = nn * chances[c], np.sqrt(nn * chances[c] * (1 - chances[c]))
        assert abs(k - mu) < 5 * sd, (c, k, mu)
        # later slots: tRef < dt, so the chance is chance * (dt - ref)/dt
        # with ref = max(0, ...) <= tRef: between (1 - tRef/dt) chance and chance
        for j in range(1, D):
            kj = int((train[j][sel] >= 1).sum())
            lo = nn * chances[c] * (1 - sp.tRef[c // 2] / sp.dt)
            assert lo - 5 * sd < kj < mu + 5 * sd, (c, j, kj)
    spk = s0 >= 1
    tsp = s0[spk] - 1.0
    assert stats.kstest(tsp, "uniform").pvalue > 1e-3
    assert np.all(v[spk] == np.asarray(sp.vR, np.float32)[t[spk]])
    assert np.all(v[~spk] == np.float32(-60.0))
    exp_tb = (np.asarray(sp.tRef, np.float32)[t[spk]] - (np.float32(1) - tsp.astype(np.float32))
              * np.float32(sp.dt)).astype(np.float32)
    # (tsp read back from 1 + tsp carries that sum's rounding, 2^-23)
    np.testing.assert_allclose(tb[spk], np.where(exp_tb > 0, exp_tb, 0), rtol=0, atol=3e-8)


def test_initial_spike_chain_order(pkg, orc):
    """With chance 1 and tRef = 1.5 dt a spike at tsp in reference slot 0
    leaves the neuron refractory through the next step of the chain when
    tsp <= 0.5 (ref = (1.5 - tsp) dt >= dt): that step is reference slot
    depthRef-1 and holds vR; with tsp > 0.5 it spikes there with chance
    (tsp - 0.5).  Slots are placed at (cur0 + depth - depthRef + j) % depth,
    older ones hold v."""
    net = _net(pkg, nblock=8, blockSize=256, nE=192, init_gFF0=None, init_gE0=None,
               init_gI0=None, init_w0=None, iModel=0, tRef=(0.1875, 0.1875))
    sp = net.spec
    D, depth, cur0 = 4, 7, 5
    p, v, w, tb, train = _spinit(pkg, orc, net, (1.0, 1.0, 1.0, 1.0), 0, D, depth, cur0)
    slot = lambda j: (cur0 + depth - D + j) % depth
    s0 = train[slot(0)]
    assert (s0 >= 1).all()
    tsp = s0 - 1.0
    nxt = train[slot(D - 1)]
    early = tsp <= 0.5
    assert np.all(nxt[early] == np.float32(sp.vR[0]))
    late = ~early
    frac = float((nxt[late] >= 1).mean())
    expect = float(np.mean(tsp[late] - 0.5))
    assert abs(frac - expect) < 5 * np.sqrt(expect * (1 - expect) / late.sum())
    # ring slots of steps before -D hold v (vR after the slot-0 spike)
    for k in range(D, depth):
        assert np.all(train[(cur0 + depth - 1 - k) % depth] == v)
