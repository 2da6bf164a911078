"""Pins of the CPU oracle against what the mathematics fixes (the reference
holds no golden vectors for this path): closed forms of the conductance ODE,
Heun's method on a linear membrane, the AdEx tonic-drive steady state, the
LIF inter-spike interval, brute-force recurrent sums, the compaction
invariants, gamma-pdf temporal weights, the logistic identities, Poisson
spike statistics and bilinear filtering."""
import math
from dataclasses import replace

import numpy as np
import pytest
from scipy import stats


# ---------------------------------------------------------------- condShape
@pytest.mark.parametrize("rise,decay", [(1.0, 3.0), (5.0, 80.0), (1.0, 5.0)])
def test_single_input_closed_form(orc, rise, decay):
    f, t = 0.7, 2.3
    g, h = orc.cond_single(0.0, 0.0, f, t, rise, decay)
    dod = rise / (decay - rise)
    g_ref = f * dod * (math.exp(-t / decay) - math.exp(-t / rise))
    h_ref = f * math.exp(-t / rise)
    assert g == pytest.approx(g_ref, rel=2e-6)
    assert h == pytest.approx(h_ref, rel=2e-6)


@pytest.mark.parametrize("rise,decay", [(1.0, 3.0), (5.0, 80.0)])
def test_decay_solves_ode(orc, rise, decay):
    """(g, h) follow h' = -h/rise, g' = (h - g)/decay; fine-step RK4 in
    float64 from the same start, and the time integral of g for a unit input
    is rise (= f*rise)."""
    g, h = orc.cond_single(0.0, 0.0, 1.0, 0.0, rise, decay)   # unit input at t=0
    dt, T = 0.125, 40.0
    G, H = float(g), float(h)
    integral = 0.0
    gs = [g]
    for _ in range(int(T / dt)):
        g, h = orc.cond_decay(g, h, dt, rise, decay)
        gs.append(g)
    # RK4 reference
    def f(y):
        return np.array([(y[1] - y[0]) / decay, -y[1] / rise])
    y = np.array([G, H])
    sub = 64
    hh = dt / sub
    for _ in range(int(T / dt) * sub):
        k1 = f(y); k2 = f(y + hh / 2 * k1); k3 = f(y + hh / 2 * k2); k4 = f(y + hh * k3)
        y = y + hh / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    assert g == pytest.approx(y[0], rel=1e-4, abs=1e-7)
    assert h == pytest.approx(y[1], rel=1e-4, abs=1e-9)
    if decay < 10:
        integral = np.trapezoid(np.array(gs, np.float64), dx=dt)
        assert integral == pytest.approx(rise, rel=2e-3)


# ---------------------------------------------------------------- helpers
def hand_network(pkg, nblock=1, bs=64, nE=48, conmat=None, ff=None, **spec_kw):
    """A Network built by hand: no LGN input unless ff=(nLGN_of, sLGN, idLGN)."""
    from visual_cortex_amd import network
    sp = network.spec_tiny(nblock=nblock, blockSize=bs, nE=nE, KE=0, KI=0, m=0,
                           nLGN_side=4, width=8, height=8, **spec_kw)
    net = network.build_network(sp)
    nV1 = nblock * bs
    if conmat is not None:
        pairs = []
        for b in range(nblock):
            for ib in range(net.maxNear):
                M = conmat[b, ib]
                pre, post = np.nonzero(M > 0)
                pairs.append((pre.astype(np.uint32), post.astype(np.uint32), M[pre, post].astype(np.float32)))
        net.rec_pairs = pairs
    if ff is None:
        net.nLGN_of = np.zeros(nV1, np.uint32)
        net.sLGN = np.zeros((1, nV1), np.float32)
        net.idLGN = np.zeros((1, nV1), np.uint32)
    return net


def osim(pkg, net, state=None):
    import oracle
    return oracle.OracleSim(net, pkg, state=state)


def flat_state(net, v):
    from visual_cortex_amd import params
    st = params.initial_state(net)
    st["v"][:] = v
    return st


# ---------------------------------------------------------------- neuron
def test_heun_linear_membrane(pkg):
    """LIF (deltaT = 0, a = b = 0), no input: Heun's amplification factor
    1 - x + x^2/2, x = dt*gL/C, toward vL (closed form)."""
    net = hand_network(pkg, bgStrength=(0.0, 0.0), deltaT=(0.0, 0.0), a=(0.0, 0.0), b=(0.0, 0.0),
                       vThres=(-20.0, -20.0))
    sp = net.spec
    sim = osim(pkg, net, flat_state(net, -55.0))
    n = 200
    sim.run(n)
    v = sim.sInfo()[:, 0].astype(np.float64)
    x = sp.dt * sp.gL[0] / sp.C[0]
    amp = 1 - x + x * x / 2
    ref = -65.0 + (-55.0 + 65.0) * amp ** np.arange(1, n + 1)
    np.testing.assert_allclose(v, ref, rtol=0, atol=2e-4)
    xi = sp.dt * sp.gL[1] / sp.C[1]
    ampi = 1 - xi + xi * xi / 2
    refi = -65.0 + 10.0 * ampi ** np.arange(1, n + 1)
    np.testing.assert_allclose(sim.sInfo()[:, sp.nE].astype(np.float64), refi, rtol=0, atol=2e-4)


def test_adex_tonic_steady_state(pkg):
    """AdEx with tonic drive: v = targetV, w = a(targetV - vL) is a fixed point
    of the continuous dynamics (depC, src/coredynamics.cuh:41-49)."""
    dep = 0.6
    net = hand_network(pkg, bgStrength=(0.0, 0.0), a=(0.5, 0.0), tonicDep=(dep, 0.0))
    sp = net.spec
    vT, vL = sp.vT[0], -65.0
    targetV = vL + (vT - vL) * dep
    st = flat_state(net, targetV)
    st["w"][:] = sp.a[0] * (targetV - vL)
    sim = osim(pkg, net, st)
    sim.run(400)
    v = sim.sInfo()[:, 0]
    assert np.max(np.abs(v - targetV)) < 2e-3
    assert abs(sim.st["w"][0] - sp.a[0] * (targetV - vL)) < 1e-3


def test_lif_interspike_interval(pkg):
    """LIF with tonic drive above threshold: ISI = tRef + tau*ln((vinf - vR) /
    (vinf - vThres)) with tau = C/gL (continuous solution; Heun at dt = 0.125 ms
    is second order)."""
    dep = 2.0   # targetV = vL + (vT - vL)*dep = -35 mV > vThres
    net = hand_network(pkg, bgStrength=(0.0, 0.0), deltaT=(0.0, 0.0), a=(0.0, 0.0), b=(0.0, 0.0),
                       vThres=(-50.0, -50.0), tonicDep=(dep, dep))
    sp = net.spec
    sim = osim(pkg, net, flat_state(net, sp.vR[0]))
    n = 4000
    sim.run(n)
    s = sim.sInfo()[:, 0]
    steps = np.nonzero(s >= 1)[0]
    tsp = (steps + (s[steps] - 1.0)) * sp.dt
    isi = np.diff(tsp)
    vinf = -65.0 + (sp.vT[0] + 65.0) * dep
    tau = sp.C[0] / sp.gL[0]
    ref = sp.tRef[0] + tau * math.log((vinf - sp.vR[0]) / (vinf - sp.vThres[0]))
    assert len(isi) > 10
    assert np.allclose(isi, ref, atol=0.02), (isi[:5], ref)


# ---------------------------------------------------------------- recurrent
def test_recurrent_sum_brute_force_and_lists(pkg):
    """One step of a hand-built 2-block network with strong drive: the
    compacted lists are the ascending spiking sets, the counts exact, and the
    recurrent increments equal a float64 brute-force sum over the dense
    matrix (tolerance of float32 summation)."""
    import ctypes as C
    rng = np.random.default_rng(5)
    bs, nb = 64, 2
    net0 = hand_network(pkg, nblock=nb, bs=bs, nE=48)
    maxNear = net0.maxNear
    M = np.where(rng.random((nb, maxNear, bs, bs)) < 0.3,
                 rng.lognormal(-1, 0.8, (nb, maxNear, bs, bs)), 0).astype(np.float32)
    net = hand_network(pkg, nblock=nb, bs=bs, nE=48, conmat=M, tonicDep=(3.0, 3.0),
                       bgStrength=(0.0, 0.0), deltaT=(0.0, 0.0), a=(0.0, 0.0), b=(0.0, 0.0),
                       vThres=(-50.0, -50.0))
    st = flat_state(net, -50.05)
    sim = osim(pkg, net, st)
    sim.run(1)                                   # everybody crosses in step 0
    s = sim.sInfo()[0]
    spk = s >= 1
    loc = np.arange(s.size) % bs
    assert spk[loc < 48].sum() > 20 and spk[loc >= 48].sum() > 10
    sp = net.spec
    npre = sim.npre
    for b in range(nb):
        loc = np.nonzero(spk[b * bs:(b + 1) * bs])[0]
        E = loc[loc < sp.nE]; I = loc[loc >= sp.nE]
        assert npre[b] == len(E) and npre[nb + b] == len(I)
        assert np.array_equal(sim.ipre[b * bs: b * bs + len(E) + len(I)], np.concatenate([E, I]))
    # brute force increments: decayed state (do_rec=0) vs applied (oracle state)
    sim2 = osim(pkg, net, flat_state(net, -50.05))
    o = sim2.o
    out = np.zeros(sp.nV1, np.float32)
    lgn_row = np.zeros(sp.nLGN, np.float32)
    fp = lambda a: a.ctypes.data_as(C.c_void_p)
    t = sim2
    o.L.orc_v1_step(C.byref(t.p), fp(t.st["v"]), fp(t.st["w"]), fp(t.st["tBack"]), fp(t.st["gFF"]),
                    fp(t.st["hFF"]), fp(t.st["gE"]), fp(t.st["hE"]), fp(t.st["gI"]), fp(t.st["hI"]),
                    fp(t.tonic), fp(t.nLGN_of), fp(t.sLGN), fp(t.idLGN), fp(lgn_row), net.maxNear,
                    fp(t.nNear), fp(t.nbr), fp(t.conMat), fp(t.ipre), fp(t.npre), fp(t.og), fp(t.oh),
                    0, fp(out), None, 0, None, None, None)
    dgE = sim.st["gE"].astype(np.float64) - t.st["gE"].astype(np.float64)
    dhI = sim.st["hI"].astype(np.float64) - t.st["hI"].astype(np.float64)
    dt = sp.dt
    ref_gE = np.zeros((len(sp.riseTimeE), sp.nV1))
    ref_hI = np.zeros(sp.nV1)
    for post in range(sp.nV1):
        b, l = divmod(post, bs)
        tpost = 0 if l < sp.nE else 1
        for ib in range(net.nNear[b]):
            bid = net.nbr[b, ib]
            for pre in range(bs):
                gpre = bid * bs + pre
                if not spk[gpre] or M[b, ib, pre, l] <= 0:
                    continue
                tsp = (s[gpre] - 1.0) * dt
                w = float(M[b, ib, pre, l])
                if pre < sp.nE:
                    for ig in range(len(sp.riseTimeE)):
                        r_, d_ = sp.riseTimeE[ig], sp.decayTimeE[ig]
                        og = r_ / (d_ - r_) * (math.exp(-(dt - tsp) / d_) - math.exp(-(dt - tsp) / r_))
                        ref_gE[ig, post] += w * sp.pE[tpost][ig] * og
                else:
                    r_ = sp.riseTimeI[0]
                    ref_hI[post] += w * sp.pI[tpost][0] * math.exp(-(dt - tsp) / r_)
    assert ref_gE.max() > 0 and ref_hI.max() > 0
    np.testing.assert_allclose(dgE, ref_gE, rtol=2e-5, atol=2e-7)
    np.testing.assert_allclose(dhI[0], ref_hI, rtol=2e-5, atol=2e-7)


# ---------------------------------------------------------------- LGN
def test_temporal_weights_are_gamma_differences(pkg):
    from visual_cortex_amd import network
    import oracle
    net = network.build_network(network.spec_tiny())
    sim = oracle.OracleSim(net, pkg)
    sp = net.spec
    tm = network.timing(sp)
    TW = sim.tables["TW"][0, 0].astype(np.float64)
    dk = tm["kernelSampleInterval"] * sp.dt
    t = (sp.nKernelSample - 1 - np.arange(sp.nKernelSample)) * dk + tm["iKernelSampleT0"] * sp.dt
    ref = stats.gamma.pdf(t, sp.nR, scale=sp.tauR) - sp.ratio * stats.gamma.pdf(t, sp.nD, scale=sp.tauD)
    ref /= np.abs(ref).sum()
    np.testing.assert_allclose(TW, ref, rtol=2e-5, atol=1e-8)
    assert abs(np.abs(TW).sum() - 1.0) < 1e-6
    assert abs(sim.tables["SW"].astype(np.float64).sum() - 1.0) < 1e-6


def test_spatial_weights_gaussian(pkg):
    from visual_cortex_amd import network
    import oracle
    net = network.build_network(network.spec_tiny())
    sim = oracle.OracleSim(net, pkg)
    sp = net.spec
    n1 = sp.nSpatialSample1D
    u = ((np.arange(n1) + 0.5) * 2.0 / n1 - 1.0) * sp.nsig / math.sqrt(2)   # w / r
    ref = np.exp(-(u[None, :] ** 2) - (u[:, None] ** 2)).ravel()
    ref /= ref.sum()
    np.testing.assert_allclose(sim.tables["SW"], ref, rtol=1e-5)


def test_uniform_frame_has_zero_response(pkg):
    from visual_cortex_amd import network
    import oracle
    net = network.build_network(network.spec_tiny(field_deg=1.0))
    sim = oracle.OracleSim(net, pkg)
    sp = net.spec
    frame = np.full((3, sp.height, sp.width), 0.37, np.float32)
    R = sim.frame_response(frame)
    assert np.max(np.abs(R)) < 1e-6


def test_logistic_identities(pkg):
    """The precached a, b of the logistic (src/LGN_props.h:163-169) give
    f(0) = spont and f(1) = 1."""
    from visual_cortex_amd import network
    net = network.build_network(network.spec_tiny())
    L, sp = net.lgn, net.spec
    X = lambda x: 1.0 / (1.0 + math.exp(sp.sharpness * (sp.c50 - x)))
    assert L["logA"][0] * X(0.0) + L["logB"][0] == pytest.approx(sp.spont, abs=1e-6)
    assert L["logA"][0] * X(1.0) + L["logB"][0] == pytest.approx(1.0, abs=1e-6)


def test_lgn_poisson_statistics(pkg):
    """Constant rate (zero temporal weights make the convolution vanish, so
    fr = frRatio*max*f(0) = frRatio*max*spont): spike counts over a long run
    are Poisson(rate T) and the inter-spike intervals exponential."""
    import oracle
    from visual_cortex_amd import network
    net = network.build_network(network.spec_tiny(nt=64 * 300))
    sim = oracle.OracleSim(net, pkg)
    sim.tables["TW"][:] = 0.0
    sp = net.spec
    sim.run(64 * 300)
    s = sim.lgn_sinfo()
    rate = float(sim.fr_hist[0][0][0])          # Hz, cell 0 (all steps equal)
    assert np.all(np.concatenate(sim.fr_hist)[:, 0] == np.float32(rate))
    assert rate == pytest.approx(sp.frRatioLGN * sim.tables["maxConvol"][0] * sp.spont, rel=1e-5)
    T = s.shape[0] * sp.dt / 1000.0
    counts = np.floor(s).sum(0)
    fr_cells = np.concatenate(sim.fr_hist)[0]
    mu = float(np.mean(fr_cells)) * T
    assert abs(counts.mean() - mu) < 4 * math.sqrt(mu / counts.size)
    # dispersion index close to 1
    assert 0.6 < counts.var() / counts.mean() < 1.5
    # exponential ISIs (first cell with enough spikes)
    c = int(np.argmax(counts))
    st = np.nonzero(s[:, c] >= 1)[0]
    tsp = (st + (s[st, c] - np.floor(s[st, c]))) * sp.dt
    isi = np.diff(tsp)
    rate_c = float(fr_cells[c])
    assert stats.kstest(isi, "expon", args=(0, 1000.0 / rate_c)).pvalue > 1e-3


def test_bilinear_closed_forms(orc):
    H, W = 6, 8
    img = (np.arange(W)[None, :] * 0.5 + np.arange(H)[:, None] * 0.25 + 1.0).astype(np.float32)
    # texel centres reproduce texels
    for j in (1, 3):
        for i in (2, 5):
            assert orc.bilinear(img, (i + 0.5) / W, (j + 0.5) / H) == img[j, i]
    # a linear ramp is interpolated exactly (interior)
    x, y = (3.3 + 0.5) / W, (2.6 + 0.5) / H
    assert orc.bilinear(img, x, y) == pytest.approx(1.0 + 3.3 * 0.5 + 2.6 * 0.25, rel=1e-6)
    # border addressing: zero far outside, half-weighted at the edge
    assert orc.bilinear(img, -0.5, 0.5) == 0.0
    assert orc.bilinear(img, 0.0, (2 + 0.5) / H) == pytest.approx(0.5 * img[2, 0], rel=1e-6)


def test_block_reduce_integer_exact(orc):
    rng = np.random.default_rng(3)
    x = rng.integers(-1000, 1000, 256).astype(np.float32)
    assert orc.block_reduce(x, 256) == np.float32(x.astype(np.int64).sum())
