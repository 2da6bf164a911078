"""Pins of the oracle's multi-spike update (compute_V_collect_spike_learnFF,
src/coredynamics.cu:427-1233, which the reference runs when dt > min tRef)
and of the dense-scan recurrent gather it pairs with (recal_G_mat_nd,
src/coredynamics.cu:2364-2607).

* For a type with tRef >= dt the multi-spike update must give the one-spike
  update's results bit for bit (the reference switches the whole network to
  it when any type has tRef < dt).
* With tRef < dt, a float64 model written from the integration scheme in
  absolute time (per step: one Heun step from each restart to the end of the
  step, conductances at the restart from their closed forms, linear
  interpolation of each crossing, restart at tsp + tRef) gives the same spike
  counts and spike times; it shares no bookkeeping with the restatement
  (new_t0, dtBack, the incremental decays and window re-reads).
* The dense-scan gather over per-neuron increments equals the compacted
  fast gather bit for bit on steps with several spikes per neuron.
"""
import ctypes as C

import numpy as np
import pytest

from test_oracle import hand_network

VL = -65.0


def _fp(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _params(pkg, net, **over):
    from visual_cortex_amd import params
    p = params.v1_params_struct(net)
    for k, v in over.items():
        arr = getattr(p, k)
        for i, x in enumerate(v):
            arr[i] = x
    return p


class Harness:
    """orc_v1_step on a hand-built network with LGN inputs given per step."""

    def __init__(self, orc, net, p, nLGN_of, sLGN, idLGN, tonic, v0):
        self.orc, self.net, self.p = orc, net, p
        sp = net.spec
        n = sp.nV1
        z = lambda *s: np.zeros(s, np.float32)
        self.st = dict(v=np.full(n, v0, np.float32), w=z(n), tBack=z(n),
                       gFF=z(2, n), hFF=z(2, n), gE=z(2, n), hE=z(2, n), gI=z(1, n), hI=z(1, n))
        self.nLGN_of = np.ascontiguousarray(nLGN_of, np.uint32)
        self.sLGN = np.ascontiguousarray(sLGN, np.float32)
        self.idLGN = np.ascontiguousarray(idLGN, np.uint32)
        self.tonic = np.ascontiguousarray(tonic, np.float32)
        self.conMat = np.ascontiguousarray(net.dense_conmat())
        self.nNear = np.ascontiguousarray(net.nNear, np.uint32)
        self.nbr = np.ascontiguousarray(net.nbr, np.uint32)
        self.ipre = np.zeros(n, np.uint32)
        self.npre = np.zeros(2 * sp.nblock, np.uint32)
        ng = 2 * sp.nE + 1 * (sp.blockSize - sp.nE)
        self.og = np.zeros(sp.nblock * ng, np.float32)
        self.oh = np.zeros(sp.nblock * ng, np.float32)

    def step(self, lgn_row, it):
        st = self.st
        out = np.zeros(self.net.spec.nV1, np.float32)
        rc = self.orc.L.orc_v1_step(
            C.byref(self.p), _fp(st["v"]), _fp(st["w"]), _fp(st["tBack"]), _fp(st["gFF"]),
            _fp(st["hFF"]), _fp(st["gE"]), _fp(st["hE"]), _fp(st["gI"]), _fp(st["hI"]),
            _fp(self.tonic), _fp(self.nLGN_of), _fp(self.sLGN), _fp(self.idLGN),
            _fp(np.ascontiguousarray(lgn_row, np.float32)), self.net.maxNear, _fp(self.nNear),
            _fp(self.nbr), _fp(self.conMat), _fp(self.ipre), _fp(self.npre), _fp(self.og),
            _fp(self.oh), it, _fp(out), None, 1, None, None, None)
        assert rc == 0
        return out


def _lif_case(pkg, orc, nsteps, dt, tRef, dep, seed):
    """64 LIF neurons (both types: deltaT = a = b = 0), tonic drive `dep`
    (a per-neuron array), 4 LGN synapses each from 16 cells whose spikes
    (1-2 per event) arrive at random times."""
    rng = np.random.default_rng(seed)
    net = hand_network(pkg, bgStrength=(0.0, 0.0), deltaT=(0.0, 0.0), a=(0.0, 0.0),
                       b=(0.0, 0.0), vThres=(-50.0, -50.0), vR=(-58.0, -58.0),
                       tRef=tRef, dt=dt, iModel=0)
    p = _params(pkg, net)
    n, m, nc = net.spec.nV1, 4, 16
    nLGN_of = np.full(n, m, np.uint32)
    sLGN = rng.uniform(0.5, 3.0, (m, n)).astype(np.float32)
    idLGN = rng.integers(0, nc, (m, n)).astype(np.uint32)
    rows = []
    for _ in range(nsteps):
        r = np.full(nc, -60.0, np.float32)             # not spiking: a voltage (< 1)
        fire = rng.random(nc) < 0.3
        nsp = rng.integers(1, 3, nc)
        frac = rng.random(nc).astype(np.float32)
        r[fire] = (nsp[fire] + frac[fire]).astype(np.float32)
        rows.append(r)
    h = Harness(orc, net, p, nLGN_of, sLGN, idLGN, dep, -58.0)
    return net, p, h, rows


def _model64(sp, p, t, tonic, fl, idl, rows, v0, mistake=None):
    """Float64 restatement of the scheme for one LIF neuron of type t: per
    step the list of (nsp, tsp) of its spikes and the end-of-step v."""
    dt = sp.dt
    gL, C_, vR, vTh, tRef, vT = (p.gL[t], p.C[t], p.vR[t], p.vThres[t], p.tRef[t], p.vT[t])
    targetV = VL + (vT - VL) * tonic
    depC = gL * (targetV - VL)
    rise = [p.riseFF[0], p.riseFF[1]]
    dec = [p.decayFF[0], p.decayFF[1]]
    pff = [p.pFF[t * 2 + 0], p.pFF[t * 2 + 1]]

    def si(f, ddt, ig):                     # unit input's (g, h) after ddt
        dod = rise[ig] / (dec[ig] - rise[ig])
        return f * dod * (np.exp(-ddt / dec[ig]) - np.exp(-ddt / rise[ig])), f * np.exp(-ddt / rise[ig])

    def dec_(g, h, ddt, ig):
        dod = rise[ig] / (dec[ig] - rise[ig])
        etd, etr = np.exp(-ddt / dec[ig]), np.exp(-ddt / rise[ig])
        return g * etd + h * dod * (etd - etr), h * etr

    g = [0.0, 0.0]
    hh = [0.0, 0.0]
    v, tBack = v0, 0.0
    out = []
    for r in rows:
        ev = []
        for f, c in zip(fl, idl):
            s = float(r[c])
            if s >= 1.0:
                k = np.floor(s)
                ev.append((float(f), k, (s - k) * dt))

        def gff(tau):                       # total FF conductance at tau in [0, dt]
            tot = 0.0
            for ig in range(2):
                a_, b_ = dec_(g[ig], hh[ig], tau, ig)
                tot += a_
                for f, k, ts in ev:
                    if ts < tau:
                        tot += si(f * pff[ig] * k, tau - ts, ig)[0]
            return tot
        spikes = []
        if tBack >= dt:
            v = vR
            tBack -= dt
        else:
            t0 = tBack if tBack > 0 else 0.0
            vs = v
            while True:
                hstep = dt - t0
                a0, b0 = gff(0.0 if mistake == "restart_g_at_0" else t0) + gL, gL * VL + depC
                a1, b1 = gff(dt) + gL, gL * VL + depC
                fk1 = (-a0 * vs + b0) / C_
                v1 = vs + hstep * fk1
                fk2 = (-a1 * v1 + b1) / C_
                ve = vs + hstep * (fk1 + fk2) / 2
                if ve > vTh:
                    ts = t0 + (vTh - vs) / (ve - vs) * hstep
                    if mistake == "tsp_from_0":
                        ts = (vTh - vs) / (ve - vs) * hstep
                    spikes.append(ts)
                    tBack = ts + tRef
                    if tBack < dt:
                        t0, vs = tBack, (ve if mistake == "restart_from_v" else vR)
                        continue
                    v = vR
                    tBack -= dt
                else:
                    v = ve
                    tBack = 0.0
                break
        # the conductances at the end of the step, for the next one
        ng, nh = [0.0, 0.0], [0.0, 0.0]
        for ig in range(2):
            ng[ig], nh[ig] = dec_(g[ig], hh[ig], dt, ig)
            for f, k, ts in ev:
                a_, b_ = si(f * pff[ig] * k, dt - ts, ig)
                ng[ig] += a_
                nh[ig] += b_
        g, hh = ng, nh
        out.append((spikes, v))
    return out


def test_multi_equals_one_spike_update_when_tref_ge_dt(pkg, orc):
    """The reference's switch (every neuron through the multi-spike update
    when dt > min tRef) changes nothing for a type with tRef >= dt: a whole
    network (AdEx E, LIF I, recurrent and LGN input, background drive, white
    noise and LGN failure) bit for bit over 150 steps."""
    import oracle
    from visual_cortex_amd import network
    net = network.build_network(network.spec_tiny(bgStrength=(0.4, 0.4), noisyDep=(0.04, 0.08),
                                                  noiseOnTonic=0, synFailFF=(0.3, 0.2),
                                                  synPerConFF=(20.0, 30.0)))
    a = oracle.OracleSim(net, pkg)
    a.run(150)
    b = oracle.OracleSim(net, pkg, tables=a.tables)
    orc.L.orc_set_multi_all(1)
    try:
        b.run(150)
    finally:
        orc.L.orc_set_multi_all(0)
    sa, sb = a.sInfo(), b.sInfo()
    assert (sa >= 1).sum() > 100
    assert np.array_equal(sa.view(np.uint32), sb.view(np.uint32))
    for k in ("v", "w", "tBack", "gFF", "hFF", "gE", "hE", "gI", "hI"):
        assert np.array_equal(a.st[k].view(np.uint32), b.st[k].view(np.uint32)), k


@pytest.mark.parametrize("dt,tRef", [(0.125, (0.03, 0.05)), (0.5, (0.1, 0.2))])
def test_multi_spike_times_against_float64_scheme(pkg, orc, dt, tRef):
    """Several spikes per step (tRef < dt, strong tonic drive, LGN input
    events inside the steps, some inside refractory windows): spike counts
    per step equal the float64 model's, spike times within 2e-4 ms, the
    sInfo encoding nsp + mean(tsp)/dt, and the end-of-step v within 2e-3
    mV; neurons whose float32 and float64 runs take a different branch at a
    threshold crossing (a tie to rounding) are counted and must be rare."""
    nsteps = 60
    rng = np.random.default_rng(5)
    dep = rng.uniform(8.0, 40.0, 64).astype(np.float32)
    net, p, h, rows = _lif_case(pkg, orc, nsteps, dt, tRef, dep, seed=11)
    sp = net.spec
    outs = np.stack([h.step(r, it) for it, r in enumerate(rows)])
    multi_steps = int((outs >= 2.0).sum())
    assert multi_steps > 100, multi_steps              # the multi-spike path is exercised
    bad = 0
    for nrn in range(sp.nV1):
        t = 0 if nrn % sp.blockSize < sp.nE else 1
        ref = _model64(sp, p, t, float(dep[nrn]), h.sLGN[:, nrn], h.idLGN[:, nrn], rows, -58.0)
        ok = True
        for it, (spk, v) in enumerate(ref):
            s = float(outs[it, nrn])
            if spk:
                k = len(spk)
                if not (s >= 1.0 and int(np.floor(s)) == k):
                    ok = False
                    break
                mean = (s - k) * dt
                if abs(mean - float(np.mean(spk))) > 2e-4:
                    ok = False
                    break
            else:
                if s >= 1.0 or abs(s - v) > 2e-3:
                    ok = False
                    break
        bad += 0 if ok else 1
    assert bad <= 2, f"{bad} of {sp.nV1} neurons differ from the float64 scheme"


@pytest.mark.parametrize("mistake", ["restart_g_at_0", "restart_from_v", "tsp_from_0"])
def test_float64_scheme_detects_bookkeeping_mistakes(pkg, orc, mistake):
    """The comparison has power: the float64 model with one plausible
    bookkeeping mistake (conductances of a restart taken at the start of the
    step; a restart from the crossing voltage instead of vR; the spike time
    interpolated from the start of the step instead of the restart) disagrees
    with the oracle on most neurons."""
    nsteps = 40
    rng = np.random.default_rng(5)
    dep = rng.uniform(8.0, 40.0, 64).astype(np.float32)
    net, p, h, rows = _lif_case(pkg, orc, nsteps, 0.125, (0.03, 0.05), dep, seed=11)
    sp = net.spec
    outs = np.stack([h.step(r, it) for it, r in enumerate(rows)])
    bad = 0
    for nrn in range(sp.nV1):
        t = 0 if nrn % sp.blockSize < sp.nE else 1
        ref = _model64(sp, p, t, float(dep[nrn]), h.sLGN[:, nrn], h.idLGN[:, nrn], rows, -58.0,
                       mistake=mistake)
        for it, (spk, v) in enumerate(ref):
            s = float(outs[it, nrn])
            k = len(spk)
            if k:
                if not (s >= 1.0 and int(np.floor(s)) == k) or \
                        abs((s - k) * sp.dt - float(np.mean(spk))) > 2e-4:
                    bad += 1
                    break
            elif s >= 1.0 or abs(s - v) > 2e-3:
                bad += 1
                break
    assert bad > sp.nV1 // 2, f"only {bad} neurons differ with the mistake {mistake}"


def test_dense_gather_equals_fast_gather_with_multi_spikes(pkg, orc):
    """recal_G_mat_nd (dense scan, per-neuron increments) and the compacted
    recal_G_mat_nd_fast give the same sums bit for bit on a step with
    several spikes per presynaptic neuron, with and without synaptic
    failure; the failure draws are keyed by the spike's step."""
    from visual_cortex_amd import network
    for fail in (False, True):
        kw = dict(synFail=(0.2, 0.3, 0.4, 0.1), synPerCon=(5.0, 7.5, 4.26, 7.88)) if fail else {}
        net = network.build_network(network.spec_tiny(nblock=3, tRef=(0.04, 0.05), **kw))
        p = _params(pkg, net)
        sp = net.spec
        rng = np.random.default_rng(7)
        n = sp.nV1
        s = rng.uniform(-70, -55, n).astype(np.float32)
        spk = rng.random(n) < 0.2
        s[spk] = (rng.integers(1, 4, spk.sum()) + rng.random(spk.sum())).astype(np.float32)
        conMat = np.ascontiguousarray(net.dense_conmat())
        nNear = np.ascontiguousarray(net.nNear, np.uint32)
        nbr = np.ascontiguousarray(net.nbr, np.uint32)
        z = lambda *sh: np.zeros(sh, np.float32)
        gE1, hE1, gI1, hI1 = z(2, n), z(2, n), z(1, n), z(1, n)
        rc = orc.L.orc_v1_rec_dense(C.byref(p), _fp(gE1), _fp(hE1), _fp(gI1), _fp(hI1),
                                    net.maxNear, _fp(nNear), _fp(nbr), _fp(conMat), _fp(s), 9)
        assert rc == 0
        # the fast gather through orc_v1_step: a step whose own outputs are s
        # is not available, so compact s by hand into the reference's lists
        bs, nE, nI = sp.blockSize, sp.nE, sp.blockSize - sp.nE
        block_ng = 2 * nE + 1 * nI
        ipre = np.zeros(n, np.uint32)
        npre = np.zeros(2 * sp.nblock, np.uint32)
        og = np.zeros(sp.nblock * block_ng, np.float32)
        oh = np.zeros(sp.nblock * block_ng, np.float32)
        for b in range(sp.nblock):
            loc = np.nonzero(s[b * bs:(b + 1) * bs] >= 1)[0]
            le, li = loc[loc < nE], loc[loc >= nE]
            cE, cI = len(le), len(li)
            npre[b], npre[sp.nblock + b] = cE, cI
            ipre[b * bs:b * bs + cE] = le
            ipre[b * bs + cE:b * bs + cE + cI] = li
            for j, l in enumerate(le):
                x = s[b * bs + l]
                k = np.float32(np.floor(x))
                ts = np.float32((x - k) * np.float32(sp.dt))
                for ig in range(2):
                    g_, h_ = orc.cond_single(0.0, 0.0, float(k), float(np.float32(sp.dt) - ts),
                                             p.riseE[ig], p.decayE[ig])
                    og[ig * cE + block_ng * b + j], oh[ig * cE + block_ng * b + j] = g_, h_
            for j, l in enumerate(li):
                x = s[b * bs + l]
                k = np.float32(np.floor(x))
                ts = np.float32((x - k) * np.float32(sp.dt))
                g_, h_ = orc.cond_single(0.0, 0.0, float(k), float(np.float32(sp.dt) - ts),
                                         p.riseI[0], p.decayI[0])
                og[block_ng * b + cE * 2 + j], oh[block_ng * b + cE * 2 + j] = g_, h_
        gE2, hE2, gI2, hI2 = z(2, n), z(2, n), z(1, n), z(1, n)
        rc = orc.L.orc_v1_rec_fast(C.byref(p), _fp(gE2), _fp(hE2), _fp(gI2), _fp(hI2),
                                   net.maxNear, _fp(nNear), _fp(nbr), _fp(conMat), _fp(ipre),
                                   _fp(npre), _fp(og), _fp(oh), 9)
        assert rc == 0
        assert np.abs(gE1).max() > 0 and np.abs(gI1).max() > 0
        for a_, b_ in ((gE1, gE2), (hE1, hE2), (gI1, gI2), (hI1, hI2)):
            assert np.array_equal(a_.view(np.uint32), b_.view(np.uint32))
