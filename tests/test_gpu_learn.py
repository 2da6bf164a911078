"""GPU parity of the feed-forward plasticity (learning = 3,
src/coredynamics.cu:1631-1992, src/discrete_input_convol.cu:1736-1821):
vc_v1_learn_ff on random inputs against the oracle through the C ABI, the
LGN trace of vc_lgn_steps_learn, and whole plastic runs (strengths, traces,
totals and every output) against the oracle bit for bit, with and without
homeostasis, with pruned strengths and rebound, clips, axonal delays and
several spikes per step."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from test_learn_oracle import LEARN  # noqa: E402


def _flush(msg):
    print(msg, flush=True)


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _tables(g):
    return dict(TW=g.TW.cpu().numpy(), SW=g.SW.cpu().numpy(), SCx=g.SCx.cpu().numpy(),
                SCy=g.SCy.cpu().numpy(), maxConvol=g.maxConvol.cpu().numpy())


def _net(**kw):
    from visual_cortex_amd import network
    base = dict(LEARN)
    base.update(kw)
    return network.build_network(network.spec_tiny(**base))


# m > 32 with homeostasis takes the one-lane-per-neuron kernel, the others
# the four-wavefront one
@pytest.mark.parametrize("homeo,rebound,m", [(0, 0, 12), (1, 0, 12), (2, 1, 12), (1, 0, 40),
                                             (0, 1, 40)])
def test_learn_kernel_random_inputs_bitwise(pkg, homeo, rebound, m):
    """vc_v1_learn_ff against orc_v1_learn_ff on random rows: spikes and
    silent neurons, 0-2 LGN spikes, zero strengths, values past the bounds."""
    import oracle
    from visual_cortex_amd import abi, params
    from visual_cortex_amd.sim import ptr
    _flush(f"test_learn_kernel_random_inputs_bitwise[homeo {homeo}, rebound {rebound}, m {m}]")
    net = _net(applyHomeo=homeo, rebound=rebound, targetFR=(4.0, 4.0), tauAvg=(300.0, 300.0),
               gmaxLGN=1.3, gminLGN=0.3, FF_InfRatio=0.9, symmetricHomeo=0, m=m)
    assert net.sLGN.shape[0] == m
    sp = net.spec
    rng = np.random.default_rng(7 + homeo + m)
    p = params.v1_params_struct(net)
    lp = params.learn_ff_struct(net)
    sL = net.sLGN.astype(np.float32).copy()
    sL[rng.random(sL.shape) < 0.1] = 0.0
    nV, nL = sp.nV1, sp.nLGN
    sInfo = np.where(rng.random(nV) < 0.3, 1.0 + rng.random(nV), -60.0 + rng.random(nV)).astype(np.float32)
    lgn = np.where(rng.random(nL) < 0.4, rng.integers(1, 3, nL) + rng.random(nL), 0.0).astype(np.float32)
    xs = (rng.random(nL) * 3).astype(np.float32)
    xk = (xs + rng.random(nL)).astype(np.float32)
    tr = {k: (rng.random(nV) * 4).astype(np.float32) for k in ("vLTD", "vTrip", "vAvg")}
    tot, inf = params.learn_ff_totals(net)
    tot = (tot * rng.uniform(0.8, 1.2, nV)).astype(np.float32)
    # oracle
    o = oracle.Oracle(abi)
    ho = {k: v.copy() for k, v in tr.items()}
    hs, ht = sL.copy(), tot.copy()
    nl = np.ascontiguousarray(net.nLGN_of, np.uint32)
    il = np.ascontiguousarray(net.idLGN, np.uint32)
    fp = lambda a: a.ctypes.data_as(C.c_void_p)
    assert o.L.orc_v1_learn_ff(C.byref(p), C.byref(lp), fp(nl), fp(hs), fp(il), fp(ho["vLTD"]),
                               fp(ho["vTrip"]), fp(ho["vAvg"]), fp(ht), fp(inf), fp(sInfo),
                               fp(lgn), fp(xs), fp(xk)) == 0
    # device
    d = torch.device("cuda")
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=d)
    g = {k: T(v) for k, v in tr.items()}
    gs, gt, gi = T(sL), T(tot), T(inf)
    gnl, gil = T(nl.view(np.int32)), T(il.view(np.int32))
    ff = abi.FFConn(ptr(gnl), ptr(gs), ptr(gil), sL.shape[0])
    ls = abi.LearnFFState(ptr(gs), ptr(g["vLTD"]), ptr(g["vTrip"]), ptr(g["vAvg"]),
                          ptr(gt) if homeo else None, ptr(gi) if homeo else None)
    # the step's LGN spike bitmap (bit c%64 of word c/64: lgn[c] >= 1)
    nW = (nL + 63) // 64
    spk = np.zeros(nW * 64, bool)
    spk[:nL] = lgn >= 1.0
    bits = np.packbits(spk.reshape(nW, 64), axis=-1, bitorder="little").reshape(-1).view(np.int64)
    rows = [T(x) for x in (sInfo, lgn, xs, xk, bits)]
    lib = abi.load_library()
    abi.check(lib.vc_v1_learn_ff(None, C.byref(p), C.byref(lp), C.byref(ff), C.byref(ls),
                                 *(ptr(r) for r in rows), nL), "vc_v1_learn_ff")
    torch.cuda.synchronize()
    assert (hs != sL).sum() > 100
    assert np.array_equal(_bits(gs.cpu().numpy()), _bits(hs))
    for k in tr:
        assert np.array_equal(_bits(g[k].cpu().numpy()), _bits(ho[k])), k
    if homeo:
        assert np.array_equal(_bits(gt.cpu().numpy()), _bits(ht))


CASES = {
    "plain": dict(),
    "target": dict(targetFR=(5.0, 5.0), tauAvg=(200.0, 200.0)),
    "homeo_mult": dict(applyHomeo=1, FF_InfRatio=0.8, decayTimeHomeoFF=50.0, gmaxLGN=1.25,
                       gminLGN=0.95, A_LGN=0.2),
    "homeo_add_asym": dict(applyHomeo=2, symmetricHomeo=0, FF_InfRatio=1.2, A_LGN=0.2),
    "delay": dict(noDelay=0, speedOfThought=0.5),
    "multi_spike": dict(dt=0.5, tRef=(0.1, 0.08), tonicDep=(1.5, 3.0), nKernelSample=20,
                        tau=20.0),
}


@pytest.mark.parametrize("case", list(CASES))
def test_plastic_run_bitwise(pkg, case):
    import oracle
    from visual_cortex_amd.sim import GpuSim
    net = _net(nt=600, **CASES[case])
    g = GpuSim(net)
    nsteps = max(g.tm["ntPerFrame"] + g.tm["ntPerFrame"] // 2 + 19, 301)
    _flush(f"test_plastic_run_bitwise[{case}]: {nsteps} steps")
    assert g.learn and not g.persistent
    o = oracle.OracleSim(net, pkg, tables=_tables(g))
    s0 = o.sLGN.copy()
    g.run(nsteps // 3)
    g.run(nsteps - nsteps // 3)
    o.run(nsteps)
    sg, so = g.sInfo(), o.sInfo()
    assert (so >= 1).sum() > 300
    diff = np.argwhere(_bits(sg) != _bits(so))
    assert diff.size == 0, f"first sInfo difference at (step, neuron) {diff[0]}"
    ls = g.learn_state_numpy()
    assert (ls["sLGN"] != s0).sum() > 50
    assert np.array_equal(_bits(ls["sLGN"]), _bits(o.sLGN))
    for k in ("vLTD", "vTrip", "vAvg"):
        assert np.array_equal(_bits(ls[k]), _bits(o.lvar[k])), k
    if o.totalFF is not None:
        assert np.array_equal(_bits(ls["totalFF"]), _bits(o.totalFF))
    if case == "multi_spike":
        assert (so >= 2).sum() > 0


def test_lgn_trace_bitwise(pkg):
    """The LGN trace rows of a period and the carried value against the
    oracle's."""
    import oracle
    from visual_cortex_amd.sim import GpuSim
    _flush("test_lgn_trace_bitwise")
    net = _net(nt=600)
    g = GpuSim(net)
    npf = g.tm["ntPerFrame"]
    o = oracle.OracleSim(net, pkg, tables=_tables(g))
    g.run(2 * npf)
    o.run(2 * npf)
    torch.cuda.synchronize()
    buf = 1                                   # period 1's rows
    assert np.array_equal(_bits(g.xStart2[buf].cpu().numpy()), _bits(o.xStart))
    assert np.array_equal(_bits(g.xSpike2[buf].cpu().numpy()), _bits(o.xSpike))
    # the spike bitmaps of the period's steps
    spk = o.lgn_hist[-1] >= 1.0
    nW = g.nWords
    want = np.zeros((npf, nW * 64), bool)
    want[:, :spk.shape[1]] = spk
    want = np.packbits(want.reshape(npf, nW, 64), axis=-1, bitorder="little")   # [npf][nW][8]
    assert spk.sum() > 20
    got = g.lbits2[buf].cpu().numpy().view(np.uint8).reshape(npf, nW, 8)
    assert np.array_equal(got, want)
    # the GPU's carried trace is that of the LGN steps it issued ahead
    o.run(npf)
    g._lgn_issue_to(3 * npf)
    g._lgn_wait(g.s_v1, 3 * npf)
    torch.cuda.synchronize()
    assert np.array_equal(_bits(g.xVar.cpu().numpy()), _bits(o.xVar))


def test_plastic_run_with_outputs_equals_run(pkg, tmp_path):
    """The per-step writer path (the CLI's) learns as run() does."""
    from visual_cortex_amd.sim import GpuSim
    _flush("test_plastic_run_with_outputs_equals_run")
    net = _net(nt=600)
    a = GpuSim(net)
    n = a.tm["ntPerFrame"] + 11
    a.run(n)
    b = GpuSim(net, record=False)
    paths = dict(rawData=str(tmp_path / "raw"), lgn_fr=str(tmp_path / "fr"),
                 lgn_sp=str(tmp_path / "sp"))
    b.run_with_outputs(n, paths, dict(rawData=False, lgn_fr=False, lgn_sp=False))
    la, lb = a.learn_state_numpy(), b.learn_state_numpy()
    for k in ("sLGN", "vLTD", "vTrip", "vAvg"):
        assert np.array_equal(_bits(la[k]), _bits(lb[k])), k


@pytest.mark.parametrize("homeo", [0, 1])
def test_plastic_snapshot_resume(pkg, tmp_path, homeo):
    """A plastic run's snapshot (the strengths, the post variables and the
    LGN trace after the reference's LGN section) resumes bit for bit without
    homeostasis; with it the totals restart from the network's initial
    strengths, as the reference's (it sums them before reading the file)."""
    from visual_cortex_amd import params, snapshot
    from visual_cortex_amd.sim import GpuSim
    _flush(f"test_plastic_snapshot_resume[homeo {homeo}]")
    net = _net(nt=900, applyHomeo=homeo, FF_InfRatio=0.8, A_LGN=0.1)
    sp = net.spec
    a = GpuSim(net)
    npf = a.tm["ntPerFrame"]
    K = max(1, 200 // npf)
    a.run(K * npf)
    path = tmp_path / "snapShot_learn.bin"
    h = snapshot.save(a, path)
    assert h["learning"] == 3
    assert path.stat().st_size == snapshot.file_size(
        sp.nV1, sp.nLGN, sp.iModel, int(a.p.ngFF), int(a.p.ngE), int(a.p.ngI),
        sp.nblock * sp.nI, learn=(sp.nblock, sp.nE, net.sLGN.shape[0]))
    n2 = npf // 2 + 37
    a.run(n2)
    ref = a.sInfo()[K * npf:]
    la = a.learn_state_numpy()
    b = GpuSim(net)
    snapshot.restore(b, path)
    if homeo:
        tot, _ = params.learn_ff_totals(net)
        assert np.array_equal(_bits(b.totalFF.cpu().numpy()), _bits(tot))
        return
    b.run(n2)
    lb = b.learn_state_numpy()
    assert (ref >= 1).sum() > 0
    assert np.array_equal(_bits(b.sInfo()), _bits(ref))
    for k in ("sLGN", "vLTD", "vTrip", "vAvg", "xVar"):
        assert np.array_equal(_bits(la[k]), _bits(lb[k])), k


def test_batched_plastic_trials_equal_serial(pkg):
    """Plastic trials batched into one network (homeostasis on) learn as each
    alone, bit for bit, on the GPU."""
    from visual_cortex_amd import batch
    from visual_cortex_amd.sim import GpuSim
    _flush("test_batched_plastic_trials_equal_serial: 3 trials, 200 steps")
    nets = [_net(nt=600, seed=40 + i, nblock=nb, ori_deg=22.5 * i, applyHomeo=1,
                 FF_InfRatio=0.9) for i, nb in enumerate((2, 4, 3))]
    g = GpuSim(batch.batch_networks(nets, device="cuda"))
    assert g.learn
    g.run(200)
    sb, lb = g.sInfo(), g.learn_state_numpy()
    off = 0
    for t, n in enumerate(nets):
        gs = GpuSim(n)
        gs.run(200)
        k = n.spec.nV1
        ls = gs.learn_state_numpy()
        assert (ls["sLGN"] != n.sLGN).sum() > 20, t
        assert np.array_equal(_bits(sb[:, off:off + k]), _bits(gs.sInfo())), t
        for key in ("sLGN", "vLTD", "vTrip", "vAvg", "totalFF"):
            a = lb[key][..., off:off + k]
            assert np.array_equal(_bits(a), _bits(ls[key])), (t, key)
        off += k
