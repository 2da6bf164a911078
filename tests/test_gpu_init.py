"""GPU parity of the initial state: rand_spInit's initial spikes and ring
(src/coredynamics.cu:5-146) through vc_v1_initial_spikes, the spike buffer
of a step's outputs (vc_spikes_compact), whole runs that start with initial
spikes (the pre-loop recal_G of src/patch.cu:5687-5770: the dense gather
without delays, the delayed gather with them, the gap sums of InhGap), and
a run from the reference's initial-state draws (src/patch.cu:4405-4570)."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _tables(g):
    return dict(TW=g.TW.cpu().numpy(), SW=g.SW.cpu().numpy(), SCx=g.SCx.cpu().numpy(),
                SCy=g.SCy.cpu().numpy(), maxConvol=g.maxConvol.cpu().numpy())


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


SP0 = dict(spE0=(40.0, 25.0), spI0=(60.0, 30.0), SCsplit=4)


def _net(**kw):
    from visual_cortex_amd import network
    base = dict(bgStrength=(0.4, 0.4), nblock=4, **SP0)
    base.update(kw)
    return network.build_network(network.spec_tiny(**base))


def test_initial_spikes_kernel_bitwise(pkg):
    """Ring (depth 9 > trainDepth 6, step 0 in ring slot 5), masks and the
    v/w/tBack of spiking neurons (AdEx) against the oracle; chances 0.5
    per step so every chain position holds spikes."""
    import oracle
    from visual_cortex_amd import abi, params
    from visual_cortex_amd.sim import ptr
    _flush("test_initial_spikes_kernel_bitwise")
    net = _net()
    sp = net.spec
    p = params.v1_params_struct(net)
    q = abi.V1SpInit()
    for i, c in enumerate((0.5, 0.3, 0.6, 0.4)):
        q.sp0[i] = c
    q.SCsplit, q.depthRef = 4, 6
    depth, cur0 = 9, 5
    st = params.initial_state(net)
    st["w"][:] = np.linspace(0.0, 2.0, sp.nV1)
    d = torch.device("cuda")
    T = {k: torch.as_tensor(np.ascontiguousarray(v), device=d) for k, v in st.items()}
    tonic = torch.as_tensor(net.tonicDep, device=d)
    s = abi.V1State(*(ptr(T[k]) for k in ("v", "w", "tBack", "gFF", "hFF", "gE", "hE", "gI",
                                          "hI")), ptr(tonic), None, None, None)
    nl = torch.as_tensor(net.nLGN_of.astype(np.int32), device=d)
    train = torch.zeros(depth, sp.nV1, device=d)
    mask = torch.zeros(depth, sp.nV1 // 64, dtype=torch.int64, device=d)
    lib = abi.load_library()
    abi.check(lib.vc_v1_initial_spikes(None, C.byref(p), C.byref(s), ptr(nl), C.byref(q),
                                       ptr(train), ptr(mask), depth, cur0),
              "vc_v1_initial_spikes")
    torch.cuda.synchronize()
    o = oracle.Oracle(abi)
    v, w, tb = st["v"].copy(), st["w"].copy(), st["tBack"].copy()
    tr = np.zeros((depth, sp.nV1), np.float32)
    nlh = np.ascontiguousarray(net.nLGN_of, np.uint32)
    assert o.L.orc_v1_initial_spikes(C.byref(p), nlh.ctypes.data, C.byref(q), v.ctypes.data,
                                     w.ctypes.data, tb.ctypes.data, tr.ctypes.data, depth,
                                     cur0) == 0
    assert (tr >= 1).sum() > 1000
    assert np.array_equal(_bits(train.cpu().numpy()), _bits(tr))
    for name, a in (("v", v), ("w", w), ("tBack", tb)):
        assert np.array_equal(_bits(T[name].cpu().numpy()), _bits(a)), name
    bits = (tr >= 1).reshape(depth, -1, 64)
    want = (bits * (np.uint64(1) << np.arange(64, dtype=np.uint64))).sum(-1, dtype=np.uint64)
    assert np.array_equal(mask.cpu().numpy().view(np.uint64), want)


def test_spikes_compact_equals_step_compaction(pkg):
    """vc_spikes_compact of a step's outputs gives the spike buffer the step
    wrote (masks and increments bit for bit), several spikes per neuron
    included."""
    from visual_cortex_amd import abi
    from visual_cortex_amd.sim import GpuSim, ptr
    _flush("test_spikes_compact_equals_step_compaction")
    net = _net(spE0=(0.0, 0.0), spI0=(0.0, 0.0), dt=0.5, tRef=(0.1, 0.08), tonicDep=(3.0, 12.0),
               tau=20.0, nKernelSample=20)
    g = GpuSim(net, persistent=False)
    sp = net.spec
    for _ in range(120):                     # up to a step with a neuron firing twice
        g.run(1)
        torch.cuda.synchronize()
        out = g.sInfo_last.clone()
        if int((out >= 2).sum()) > 0:
            break
    assert int((out >= 2).sum()) > 0
    mk = torch.zeros(sp.nV1 // 64, dtype=torch.int64, device="cuda")
    og = torch.zeros(abi.VC_MAX_NG, sp.nV1, device="cuda")
    oh = torch.zeros(abi.VC_MAX_NG, sp.nV1, device="cuda")
    buf = abi.Spikes(ptr(mk), ptr(og), ptr(oh))
    abi.check(g.lib.vc_spikes_compact(None, C.byref(g.p), ptr(out), C.byref(buf)),
              "vc_spikes_compact")
    torch.cuda.synchronize()
    assert torch.equal(mk, g.masks[g.cur])
    m = mk.cpu().numpy().view(np.uint64)
    nsp = sum(bin(int(x)).count("1") for x in m)
    assert nsp > 0
    # compare the increments at the occupied slots only (the rest is stale)
    occ = np.zeros(sp.nV1, bool)
    for t_, x in enumerate(m):
        r = bin(int(x)).count("1")
        occ[t_ * 64:t_ * 64 + r] = True
    for a, b in ((og, g.og[g.cur]), (oh, g.oh[g.cur])):
        for ig in range(2):
            assert np.array_equal(_bits(a[ig].cpu().numpy()[occ]), _bits(b[ig].cpu().numpy()[occ]))


CASES = {
    "nodelay": dict(),
    "nodelay_adex_lif": dict(spE0=(200.0, 100.0), spI0=(300.0, 150.0)),
    "delay": dict(noDelay=0, speedOfThought=0.5),
    "gap": dict(nGap=6, InhGap=1),
    "gap_delay": dict(nGap=6, InhGap=1, noDelay=0, speedOfThought=0.5),
}


# the persistent launch runs without delays and gap junctions only
@pytest.mark.parametrize("case,persistent", [(c, False) for c in CASES] + [("nodelay", True)])
def test_run_with_initial_spikes_bitwise(pkg, case, persistent):
    import oracle
    from visual_cortex_amd.sim import GpuSim
    nsteps = 96
    _flush(f"test_run_with_initial_spikes_bitwise[{case}, persistent {persistent}]: {nsteps} steps")
    net = _net(**CASES[case])
    g = GpuSim(net, persistent=persistent)
    o = oracle.OracleSim(net, pkg, tables=_tables(g))
    for k in ("v", "w", "tBack"):                             # after rand_spInit
        assert np.array_equal(_bits(g.state[k].cpu().numpy()), _bits(o.st[k])), k
    assert (o.st["tBack"] > 0).sum() > 0
    g.run(nsteps)
    o.run(nsteps)
    sg, so = g.sInfo(), o.sInfo()
    diff = np.argwhere(_bits(sg) != _bits(so))
    assert diff.size == 0, f"first sInfo difference at (step, neuron) {diff[0]}"
    # the initial spikes matter: the same network without them runs otherwise
    from dataclasses import replace
    from visual_cortex_amd import network
    net0 = network.build_network(replace(net.spec, spE0=(0.0, 0.0), spI0=(0.0, 0.0)))
    o0 = oracle.OracleSim(net0, pkg, tables=_tables(g))
    o0.run(4)
    assert not np.array_equal(_bits(o0.sInfo()[:4]), _bits(so[:4]))


def test_run_from_reference_initial_state_bitwise(pkg):
    """The reference's initial-state draws (w0, gFF0, gE0, gI0, use_v0): the
    GPU run from them equals the oracle's."""
    import oracle
    from visual_cortex_amd.sim import GpuSim
    _flush("test_run_from_reference_initial_state_bitwise")
    init = dict(init_w0=(0.2, 0.05, 0.0, 0.0),
                init_gFF0=(0.05, 0.5, 0.02, 0.5, 0.01, 0.5, 0.005, 0.5) * 2,
                init_gE0=(0.05, 0.5, 0.02, 0.5, 0.01, 0.5, 0.005, 0.5) * 2,
                init_gI0=(0.1, 0.5, 0.05, 0.5) * 2, use_v0=1)
    net = _net(spE0=(0.0, 0.0), spI0=(0.0, 0.0), **init)
    g = GpuSim(net)
    o = oracle.OracleSim(net, pkg, tables=_tables(g))
    assert float(np.abs(o.st["gE"]).max()) > 0
    g.run(80)
    o.run(80)
    assert np.array_equal(_bits(g.sInfo()), _bits(o.sInfo()))
