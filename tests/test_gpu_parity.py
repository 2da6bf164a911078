"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs.  Integer results (spike counts, spike steps, compaction)
must be identical; with identical LGN tables the whole per-step path is
bit-identical by construction (same binary32 operation sequences), which is
what these tests assert.  The one-time LGN tables use the platform
transcendentals and are compared with a relative tolerance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def tiny(pkg):
    from visual_cortex_amd import network
    return network.build_network(network.spec_tiny())


def test_lgn_tables_match_oracle(pkg, tiny):
    import oracle
    from visual_cortex_amd.sim import GpuSim
    _flush("test_lgn_tables_match_oracle: tiny spec")
    g = GpuSim(tiny)
    o = oracle.OracleSim(tiny, pkg)
    for k, gt in (("TW", g.TW), ("SW", g.SW), ("SCx", g.SCx), ("SCy", g.SCy), ("maxConvol", g.maxConvol)):
        a = gt.cpu().numpy()
        b = o.tables[k]
        np.testing.assert_allclose(a, b, rtol=2e-6, atol=2e-7, err_msg=k)


def test_frame_response_bitwise(pkg, tiny):
    import oracle
    from visual_cortex_amd.sim import GpuSim
    _flush("test_frame_response_bitwise: tiny spec")
    g = GpuSim(tiny)
    tables = dict(TW=g.TW.cpu().numpy(), SW=g.SW.cpu().numpy(), SCx=g.SCx.cpu().numpy(),
                  SCy=g.SCy.cpu().numpy(), maxConvol=g.maxConvol.cpu().numpy())
    o = oracle.OracleSim(tiny, pkg, tables=tables)
    for f in range(3):
        g._frame_response(g.frames[f], 1)
        Rg = g.Rring[1].cpu().numpy()
        Ro = o.frame_response(tiny.frames[f])
        assert np.array_equal(Rg.view(np.uint32), Ro.view(np.uint32)), f


@pytest.mark.parametrize("n1d", [8, 16])
def test_frame_response_geometry_equals_tables(pkg, n1d):
    """The frame responses computed from the per-region sample geometry
    (coordinates regenerated in the kernel) equal those read from the SCx/SCy
    tables and the oracle's bit for bit, for 64 and 256 samples."""
    import oracle
    import torch
    from visual_cortex_amd import network
    from visual_cortex_amd.abi import check
    from visual_cortex_amd.sim import GpuSim, ptr
    _flush(f"test_frame_response_geometry_equals_tables: {n1d}x{n1d} samples")
    net = network.build_network(network.spec_tiny(nSpatialSample1D=n1d, nLGN_side=12))
    g = GpuSim(net)
    sp = net.spec
    tables = dict(TW=g.TW.cpu().numpy(), SW=g.SW.cpu().numpy(), SCx=g.SCx.cpu().numpy(),
                  SCy=g.SCy.cpu().numpy(), maxConvol=g.maxConvol.cpu().numpy())
    o = oracle.OracleSim(net, pkg, tables=tables)
    Rt = torch.zeros(sp.nLGN, 2, device="cuda")
    for f in range(3):
        check(g.lib.vc_lgn_frame_response(None, ptr(g.frames[f]), sp.width, sp.height, sp.nLGN,
                                          n1d, ptr(g.SCx), ptr(g.SCy), ptr(g.SW),
                                          ptr(g.coneType), None, 0, ptr(Rt)),
              "vc_lgn_frame_response")
        g._frame_response(g.frames[f], 1)                     # the geometry path
        torch.cuda.synchronize()
        Rg = g.Rring[1].cpu().numpy()
        assert np.abs(Rg).max() > 0
        assert np.array_equal(Rg.view(np.uint32), Rt.cpu().numpy().view(np.uint32)), f
        Ro = o.frame_response(net.frames[f])
        assert np.array_equal(Rg.view(np.uint32), Ro.view(np.uint32)), f
    # a partial range of cells writes those cells only
    Rp = torch.full((sp.nLGN, 2), -7.0, device="cuda")
    frame = g.frames[2]
    check(g.lib.vc_lgn_frame_response_geom(None, ptr(frame), sp.width, sp.height, sp.nLGN, 5, 77,
                                           n1d, ptr(g.geom), *g.view, ptr(g.SW), ptr(g.coneType),
                                           None, 0, ptr(Rp)), "vc_lgn_frame_response_geom")
    torch.cuda.synchronize()
    Rp = Rp.cpu().numpy()
    assert np.array_equal(Rp[5:77].view(np.uint32), Rg[5:77].view(np.uint32))
    assert (Rp[:5] == -7.0).all() and (Rp[77:] == -7.0).all()


CASES = {
    # name: (spec factory kwargs, steps)
    "tiny": (dict(), 400),
    # synchronous start under strong tonic drive: hundreds of spikes per
    # neighbourhood in one step (several 64-spike batch flushes)
    "burst": (dict(tonicDep=(2.0, 2.0), sE=1.5), 200),
    # E/I boundary inside a 64-neuron tile (mixed tiles), 3 blocks
    "mixed_tile": (dict(nE=200, nblock=3, bgStrength=(0.4, 0.4)), 300),
    # one block of 64 neurons, one tile
    "one_tile": (dict(nblock=1, blockSize=64, nE=48, KE=20, KI=5, bgStrength=(0.4, 0.4)), 300),
    # a temporal kernel of 32 taps (the rate kernel's generic tap count)
    "taps32": (dict(nKernelSample=32, bgStrength=(0.4, 0.4)), 200),
}


@pytest.mark.parametrize("case,events,persistent,dense",
                         [(c, True, True, True) for c in CASES if c != "taps32"]
                         + [(c, True, False, False) for c in CASES]
                         + [("tiny", True, True, False), ("burst", True, False, True),
                            ("tiny", False, True, True), ("burst", False, True, False),
                            ("tiny", False, False, False), ("burst", False, False, True)])
def test_full_path_bitwise(pkg, case, events, persistent, dense):
    import oracle
    from visual_cortex_amd import network
    from visual_cortex_amd.sim import GpuSim
    kw, nsteps = CASES[case]
    _flush(f"test_full_path_bitwise[{case}, events={events}, persistent={persistent}, "
           f"dense={dense}]: {kw} {nsteps} steps")
    net = network.build_network(network.spec_tiny(**kw))
    state = None
    if case == "burst":
        from visual_cortex_amd import params
        state = params.initial_state(net)
        state["v"][:] = -52.0
    g = GpuSim(net, state=state, ff_events=events, persistent=persistent, dense_weights=dense)
    tables = dict(TW=g.TW.cpu().numpy(), SW=g.SW.cpu().numpy(), SCx=g.SCx.cpu().numpy(),
                  SCy=g.SCy.cpu().numpy(), maxConvol=g.maxConvol.cpu().numpy())
    o = oracle.OracleSim(net, pkg, tables=tables, state=state)
    g.run(nsteps)
    o.run(nsteps)
    assert g.persistent == persistent          # the requested V1 launch kind ran
    lg, lo = g.lgn_sinfo(), o.lgn_sinfo()
    assert np.array_equal(lg.view(np.uint32), lo.view(np.uint32)), "LGN sInfo differs"
    sg, so = g.sInfo(), o.sInfo()
    nsp_g, nsp_o = (sg >= 1).sum(), (so >= 1).sum()
    per_step = (so >= 1).sum(1)
    _flush(f"  V1 spikes gpu={nsp_g} oracle={nsp_o}, max per step {per_step.max()}")
    assert nsp_o > 0, "no V1 spikes: the test would not exercise the recurrent path"
    if case == "burst":
        assert per_step.max() > 64
    diff = np.argwhere(sg.view(np.uint32) != so.view(np.uint32))
    assert diff.size == 0, f"first sInfo difference at (step, neuron) {diff[0]}"
    g.finalize()
    sgs = g.state_numpy()
    for k in ("v", "w", "tBack", "gFF", "hFF", "gE", "hE", "gI", "hI"):
        assert np.array_equal(sgs[k].view(np.uint32), o.st[k].view(np.uint32)), k


def test_c1_first_200_steps_bitwise(pkg):
    """The small plain-forward workload (4 blocks x 1024, 1024 LGN cells,
    128x128 px) for its first 200 steps."""
    import oracle
    from visual_cortex_amd import network
    from visual_cortex_amd.sim import GpuSim
    _flush("test_c1_first_200_steps_bitwise")
    net = network.build_network(network.spec_c1())
    g = GpuSim(net)
    tables = dict(TW=g.TW.cpu().numpy(), SW=g.SW.cpu().numpy(), SCx=g.SCx.cpu().numpy(),
                  SCy=g.SCy.cpu().numpy(), maxConvol=g.maxConvol.cpu().numpy())
    o = oracle.OracleSim(net, pkg, tables=tables)
    g.run(200)
    o.run(200)
    assert g.persistent
    sg, so = g.sInfo(), o.sInfo()
    assert (so >= 1).sum() > 0
    assert np.array_equal(sg.view(np.uint32), so.view(np.uint32))


def test_short_runs_bitwise(pkg):
    """Runs of a few steps at a time (the LGN stage issued per run, in chunks
    shorter than a frame period: the lane-per-cell rate kernel) give the
    oracle's V1 outputs and LGN state bit for bit."""
    import torch
    import oracle
    from visual_cortex_amd import network
    from visual_cortex_amd.sim import GpuSim
    _flush("test_short_runs_bitwise: tiny, runs of 5/20/7/32/13/1/64/... steps")
    net = network.build_network(network.spec_tiny(bgStrength=(0.4, 0.4)))
    sp = net.spec
    g = GpuSim(net, record=False)
    tables = dict(TW=g.TW.cpu().numpy(), SW=g.SW.cpu().numpy(), SCx=g.SCx.cpu().numpy(),
                  SCy=g.SCy.cpu().numpy(), maxConvol=g.maxConvol.cpu().numpy())
    o = oracle.OracleSim(net, pkg, tables=tables)
    sizes = [5, 20, 7, 32, 13, 1, 64, 30, 2, 26, 56]      # 256 steps: the LGN state of
    assert sum(sizes) % g.tm["ntPerFrame"] == 0               # both sides is at a period end
    outs = []
    for n in sizes:
        out = torch.empty(n, sp.nV1, dtype=torch.float32, device="cuda")
        g.run(n, out=out)
        outs.append(out)
    T = sum(sizes)
    o.run(T)
    sg = torch.cat(outs).cpu().numpy()
    so = o.sInfo()
    assert int((so >= 1).sum()) > 100
    diff = np.argwhere(sg.view(np.uint32) != so.view(np.uint32))
    assert diff.size == 0, f"first sInfo difference at (step, neuron) {diff[0]}"
    torch.cuda.synchronize()
    assert np.array_equal(g.lTR.cpu().numpy().view(np.uint32), o.lTR.view(np.uint32))
    assert np.array_equal(g.lNL.cpu().numpy().view(np.uint32), o.lNL.view(np.uint32))
