"""The allocation-free persistent entry point (vc_v1_run_persistent_noalloc)
against vc_v1_run_persistent, through the C ABI: two runs in a row on one
GPU, the second on a smaller network, each compared bit for bit with the
entry point that takes a caller buffer (every V1 output of every step and the
final state).  Its sizes limit is checked from the host."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(net, noalloc, nsteps):
    from visual_cortex_amd.sim import GpuSim
    g = GpuSim(net, noalloc=noalloc)
    g.run(nsteps)
    assert g.persistent, "the persistent launch was not used"
    return g.sInfo(), g.state_numpy()


@pytest.mark.parametrize("order", ["large_then_small"])
def test_noalloc_matches_caller_buffer(pkg, order):
    from visual_cortex_amd import network
    cases = [(dict(nblock=4, bgStrength=(0.4, 0.4)), 300),     # 1024 neurons, 300 steps
             (dict(nblock=2, bgStrength=(0.4, 0.4)), 200)]     # smaller, second
    for kw, nsteps in cases:
        print(f"test_noalloc_matches_caller_buffer: spec_tiny {kw}, {nsteps} steps", flush=True)
        net = network.build_network(network.spec_tiny(**kw))
        s_ref, st_ref = _run(net, False, nsteps)
        s_na, st_na = _run(net, True, nsteps)
        nsp = int((s_ref >= 1).sum())
        assert nsp > 100, f"only {nsp} spikes: the comparison would be vacuous"
        assert np.array_equal(s_ref.view(np.uint32), s_na.view(np.uint32))
        for k in st_ref:
            assert np.array_equal(st_ref[k].view(np.uint32), st_na[k].view(np.uint32)), k


def test_noalloc_refuses_too_many_neurons(pkg):
    from visual_cortex_amd import abi
    lib = abi.load_library()
    mx = int(lib.vc_v1_run_persistent_noalloc_max_v1())
    assert mx == 131072
    p = abi.V1Params()
    p.nV1, p.nblock, p.blockSize = mx + 1024, (mx + 1024) // 1024, 1024
    rc = lib.vc_v1_run_persistent_noalloc(None, C.byref(p), None, None, None, None, 0, None, 0,
                                          None, 0, 0, 1, None, 0, None)
    assert rc == abi.VC_E_INVALID
    assert b"exceeds" in lib.vc_last_error()
