"""Pins of the float64 V1 reference used to check the GPU spike-time
adjoint: (1) its forward is the same dynamics as the float32 oracle (same
spike pattern, sInfo within float32 rounding drift); (2) its tape gradient
matches central finite differences of the float64 loss at a fixed spike
pattern, for recurrent and LGN->V1 weights."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def case(pkg, orc):
    import oracle as O
    from visual_cortex_amd import network as N, params as P
    spec = N.lif(N.spec_tiny(nt=160, tonicDep=(0.98, 0.98), bgStrength=(0.3, 0.3),
                             frRatioLGN=200.0, sLGN=3.0))
    net = N.build_network(spec)
    T = 160
    osim = O.OracleSim(net, pkg)
    osim.run(T)
    s32, lgn = osim.sInfo(), osim.lgn_sinfo()[:T]
    st = P.initial_state(net)
    init = np.zeros((14, spec.nV1), np.float32)
    init[0], init[1] = st["v"], st["tBack"]
    g64 = O.V1Grad64(orc, net, P, lgn, init)
    return spec, g64, s32


def test_v1_f64_forward_is_the_oracle_dynamics(case):
    spec, g64, s32 = case
    s64 = g64.run()
    assert ((s64 >= 1) == (s32 >= 1)).all()
    assert (s32 >= 1).sum() > 200                       # the spike paths are exercised
    spk = s32 >= 1
    # float32 rounding drifts through the recurrent network: spike phases
    # (fractions of dt) and voltages agree closely, not bitwise
    dph = np.abs(s64[spk] - s32[spk])
    assert np.percentile(dph, 99) < 1e-3 and dph.max() < 0.05
    dv = np.abs(s64[~spk] - s32[~spk])
    assert np.percentile(dv, 99) < 1e-2 and dv.max() < 0.5      # mV


def test_v1_tape_gradient_finite_differences(case):
    import oracle as O
    spec, g64, s32 = case
    pat = s32 >= 1
    dt, T = spec.dt, s32.shape[0]
    tgt = np.random.default_rng(0).random((T // 40, spec.nV1)) * 0.5
    s64 = g64.run(pattern=pat)
    L, gOut = O.soft_count_loss64(s64, tgt, dt, 40, 1.0)
    _, dsl, dcm = g64.run(pattern=pat, gOut=gOut)
    rng = np.random.default_rng(1)
    eps = 1e-6
    checked = 0
    for kind, arr, grad in (("w", g64.conMat, dcm), ("s", g64.sLGN, dsl)):
        nz = np.argwhere(grad != 0)
        assert len(nz) > 50
        for idx in map(tuple, nz[rng.choice(len(nz), 16, replace=False)]):
            vals = []
            for sgn in (1, -1):
                a = arr.copy()
                a[idx] += sgn * eps
                out = g64.run(pattern=pat, **({"conMat": a} if kind == "w" else {"sLGN": a}))
                vals.append(O.soft_count_loss64(out, tgt, dt, 40, 1.0)[0])
            fd = (vals[0] - vals[1]) / (2 * eps)
            assert abs(grad[idx] - fd) <= 1e-4 * abs(fd) + 1e-7 * np.abs(grad).max(), (kind, idx)
            checked += 1
    assert checked == 32


def test_soft_count_loss_gradient(case):
    """d loss / d sInfo of the float64 loss against finite differences."""
    import oracle as O
    spec, g64, s32 = case
    s = s32.astype(np.float64)
    T = s.shape[0]
    tgt = np.random.default_rng(2).random((T // 40, spec.nV1))
    L, g = O.soft_count_loss64(s, tgt, spec.dt, 40, 1.0)
    it, n = np.nonzero(s >= 1)
    for k in np.random.default_rng(3).choice(it.size, 10, replace=False):
        e = 1e-5
        sp_ = s.copy(); sp_[it[k], n[k]] += e
        sm_ = s.copy(); sm_[it[k], n[k]] -= e
        fd = (O.soft_count_loss64(sp_, tgt, spec.dt, 40, 1.0)[0] -
              O.soft_count_loss64(sm_, tgt, spec.dt, 40, 1.0)[0]) / (2 * e)
        assert g[it[k], n[k]] == pytest.approx(fd, rel=1e-5, abs=1e-12)


def test_v1_tangent_mode_equals_tape_and_finite_differences(case):
    """The forward-mode derivative (dual numbers) along a random direction of
    all weights equals the tape's reverse-mode gradient dotted with that
    direction, and the central difference of the float64 loss along it; with
    the trajectory's own states as the anchor it is unchanged, and anchored to
    the float32 trajectory it stays close over this short run."""
    import oracle as O
    spec, g64, s32 = case
    pat = s32 >= 1
    dt, T = spec.dt, s32.shape[0]
    tgt = np.random.default_rng(0).random((T // 40, spec.nV1)) * 0.5
    s64 = g64.run(pattern=pat)
    L, gOut = O.soft_count_loss64(s64, tgt, dt, 40, 1.0)
    _, dsl, dcm = g64.run(pattern=pat, gOut=gOut)
    rng = np.random.default_rng(4)
    uc = rng.standard_normal(g64.conMat.shape) * (g64.conMat != 0) * np.abs(g64.conMat)
    us = rng.standard_normal(g64.sLGN.shape) * np.abs(g64.sLGN)
    dOut = g64.tangent(pat, uConMat=uc, uSLGN=us)
    fwd = float((gOut * dOut).sum())
    rev = float((dcm * uc).sum() + (dsl * us).sum())
    assert abs(fwd) > 1e-6
    assert fwd == pytest.approx(rev, rel=1e-9)
    e = 1e-7
    lp = O.soft_count_loss64(g64.run(conMat=g64.conMat + e * uc, sLGN=g64.sLGN + e * us,
                                     pattern=pat), tgt, dt, 40, 1.0)[0]
    lm = O.soft_count_loss64(g64.run(conMat=g64.conMat - e * uc, sLGN=g64.sLGN - e * us,
                                     pattern=pat), tgt, dt, 40, 1.0)[0]
    assert fwd == pytest.approx((lp - lm) / (2 * e), rel=1e-5)
    # anchored to the float32 trajectory (the oracle's state at the start of
    # every step): the same spike pattern, a trajectory within float32
    # rounding drift of the float64 one, hence nearly the same derivative
    anchor = _float32_trajectory(g64, T)
    d_anc = g64.tangent(pat, uConMat=uc, uSLGN=us, anchor=anchor)
    fa = float((gOut * d_anc).sum())
    assert fa == pytest.approx(fwd, rel=2e-3)
    assert not np.array_equal(d_anc, dOut)          # the anchor was applied


def _float32_trajectory(g64, T):
    """[T][14][nV1]: the float32 oracle's state at the start of every step
    (snapshot rows; the cortical conductances after the step's recurrent
    increments, which the oracle adds at the end of the step before)."""
    import oracle as O
    import importlib
    pkg = importlib.import_module("visual_cortex_amd")
    net = g64.net
    nV1 = net.spec.nV1
    o = O.OracleSim(net, pkg)
    out = np.zeros((T, 14, nV1), np.float32)
    rows = (("v", 0, 1), ("tBack", 1, 1), ("gFF", 2, 2), ("hFF", 4, 2), ("gE", 6, 2),
            ("hE", 8, 2), ("gI", 10, 2), ("hI", 12, 2))
    for it in range(T):
        for k, r0, nr in rows:
            a = o.st[k].reshape(-1, nV1)
            out[it, r0:r0 + min(nr, a.shape[0])] = a[:nr]
        o.run(1, record=False)
    return out
