"""Pins of the float64 LGN loss gradient of the oracle (the reference the GPU
adjoint is checked against) by central finite differences of the float64
loss itself, on a small LGN layer, over stimulus pixels and every kind of
per-cell parameter.  Coordinates whose +-eps runs change the activation
pattern (a clipped contrast sample or a rectified step) are skipped: the
loss has kinks there."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def case(pkg, orc):
    import oracle as O
    from visual_cortex_amd import network as N
    spec = N.spec_tiny(nt=300)
    lgn = N.lgn_layer(spec)
    tab = O.lgn_tables(orc, spec, lgn, N)
    g64 = O.LgnGrad64(orc, spec, lgn, tab, 300, N)
    frames = N.white_noise_frames(spec, g64.nF, seed=1)
    th = g64.theta0.astype(np.float64)
    tgt = g64.rates(N.grating_frames(spec, g64.nF), th).astype(np.float32)
    df, dth = g64.grad(frames, th, tgt)
    return g64, frames, th, tgt, df, dth


def test_oracle_lgn_grad_pixels(case):
    g64, frames, th, tgt, df, dth = case
    rng = np.random.default_rng(0)
    nz = np.argwhere(df != 0)
    assert len(nz) > 1000
    eps, checked = 1e-5, 0
    for (f, ch, y, x) in nz[rng.choice(len(nz), 40, replace=False)]:
        cells = g64.cells_touching(x, y, ch)
        fp = frames.astype(np.float64); fp[f, ch, y, x] += eps
        fm = frames.astype(np.float64); fm[f, ch, y, x] -= eps
        lp, pp = g64.loss(fp, th, tgt, cells)
        lm, pm = g64.loss(fm, th, tgt, cells)
        if pp != pm:
            continue
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - df[f, ch, y, x]) <= 1e-6 * abs(fd) + 1e-9 * abs(df).max()
        checked += 1
    assert checked >= 36


@pytest.mark.parametrize("k", range(6))
def test_oracle_lgn_grad_theta(case, k):
    g64, frames, th, tgt, df, dth = case
    rng = np.random.default_rng(k)
    eps, checked = 1e-5, 0
    for c in rng.choice(th.shape[0], 16, replace=False):
        tp = th.copy(); tp[c, k] += eps
        tm = th.copy(); tm[c, k] -= eps
        lp, pp = g64.loss(frames, tp, tgt, [c])
        lm, pm = g64.loss(frames, tm, tgt, [c])
        if pp != pm:
            continue
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - dth[c, k]) <= 1e-6 * abs(fd) + 1e-9 * abs(dth[:, k]).max()
        checked += 1
    assert checked >= 14


def test_oracle_lgn_loss_is_sum_over_cells(case):
    """The cell-restricted loss used by the finite differences adds up to the
    full loss, and the rates are the per-cell quantities it is built from."""
    g64, frames, th, tgt, df, dth = case
    full, _ = g64.loss(frames, th, tgt)
    n = th.shape[0]
    parts = sum(g64.loss(frames, th, tgt, [c])[0] for c in range(n))
    assert parts == pytest.approx(full, rel=1e-12)
    fr = g64.rates(frames, th)
    assert np.mean((fr - tgt.astype(np.float64)) ** 2) == pytest.approx(full, rel=1e-12)
