"""Rates and orientation-tuning measures (CPU), pinned by closed forms:
for r(theta) = b + a cos 2(theta - theta0) sampled at nOri orientations,
gOSI = a / 2b exactly (the second circular moment of the samples), and with
theta0 on the grid the aligned curve peaks at nOri//2 with OSI = a / b."""
import numpy as np
import pytest


def _cos_curves(nOri, b, a, pref):
    th = np.arange(nOri) * np.pi / nOri
    return np.stack([b + a * np.cos(2 * (th - pref[i] * np.pi / nOri))
                     for i in range(len(pref))], axis=1)          # [nOri, n]


@pytest.mark.parametrize("nOri", [8, 12, 7])
def test_gosi_and_osi_closed_forms(pkg, nOri):
    from visual_cortex_amd import analysis as A
    pref = np.array([0, 1, nOri - 1, nOri // 2, 3])
    b, a = 10.0, 4.0
    fr = _cos_curves(nOri, b, a, pref)
    np.testing.assert_allclose(A.gosi(fr), a / (2 * b), rtol=1e-12)
    assert np.array_equal(A.preferred(fr), pref)
    al = A.align(fr)
    assert al.shape == (len(pref), nOri + 1)
    mid = nOri // 2 if nOri % 2 == 0 else (nOri - 1) // 2
    np.testing.assert_allclose(al[:, mid], b + a)                 # preferred in the middle
    np.testing.assert_allclose(al[:, -1], al[:, 0])               # wrap column
    if nOri % 2 == 0:
        np.testing.assert_allclose(A.osi(al), a / b, rtol=1e-12)  # orthogonal at column 0
    # untuned and silent neurons
    flat = np.full((nOri, 2), 5.0)
    flat[:, 1] = 0.0
    np.testing.assert_allclose(A.gosi(flat), 0.0, atol=1e-12)
    assert A.osi(A.align(flat))[1] == 0.0


def test_rates_count_multi_spike_steps(pkg):
    from visual_cortex_amd import analysis as A
    s = np.full((800, 3), -60.0, np.float32)      # 100 ms at dt 0.125
    s[10, 0] = 1.25                                # one spike
    s[20, 1] = 2.5                                 # two spikes in one step
    s[30, 1] = 1.0
    np.testing.assert_array_equal(A.spike_counts(s), [1, 3, 0])
    np.testing.assert_allclose(A.rates(s, 0.125), [10.0, 30.0, 0.0])
