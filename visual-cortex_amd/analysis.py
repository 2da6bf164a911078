"""Rates and orientation tuning from the simulator's output, with the
reference analysis's definitions (src/readPatchOutput.py:291-351 spike
extraction, src/getTuningCurve.py:225-240 preferred orientation,
:1095-1126 aligned tuning curves, :1709-1731 OSI and gOSI).

    fr = tuning_curves([rates(sInfo_k, dt) for sInfo_k in runs])   # [nOri, nV1] Hz
    gosi(fr), osi(align(fr))

Orientation k of nOri is k * 180 / nOri degrees.
"""
from __future__ import annotations

import numpy as np


def spike_counts(sInfo: np.ndarray) -> np.ndarray:
    """Spikes per neuron of a run, sInfo [nt, nV1]: a step with sInfo >= 1
    holds floor(sInfo) spikes (src/readPatchOutput.py:318-336)."""
    s = np.asarray(sInfo)
    return np.where(s >= 1.0, np.floor(s), 0.0).sum(axis=0)


def rates(sInfo: np.ndarray, dt_ms: float) -> np.ndarray:
    """Firing rate of every neuron over the run, in Hz."""
    s = np.asarray(sInfo)
    return spike_counts(s) / (s.shape[0] * dt_ms / 1000.0)


def tuning_curves(rates_by_ori) -> np.ndarray:
    """[nOri, nV1] from one rate vector per orientation."""
    return np.stack([np.asarray(r, dtype=np.float64) for r in rates_by_ori])


def preferred(fr: np.ndarray) -> np.ndarray:
    """Index of the orientation with the highest rate (src/getTuningCurve.py:236)."""
    return np.argmax(fr, axis=0)


def align(fr: np.ndarray, ipref: np.ndarray | None = None) -> np.ndarray:
    """Tuning curves [nV1, nOri+1] rotated so that the preferred orientation
    sits at column nOri//2 (for odd nOri, (nOri-1)//2, shifted by one when
    the next orientation beats the previous one), with column 0 repeated at
    the end (src/getTuningCurve.py:1095-1126)."""
    nOri, n = fr.shape
    if ipref is None:
        ipref = preferred(fr)
    mid = nOri // 2 if nOri % 2 == 0 else (nOri - 1) // 2
    idx = np.empty((n, nOri + 1), dtype=np.int64)
    base = np.arange(nOri)
    for i in range(n):
        row = np.roll(base, mid - int(ipref[i]))
        if nOri % 2 == 1:
            inext = (ipref[i] + 1) % nOri
            iprev = (ipref[i] + nOri - 1) % nOri
            if fr[iprev, i] > fr[inext, i]:
                row = np.roll(row, 1)
        idx[i, :nOri] = row
        idx[i, nOri] = row[0]
    return fr[idx, np.arange(n)[:, None]]


def osi(aligned: np.ndarray) -> np.ndarray:
    """(r_pref - r_orth) / (r_pref + r_orth) on aligned curves [n, nOri+1],
    0 where the preferred rate is 0 (src/getTuningCurve.py:1721-1731)."""
    nOri = aligned.shape[1] - 1
    lo = aligned[:, 0]
    hi = aligned[:, nOri // 2] if nOri % 2 == 0 else aligned[:, (nOri - 1) // 2]
    out = np.zeros(aligned.shape[0])
    pick = hi > 0
    out[pick] = (hi[pick] - lo[pick]) / (lo[pick] + hi[pick])
    return out


def gosi(fr: np.ndarray) -> np.ndarray:
    """Global OSI |sum_k r_k e^{2 i theta_k}| / sum_k r_k (one minus the
    circular variance) of curves [nOri, n], 0 for silent neurons
    (src/getTuningCurve.py:1709-1719)."""
    nOri, n = fr.shape
    ph = np.exp(1j * 2.0 * np.pi * np.arange(nOri) / nOri)
    out = np.zeros(n)
    pick = fr.max(axis=0) > 0
    num = np.abs((fr * ph[:, None]).sum(axis=0))
    out[pick] = num[pick] / fr.sum(axis=0)[pick]
    return out
