"""Shared helpers of the test suite (comparisons in the reference's conventions)."""
import numpy as np


def interior(a):
    return a[1:-1, 1:-1, 1:-1]


def rel_maxdiff_nomean(a, b):
    """max|(a - mean a) - (b - mean b)| / max|b - mean b| over the interior: the 1e-12
    pressure criterion after removing the Neumann null-space constant."""
    ia = interior(a) - interior(a).mean()
    ib = interior(b) - interior(b).mean()
    return float(np.abs(ia - ib).max() / np.abs(ib).max())


def packed_rfft(x):
    """numpy rfft repacked in fftpack's RFFTF order [Re0, Re1, Im1, ..., Re_{L/2}]."""
    L = x.shape[-1]
    c = np.fft.rfft(x, axis=-1)
    out = np.zeros_like(x, dtype=np.float64)
    out[..., 0] = c[..., 0].real
    for k in range(1, (L - 1) // 2 + 1):
        out[..., 2 * k - 1] = c[..., k].real
        out[..., 2 * k] = c[..., k].imag
    if L % 2 == 0:
        out[..., L - 1] = c[..., L // 2].real
    return out
