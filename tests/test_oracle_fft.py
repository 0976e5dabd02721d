"""The oracle's fftpack-convention real FFT (src/fftpack.f:140-240) pinned against the
reference's own RFFTF/RFFTB outputs (golden fixture written from oracle/_ref, the
reference's fftpack.f compiled from its sources), numpy's rfft and a brute-force
long-double DFT."""
import os

import numpy as np
import pytest

import oracle
from helpers import packed_rfft

GOLD = os.path.join(os.path.dirname(__file__), "golden", "fftpack_rfft.npz")


def _lengths(z):
    return sorted(int(k.split("_")[1]) for k in z.files if k.startswith("x_"))


def test_rfftf_matches_reference_golden():
    z = np.load(GOLD)
    for L in _lengths(z):
        x = z["x_%d" % L]
        got = oracle.rfftf(x)
        ref = z["rfftf_%d" % L]
        assert np.abs(got - ref).max() <= 4e-15 * max(1.0, np.abs(ref).max()) * np.log2(max(L, 2)), L


def test_rfftb_matches_reference_golden():
    z = np.load(GOLD)
    for L in _lengths(z):
        x = z["x_%d" % L]
        got = oracle.rfftb(x)
        ref = z["rfftb_%d" % L]
        assert np.abs(got - ref).max() <= 8e-15 * max(1.0, np.abs(ref).max()) * np.log2(max(L, 2)), L


@pytest.mark.parametrize("L", [2, 6, 8, 10, 32, 96, 100, 256])
def test_rfftf_matches_numpy_packing(L):
    x = np.random.default_rng(L).standard_normal((4, L))
    np.testing.assert_allclose(oracle.rfftf(x), packed_rfft(x), rtol=0, atol=1e-13 * L)


@pytest.mark.parametrize("L", [6, 8, 12, 20])
def test_rfftf_bruteforce_longdouble(L):
    """The documented sums themselves (fftpack.f:170-190), evaluated in extended precision."""
    x = np.random.default_rng(100 + L).standard_normal(L)
    xl = x.astype(np.longdouble)
    r = np.zeros(L, dtype=np.longdouble)
    j = np.arange(L, dtype=np.longdouble)
    pi = np.longdouble(np.pi)
    r[0] = xl.sum()
    for k in range(1, (L - 1) // 2 + 1):
        r[2 * k - 1] = (xl * np.cos(2 * pi * k * j / L)).sum()
        r[2 * k] = -(xl * np.sin(2 * pi * k * j / L)).sum()
    if L % 2 == 0:
        r[L - 1] = (xl * (-1.0) ** np.arange(L)).sum()
    np.testing.assert_allclose(oracle.rfftf(x), r.astype(np.float64), rtol=0, atol=5e-15 * L)


@pytest.mark.parametrize("L", [2, 8, 30, 128, 384])
def test_roundtrip_is_L_times_identity(L):
    """RFFTB(RFFTF(x)) = L x (fftpack.f:196-199)."""
    x = np.random.default_rng(7).standard_normal((3, L))
    np.testing.assert_allclose(oracle.rfftb(oracle.rfftf(x)), L * x, rtol=0, atol=1e-13 * L)


def test_planted_sign_error_is_caught():
    """A plausible mistake (the e^{+i} sign convention) must fail the golden check."""
    z = np.load(GOLD)
    x = z["x_32"]
    wrong = oracle.rfftf(x)
    wrong[:, 2::2][:, :-1] *= -1.0  # flip the imaginary parts
    assert np.abs(wrong - z["rfftf_32"]).max() > 1e-3


def test_reference_library_direct_when_built():
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref/libfftpack_ref.so not built here")
    x = np.random.default_rng(3).standard_normal((5, 96))
    np.testing.assert_allclose(oracle.rfftf(x), oracle.ref_rfftf(x), rtol=0, atol=1e-13)
    np.testing.assert_allclose(oracle.rfftb(x), oracle.ref_rfftb(x), rtol=0, atol=1e-12)
