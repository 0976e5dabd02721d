"""The oracle's DIRECT against outputs of the reference's own DIRECT (tests/golden/direct_ref.npz,
written by tests/golden/make_direct_ref_golden.py from src/direct.f + pdc2dn.f + fftpack.f +
numrec.f compiled from the reference's sources and run under MPI with 1-4 ranks).

Both sides solve the same system; the comparison is after the plain-mean removal DIRECT applies
(direct.f:894-898). Tolerances per case: the uniform-x2 cases pin the whole path (FFT packing,
wavenumbers AK, per-mode solve, mean removal, ghost rules) to 1e-13; on stretched x2 the
reference's IPRES=4 solver (pdc2dn) itself leaves a relative residual of ~2e-8 (stored with
the case), so the bar there is 50x the reference's own residual."""
import numpy as np
import pytest

import oracle
from helpers import golden_case, golden_cases, golden_rhs


def mean_free(a):
    x = a[1:-1, 1:-1, 1:-1]
    return x - x.mean()


def rel_residual(g, dp, rhs):
    b = rhs[1:-1, 1:-1, 1:-1]
    return np.abs(oracle.laplacian(g, dp)[1:-1, 1:-1, 1:-1] - b).max() / np.abs(b).max()


def uniform_x2(g):
    return bool(np.all(g.cp[1:-1] == g.cp[1]) and np.all(g.cw == g.cw[1]))


def tolerance(g, ref_dp, rhs):
    """1e-13 where x2 is uniform; 50x the reference's own residual where it is stretched."""
    if uniform_x2(g):
        return 1e-13
    return max(1e-13, 50.0 * rel_residual(g, ref_dp, rhs))


@pytest.mark.parametrize("name", golden_cases())
def test_oracle_matches_reference_direct(name):
    g, ref_dp = golden_case(name)
    rhs = golden_rhs(name)
    got = oracle.direct(g, rhs, extended=True)
    a, b = mean_free(got), mean_free(ref_dp)
    err = np.abs(a - b).max() / np.abs(b).max()
    tol = tolerance(g, ref_dp, rhs)
    assert err <= tol, (name, err, tol)
    # the oracle solves the system to long-double accuracy
    assert rel_residual(g, got, rhs) <= 1e-13
    # ghost rules of direct.f:1090-1106 hold on the reference's output as on the oracle's
    if g.mp == 1:
        np.testing.assert_array_equal(ref_dp[0], ref_dp[1])
        np.testing.assert_array_equal(got[0], got[1])
    np.testing.assert_array_equal(ref_dp[:, 0], ref_dp[:, -2])
    np.testing.assert_array_equal(got[:, 0], got[:, -2])
    assert not ref_dp[:, :, 0].any() and not got[:, :, 0].any()


def test_reference_residuals():
    """What the loose bar rests on: the reference's residual is FP64 roundoff on every
    uniform-x2 case and 1e-12..1e-7 (pdc2dn, IPRES=4) on the stretched ones."""
    for name in golden_cases():
        g, ref_dp = golden_case(name)
        res = rel_residual(g, ref_dp, golden_rhs(name))
        if uniform_x2(g):
            assert res <= 2e-14, (name, res)
        else:
            assert res <= 1e-7, (name, res)


def test_random_64x64x32_samples():
    """The 64x64x32 workload (U(-1,1) velocities, seed 1, rebuilt here from its seed): the
    reference's answer at 4096 sampled cells and its norms, solved on 4 ranks."""
    d = np.load(golden_case.path)
    from eddyproj import configs

    g, us, vs, ws = configs.random_64x64x32()
    rhs = configs.rhs_from_velocity(g, us, vs, ws)
    oracle.set_threads(4)
    got = mean_free(oracle.direct(g, rhs))
    idx = d["random_64x64x32/sample_index"]
    ref = d["random_64x64x32/sample_value"]
    scale = d["random_64x64x32/norms"][1]
    assert np.abs(got.reshape(-1)[idx] - ref).max() <= 1e-13 * scale
    np.testing.assert_allclose([np.sqrt((got * got).sum()), np.abs(got).max()], d["random_64x64x32/norms"],
                               rtol=1e-13)


def test_planted_mistake_fails_fixture():
    """A sign error in the y-mode coupling (AK of the wrong sign) planted in the oracle's
    answer, through a perturbed solve of the uniform case, breaks the comparison."""
    name = "walls_16x8x8"
    g, ref_dp = golden_case(name)
    rhs = golden_rhs(name)
    g.delysq = -g.delysq  # AK -> -AK
    try:
        bad = oracle.direct(g, rhs, extended=True)
    except RuntimeError:
        return
    err = np.abs(mean_free(bad) - mean_free(ref_dp)).max() / np.abs(mean_free(ref_dp)).max()
    assert err > 1e-3
