"""The resident loop's workload noise in the oracle: Philox4x32-10 against the published
known-answer vectors, the Box-Muller draws against N(0,1), which faces the noise touches,
and the CPU resident step's post-projection divergence."""
import numpy as np
import pytest
from scipy import stats

import oracle
from eddyproj import configs, grid

# Random123 known-answer vectors for philox4x32 with 10 rounds (kat_vectors of the
# Random123 distribution, Salmon et al., SC'11): (counter, key, output).
PHILOX_KAT = [
    ([0x00000000, 0x00000000, 0x00000000, 0x00000000], [0x00000000, 0x00000000],
     [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff], [0xffffffff, 0xffffffff],
     [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,key,out", PHILOX_KAT)
def test_philox_known_answers(ctr, key, out):
    assert oracle.philox4x32_10(ctr, key) == out


def test_normals_are_standard_normal_and_keyed():
    cells = np.arange(0, 40000 * 37, 37)
    g = np.array([oracle.normals(c, 3, 5) for c in cells])
    for col in range(3):
        ks = stats.kstest(g[:, col], "norm")
        assert ks.pvalue > 1e-3, (col, ks)
    # the three draws of a cell are uncorrelated, and so are two steps / two seeds
    c = np.corrcoef(g.T)
    assert np.abs(c - np.eye(3)).max() < 0.03
    g4 = np.array([oracle.normals(c, 4, 5) for c in cells[:5000]])
    g6 = np.array([oracle.normals(c, 3, 6) for c in cells[:5000]])
    assert abs(np.corrcoef(g[:5000, 0], g4[:, 0])[0, 1]) < 0.05
    assert abs(np.corrcoef(g[:5000, 0], g6[:, 0])[0, 1]) < 0.05
    np.testing.assert_array_equal(g[:10], np.array([oracle.normals(c, 3, 5) for c in cells[:10]]))


def test_perturb_touches_exactly_the_corrected_faces():
    g = grid.make_grid(12, 10, 8, x_beta=1.5)
    nx, ny, nz = g.shape()
    u, v, w = configs.random_uniform_velocity(g, 9)
    us, vs, ws = oracle.perturb(g, u, v, w, 1e-2, 5, 11)
    du, dv, dw = us - u, vs - v, ws - w
    # U faces I = 2..NX-2, W faces K = 2..NZ-2, V everywhere in the interior (1-based)
    mask_u = np.zeros(u.shape, bool)
    mask_u[1:nx - 2, :, 1:nz - 1] = True
    mask_v = np.zeros(u.shape, bool)
    mask_v[1:nx - 1, :, 1:nz - 1] = True
    mask_w = np.zeros(u.shape, bool)
    mask_w[1:nx - 1, :, 1:nz - 2] = True
    for d, m in ((du, mask_u), (dv, mask_v), (dw, mask_w)):
        assert np.all(d[~m] == 0.0)
        assert np.all(d[m] != 0.0)
        # periodic ghost rows carry their images' values
        np.testing.assert_array_equal(d[:, 0, :], d[:, ny - 2, :])
        np.testing.assert_array_equal(d[:, ny - 1, :], d[:, 1, :])
    # the increment at a cell is sigma times that cell's draws
    I, J, K = 3, 4, 5
    gg = oracle.normals(I + nx * (J + ny * K), 11, 5)
    np.testing.assert_allclose([du[I, J, K], dv[I, J, K], dw[I, J, K]], 1e-2 * gg, rtol=0, atol=1e-15)  # u + s g - u rounds
    assert 0.5e-2 < np.std(dv[mask_v]) < 2e-2


def test_resident_steps_leave_roundoff_divergence():
    g = grid.make_grid(24, 16, 12, x_beta=1.5)
    u, v, w = configs.wake(g)
    d0 = configs.divergence(g, *oracle.perturb(g, u, v, w, 1e-2, 5, 0))
    u1, v1, w1, log = oracle.resident_steps(g, u, v, w, 3, 1e-2, 5, configs.DT)
    assert log.shape == (3, 2)
    assert np.abs(log).max() <= 1e-10 * np.abs(d0).max()
    # periodic ghost rows stay consistent and the wall-normal wall faces are untouched
    np.testing.assert_array_equal(u1[:, 0, :], u1[:, -2, :])
    np.testing.assert_array_equal(u1[-2, :, :], u[-2, :, :])
    np.testing.assert_array_equal(w1[:, :, -2], w[:, :, -2])
