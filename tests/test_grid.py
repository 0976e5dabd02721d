"""Metric construction (src/grid.f) and the synthetic workloads."""
import os

import numpy as np

import oracle
from eddyproj import configs, grid


def test_stretched_metrics_closed_form_and_oracle():
    f = grid.faces_tanh_onesided(32, 4.0, 1.5)
    ap, au, xu, xc = grid.metrics_stretched(f)
    nx = 34
    assert ap.size == nx and au.size == nx
    np.testing.assert_allclose(ap[1:-1], 1.0 / np.diff(f), rtol=1e-14)             # grid.f:56
    np.testing.assert_allclose(au[1:-2], 1.0 / np.diff(xc[1:-1]), rtol=1e-13)      # centre spacing
    oap, oau = oracle.metrics_from_faces(nx, f)
    np.testing.assert_allclose(ap, oap, rtol=1e-15)
    np.testing.assert_allclose(au, oau, rtol=1e-15)
    assert f[0] == 0.0 and abs(f[-1] - 4.0) < 1e-15


def test_uniform_metrics_and_two_sided():
    ap, au, xu, _ = grid.metrics_uniform(16, 2.0)
    assert np.all(ap == 8.0) and np.all(au == 8.0)
    f = grid.faces_tanh_twosided(64, 2.0, 1.2)
    np.testing.assert_allclose(f, 2.0 - f[::-1], atol=1e-15)  # symmetric about the mid-plane


def test_config_shapes_and_compatibility():
    g, us, vs, ws = configs.random_64x64x32()
    assert (g.nx, g.ny, g.nz) == (66, 34, 66)
    assert (g.M, g.N, g.L) == (64, 64, 32)
    assert g.z_uniform()
    dp = configs.rhs_from_velocity(g, us, vs, ws)
    vol = g.cell_volume()
    b = dp[1:-1, 1:-1, 1:-1]
    assert abs((vol * b).sum()) <= 1e-12 * (vol * np.abs(b)).sum()
    np.testing.assert_array_equal(vs[:, 0], vs[:, -2])


def test_batch_sizes_are_seeded():
    a = configs.batch48_sizes()
    assert len(a) == 48 and a == configs.batch48_sizes()
    assert all(m in configs.BATCH_NX1 and n in configs.BATCH_NX2 and l in configs.BATCH_NX3 for m, n, l, _ in a)


GRIDS = configs.DATA


def test_production_grid_files():
    """The reference's production grids (run/x1_grid_half.in, run/x3_grid.in; data files kept
    in eddyproj/data): counts of run/eddy.input:327 (nx=302, nz=1026), the radial grid starting on
    the axis, and metrics equal to the oracle's restatement of grid.f on both axes."""
    xf = grid.read_grid_file(os.path.join(GRIDS, "x1_grid_half.in"))
    zf = grid.read_grid_file(os.path.join(GRIDS, "x3_grid.in"))
    assert xf.size == 301 and zf.size == 1025
    assert xf[0] == 0.0 and xf[-1] == 60.0 and zf[0] == -20.0 and zf[-1] == 20.0
    g = grid.from_faces(xf, zf, 128, 2.0 * np.pi, cyl=True)
    assert (g.nx, g.ny, g.nz) == (302, 130, 1026)
    assert g.ru[0] == 0.0 and np.all(g.rp[1:-1] > 0) and not g.z_uniform()
    for n, f, a, b in ((g.nx, xf, g.ap, g.au), (g.nz, zf, g.cp, g.cw)):
        oa, ob = oracle.metrics_from_faces(n, f)
        np.testing.assert_allclose(a, oa, rtol=1e-15)
        np.testing.assert_allclose(b, ob, rtol=1e-15)
