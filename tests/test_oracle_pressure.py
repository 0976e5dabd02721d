"""The oracle's pressure update (Eddy6.f:2772-2846) pinned by what the statements fix: the
reference value zeroes the mean of P(2, 2:NY-1, 2), the boundary copies and their order, and
the implicit corrections on fields whose differences are known in closed form."""
import numpy as np
import pytest

import oracle
from eddyproj import grid
from helpers import interior


def fields(g, seed):
    rng = np.random.default_rng(seed)
    p = np.asfortranarray(rng.standard_normal(g.shape()))
    dp = np.asfortranarray(rng.standard_normal(g.shape()))
    return p, dp


def test_reference_value_and_wall_ghosts():
    g = grid.make_grid(10, 6, 8, x_beta=1.5)
    p, dp = fields(g, 1)
    q = oracle.pressure_update(g, p, dp, 1e-3)
    # P(2, 2:NY-1, 2) has zero mean; every interior cell moved by DP minus the same constant
    assert abs(q[1, 1:-1, 1].mean()) <= 1e-15
    shift = interior(q) - interior(p + dp)
    assert np.ptp(shift) <= 1e-14
    np.testing.assert_allclose(shift.mean(), -(p + dp)[1, 1:-1, 1].mean(), rtol=1e-13)
    # x walls (MP=1): P(1) = -P(IX1), P(NX) = -P(IX2) on every interior-row (y and z ghosts
    # are overwritten afterwards)
    np.testing.assert_array_equal(q[0, 1:-1, 1:-1], -q[1, 1:-1, 1:-1])
    np.testing.assert_array_equal(q[-1, 1:-1, 1:-1], -q[-2, 1:-1, 1:-1])
    # y periodic (LP=0) over all x, including the x ghosts: applied after x
    np.testing.assert_array_equal(q[:, 0, 1:-1], q[:, -2, 1:-1])
    np.testing.assert_array_equal(q[:, -1, 1:-1], q[:, 1, 1:-1])
    # x2 Neumann (ITYPE 5/6 wall types) over everything: applied last
    np.testing.assert_array_equal(q[:, :, 0], q[:, :, 1])
    np.testing.assert_array_equal(q[:, :, -1], q[:, :, -2])
    # corner: y copy of the x copy
    assert q[0, 0, 3] == -q[1, -2, 3]


def test_periodic_x2_and_axis_reflection():
    g = grid.make_grid(10, 6, 8, x_beta=1.5, cyl=True, lx3=2 * np.pi)
    p, dp = fields(g, 2)
    q = oracle.pressure_update(g, p, dp, 1e-3, itype1=300, itype5=500, itype6=500)
    np.testing.assert_array_equal(q[:, :, 0], q[:, :, -2])
    np.testing.assert_array_equal(q[:, :, -1], q[:, :, 1])
    L = g.L
    for J in range(1, g.ny - 1):  # 0-based interior row; JSYM is half a period away
        js = J + L // 2
        if js > L:
            js -= L
        np.testing.assert_array_equal(q[0, J, 1:-1], q[1, js, 1:-1])


@pytest.mark.parametrize("cyl", [False, True])
def test_implicit_corrections_closed_form(cyl):
    g = grid.make_grid(12, 6, 16, x_beta=1.5, cyl=cyl, lx3=2 * np.pi if cyl else 2.0)
    dt, ru1 = 2e-3, 0.01
    rng = np.random.default_rng(3)
    mask = (rng.random(g.nx) > 0.4).astype(float)
    p = np.zeros(g.shape(), order="F")
    # DP = cos(2 pi j / L) along y: its periodic second difference is (2 cos(2 pi/L) - 2) DP
    j = np.arange(g.ny) - 1
    dp = np.asfortranarray(np.broadcast_to(np.cos(2 * np.pi * j / g.L)[None, :, None], g.shape()).copy())
    q = oracle.pressure_update(g, p, dp, dt, ru1=ru1, rvimply=mask, imply=1)
    lam = 2.0 * np.cos(2 * np.pi / g.L) - 2.0
    rp = g.rp[1:-1, None, None]
    expect = interior(dp) - 0.5 * dt * ru1 / (g.dely * rp) ** 2 * lam * interior(dp) * mask[1:-1, None, None]
    pref = expect[0, :, 0].mean()
    np.testing.assert_allclose(interior(q), expect - pref, rtol=0, atol=1e-14)
    # DP constant: both implicit terms vanish (differences of a constant), P = DP - PREF = 0
    dpc = np.full(g.shape(), 0.7, order="F")
    q = oracle.pressure_update(g, p, dpc, dt, ru1=ru1, rvimply=mask, rvimplx=mask, imply=1, implx=1)
    assert np.abs(interior(q)).max() <= 1e-15


def test_implicit_x_on_linear_and_quadratic_profiles():
    """Uniform x1 grid, RU = RP = 1, AP = AU = 1/dx: the x correction is
    0.5 dt RU1 (DP(I+1) - 2 DP(I) + DP(I-1)) / dx^2, zero for a linear DP and
    0.5 dt RU1 * 2 c / dx^2 for DP = c x^2 (index units)."""
    g = grid.make_grid(12, 6, 8, x_stretch=False)
    dt, ru1, c = 1e-3, 0.05, 0.3
    mask = np.ones(g.nx)
    x = np.arange(g.nx, dtype=float)
    p = np.zeros(g.shape(), order="F")
    for prof, second in ((x, 0.0), (c * x * x, 2.0 * c)):
        dp = np.asfortranarray(np.broadcast_to(prof[:, None, None], g.shape()).copy())
        q = oracle.pressure_update(g, p, dp, dt, ru1=ru1, rvimplx=mask, implx=1)
        expect = interior(dp) - 0.5 * dt * ru1 * g.ap[1] * g.au[1] * second
        expect = expect - expect[0, :, 0].mean()
        np.testing.assert_allclose(interior(q), expect, rtol=0, atol=1e-12)
