"""The oracle's DIRECT, DIVERGENCE, CORRECT and operator pinned by mathematics: a dense 3-D
solve of the discrete system on tiny grids, an independent numpy restatement of the 7-point
operator, closed-form eigenpairs, the projection identity DIV(CORRECT(u)) = DIV(u) - dt L(phi)
and the solution properties direct.f fixes (plain-mean zero, ghost fill)."""
import numpy as np
import pytest

import oracle
from eddyproj import configs, grid
from helpers import interior


def numpy_operator(g, phi):
    """L(phi) on the interior from the metric arrays (direct.f:344-352, 370-378, 536); with
    MP=0 / NP=0 the x1 / x2 differences wrap around (AM(1), CM(M) / AN(1), CN(N) kept)."""
    M, L, N = g.M, g.L, g.N
    ap, au, cp, cw, rp, ru = g.ap, g.au, g.cp, g.cw, g.rp, g.ru
    p = phi[1:-1, 1:-1, 1:-1]
    out = np.zeros_like(p)
    for i in range(M):
        I = i + 1
        if i > 0 or g.mp == 0:
            out[i] += ap[I] / rp[I] * au[I - 1] * ru[I - 1] * (p[i - 1] - p[i])
        if i < M - 1 or g.mp == 0:
            out[i] += ap[I] / rp[I] * au[I] * ru[I] * (p[(i + 1) % M] - p[i])
        out[i] += (np.roll(p[i], -1, axis=0) - 2 * p[i] + np.roll(p[i], 1, axis=0)) / g.delysq / rp[I] ** 2
    for k in range(N):
        K = k + 1
        if k > 0 or g.np == 0:
            out[:, :, k] += cp[K] * cw[K - 1] * (p[:, :, k - 1] - p[:, :, k])
        if k < N - 1 or g.np == 0:
            out[:, :, k] += cp[K] * cw[K] * (p[:, :, (k + 1) % N] - p[:, :, k])
    return out


def dense_solution(g, dp):
    """The discrete system solved densely (plain-mean constraint appended): tiny grids only."""
    n = g.M * g.L * g.N
    A = np.zeros((n, n))
    for t in range(n):
        e = np.zeros(g.shape(), order="F")
        interior(e).flat[t] = 1.0
        A[:, t] = numpy_operator(g, e).ravel()
    b = interior(dp).ravel()
    A2 = np.vstack([A, np.ones((1, n))])
    return np.linalg.lstsq(A2, np.concatenate([b, [0.0]]), rcond=None)[0]


def periodic_rhs(g, seed):
    rng = np.random.default_rng(seed)
    b = rng.standard_normal((g.M, g.L, g.N))
    vol = g.cell_volume()
    b -= (vol * b).sum() / vol.sum()
    dp = np.zeros(g.shape(), order="F")
    interior(dp)[:] = b
    return dp


def tiny(m=6, n=5, l=4, zb=1.2, seed=7, cyl=False):
    g = grid.make_grid(m, n, l, x_beta=1.7, z_beta=zb, cyl=cyl)
    us, vs, ws = configs.random_uniform_velocity(g, seed)
    return g, configs.rhs_from_velocity(g, us, vs, ws)


@pytest.mark.parametrize("zb,cyl", [(None, False), (1.2, False), (1.2, True)])
def test_operator_matches_numpy_restatement(zb, cyl):
    g, _ = tiny(zb=zb, cyl=cyl)
    phi = np.asfortranarray(np.random.default_rng(1).standard_normal(g.shape()))
    np.testing.assert_allclose(interior(oracle.laplacian(g, phi)), numpy_operator(g, phi), rtol=1e-13, atol=1e-10)


@pytest.mark.parametrize("extended", [False, True])
@pytest.mark.parametrize("zb,l,cyl", [(None, 4, False), (1.2, 4, False), (1.2, 6, False), (None, 2, False),
                                      (None, 4, True), (1.2, 6, True)])
def test_direct_matches_dense_solve(zb, l, cyl, extended):
    """cyl: x1 is the radius from the axis, RU = XU (RU = 0 on the axis face), RP = XC."""
    g, dp = tiny(zb=zb, l=l, cyl=cyl)
    phi = oracle.direct(g, dp, extended=extended)
    n = g.M * g.L * g.N
    A = np.zeros((n, n))
    for t in range(n):
        e = np.zeros(g.shape(), order="F")
        interior(e).flat[t] = 1.0
        A[:, t] = numpy_operator(g, e).ravel()
    b = interior(dp).ravel()
    # unique solution with zero plain mean (direct.f:895-898)
    A2 = np.vstack([A, np.ones((1, n))])
    sol = np.linalg.lstsq(A2, np.concatenate([b, [0.0]]), rcond=None)[0]
    assert np.abs(interior(phi).ravel() - sol).max() <= 1e-12 * np.abs(sol).max()


def test_direct_ghosts_and_mean():
    g, dp = tiny()
    phi = oracle.direct(g, dp)
    assert abs(interior(phi).mean()) < 1e-13 * np.abs(phi).max()
    np.testing.assert_array_equal(phi[0], phi[1])            # MP=1: DP(1)=DP(2)
    np.testing.assert_array_equal(phi[-1], phi[-2])          # DP(NX)=DP(NX-1)
    np.testing.assert_array_equal(phi[:, 0], phi[:, -2])     # LP=0: DP(:,1)=DP(:,NY-1)
    np.testing.assert_array_equal(phi[:, -1], phi[:, 1])     # DP(:,NY)=DP(:,2)
    assert not phi[:, :, 0].any() and not phi[:, :, -1].any()  # z ghosts zero


def test_direct_random_64x64x32_residual():
    g, us, vs, ws = configs.random_64x64x32()
    dp = configs.rhs_from_velocity(g, us, vs, ws)
    oracle.set_threads(4)
    phi = oracle.direct(g, dp)
    r = interior(oracle.laplacian(g, phi)) - interior(dp)
    assert np.linalg.norm(r) <= 1e-13 * np.linalg.norm(interior(dp))


def test_planted_operator_mistake_fails_dense_check():
    """Planting a dropped term (the y coupling) in the oracle's answer breaks the residual."""
    g, dp = tiny()
    phi = oracle.direct(g, dp)
    bad = phi.copy(order="F")
    interior(bad)[:] = interior(bad) + 1e-3 * np.sin(np.arange(interior(bad).size)).reshape(interior(bad).shape)
    r = numpy_operator(g, bad) - interior(dp)
    assert np.abs(r).max() > 1e-6 * np.abs(interior(dp)).max()


def test_tridiag_eig_against_numpy_and_closed_form():
    n = 12
    d = np.full(n, -2.0)
    d[0] = d[-1] = -1.0
    e = np.ones(n - 1)
    lam, z = oracle.tridiag_eig(d, e)
    exact = -4.0 * np.sin(np.pi * np.arange(n) / (2 * n)) ** 2
    np.testing.assert_allclose(np.sort(lam), np.sort(exact), atol=1e-14)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    np.testing.assert_allclose(T @ z, z * lam, atol=1e-13)
    np.testing.assert_allclose(z.T @ z, np.eye(n), atol=1e-14)
    rng = np.random.default_rng(2)
    d2, e2 = rng.standard_normal(9), rng.standard_normal(8)
    lam2, _ = oracle.tridiag_eig(d2, e2)
    np.testing.assert_allclose(lam2, np.linalg.eigvalsh(np.diag(d2) + np.diag(e2, 1) + np.diag(e2, -1)), atol=1e-13)


def test_mode_solve_matches_dense_2d():
    g, _ = tiny(m=7, n=6, l=8)
    M, N = g.M, g.N
    for J in (1, 2, 3, 8):
        rhs = np.random.default_rng(J).standard_normal((M, N))
        if J == 1:  # compatible RHS for the singular mode: zero volume-weighted sum
            w = (1.0 / g.ap[1:-1])[:, None] * (1.0 / g.cp[1:-1])[None, :]
            rhs -= (w * rhs).sum() / w.sum()
        got = oracle.mode_solve(g, J, rhs)
        kw = J // 2
        ak = 2.0 * (1.0 - np.cos(2 * np.pi * kw / g.L)) / g.delysq
        A = np.zeros((M * N, M * N))
        idx = lambda i, k: k + N * i  # noqa: E731  (RHS1D(k + N*(i-1)), direct.f:755-762)
        for i in range(M):
            I = i + 1
            am = g.ap[I] * g.au[I - 1] if i > 0 else 0.0
            cm = g.ap[I] * g.au[I] if i < M - 1 else 0.0
            for k in range(N):
                K = k + 1
                an = g.cp[K] * g.cw[K - 1] if k > 0 else 0.0
                cn = g.cp[K] * g.cw[K] if k < N - 1 else 0.0
                r = idx(i, k)
                A[r, r] = -am - cm - an - cn - ak
                if i > 0:
                    A[r, idx(i - 1, k)] = am
                if i < M - 1:
                    A[r, idx(i + 1, k)] = cm
                if k > 0:
                    A[r, idx(i, k - 1)] = an
                if k < N - 1:
                    A[r, idx(i, k + 1)] = cn
        b = rhs.ravel()
        if J == 1:
            sol = np.linalg.lstsq(np.vstack([A, np.ones((1, M * N))]), np.concatenate([b, [0.0]]), rcond=None)[0]
        else:
            sol = np.linalg.solve(A, b)
        assert np.abs(got.ravel() - sol).max() <= 1e-12 * np.abs(sol).max(), J


def test_projection_identity_and_divergence_stencil():
    g, us, vs, ws = configs.random_64x64x32()
    dt = configs.DT
    div = oracle.divergence(g, us, vs, ws)
    np.testing.assert_allclose(interior(div), configs.divergence(g, us, vs, ws), rtol=1e-14, atol=1e-12)
    dp = configs.rhs_from_velocity(g, us, vs, ws)
    phi = oracle.direct(g, dp)
    uo, vo, wo = oracle.correct(g, us, vs, ws, phi, dt)
    for a in (uo, vo, wo):  # periodic y ghosts, as the velocity BCs after CORRECT set them
        configs._periodic_y(a)
    div2 = interior(oracle.divergence(g, uo, vo, wo))
    # DIV(CORRECT(u)) = DIV(u) - dt L(phi) = dt * (volume-weighted mean of DIV(u)/dt)
    vol = g.cell_volume()
    c = (vol * interior(div)).sum() / vol.sum()
    assert np.abs(div2 - c).max() <= 1e-9 * np.abs(interior(div)).max()


def test_projection_identity_cylindrical():
    """Cylindrical metrics (grid.f:462-466): the /RP and RU factors of DIVERGENCE and CORRECT
    and the AK/RP^2 mode term of DIRECT; the identity DIV(CORRECT(u)) = dt * mean holds."""
    g = grid.make_grid(24, 12, 16, x_beta=1.5, cyl=True, lx3=2.0 * np.pi)
    us, vs, ws = configs.random_uniform_velocity(g, 4)
    dt = configs.DT
    div = oracle.divergence(g, us, vs, ws)
    np.testing.assert_allclose(interior(div), configs.divergence(g, us, vs, ws), rtol=1e-14, atol=1e-12)
    dp = configs.rhs_from_velocity(g, us, vs, ws)
    phi = oracle.direct(g, dp, extended=True)
    uo, vo, wo = oracle.correct(g, us, vs, ws, phi, dt)
    for a in (uo, vo, wo):
        configs._periodic_y(a)
    div2 = interior(oracle.divergence(g, uo, vo, wo))
    vol = g.cell_volume()
    c = (vol * interior(div)).sum() / vol.sum()
    assert np.abs(div2 - c).max() <= 1e-9 * np.abs(interior(div)).max()


def test_correct_faces_follow_setup_bounds():
    g, us, vs, ws = configs.random_64x64x32()
    phi = np.asfortranarray(np.random.default_rng(5).standard_normal(g.shape()))
    uo, vo, wo = oracle.correct(g, us, vs, ws, phi, 1e-3)
    np.testing.assert_array_equal(uo[0], us[0])      # wall face I=1 kept (IBU=2)
    assert not np.array_equal(uo[1:-2, 1:-1, 1:-1], us[1:-2, 1:-1, 1:-1])
    np.testing.assert_array_equal(uo[-2], us[-2])
    np.testing.assert_array_equal(wo[:, :, -2], ws[:, :, -2])  # K=NZ-1 wall kept
    np.testing.assert_array_equal(wo[:, :, 0], ws[:, :, 0])
    i, j, k = 5, 7, 9
    exp_u = us[i, j, k] - g.au[i] * (phi[i + 1, j, k] - phi[i, j, k]) * 1e-3
    assert uo[i, j, k] == pytest.approx(exp_u, rel=1e-15)


def test_fp64_and_extended_oracles_agree_on_stretched_grid():
    """The FP64 build and the long-double build solve the same system; their difference is
    the FP64 build's roundoff (amplified by the conditioning of the stretched operator)."""
    g = grid.make_grid(64, 48, 32, x_beta=2.0, z_beta=1.2)
    us, vs, ws = configs.random_uniform_velocity(g, 9)
    dp = configs.rhs_from_velocity(g, us, vs, ws)
    a = oracle.direct(g, dp)
    b = oracle.direct(g, dp, extended=True)
    ia, ib = interior(a), interior(b)
    assert np.abs(ia - ib).max() <= 2e-12 * np.abs(ib).max()
    r = interior(oracle.laplacian(g, b)) - interior(dp)
    assert np.linalg.norm(r) <= 1e-13 * np.linalg.norm(interior(dp))


# ------------------------------------------------------------------ periodic x1 / x2 (MP=0, NP=0)
def periodic_grid(mp, np_, m=6, n=5, l=4, zb=None, xb=1.7, cw_wrap=True):
    """Tiny grid with periodic x1 (mp=0) and/or x2 (np_=0). A periodic axis is uniform unless
    its stretching is asked for; cw_wrap makes CW(1) = CW(N+1) (a symmetric periodic x2)."""
    g = grid.make_grid(m, n, l, x_beta=xb, z_beta=zb, x_stretch=(mp == 1 or xb is not None))
    g.mp, g.np = mp, np_
    if np_ == 0 and zb is not None and cw_wrap:
        g.cw[0] = g.cw[-2]
    return g


@pytest.mark.parametrize("extended", [False, True])
@pytest.mark.parametrize("mp,np_,zb,xb,cw_wrap", [
    (0, 1, None, None, True),     # periodic uniform x1, walls x2
    (0, 1, 1.2, None, True),      # periodic x1, stretched walled x2
    (1, 0, None, 1.7, True),      # walls x1 (stretched), periodic uniform x2 (IPRES=3's case)
    (1, 0, 1.2, 1.7, True),       # periodic stretched x2, symmetric (CW(1) = CW(N+1))
    (1, 0, 1.2, 1.7, False),      # periodic stretched x2 with CW(1) != CW(N+1): x1 eigenbasis path
    (0, 0, None, None, True),     # doubly periodic
])
def test_periodic_direct_matches_dense_solve(mp, np_, zb, xb, cw_wrap, extended):
    g = periodic_grid(mp, np_, zb=zb, xb=xb, cw_wrap=cw_wrap)
    dp = periodic_rhs(g, 3)
    if np_ == 0 and not cw_wrap:
        # the operator is not symmetric under the volume weights: make b compatible with the
        # left null vector of the dense operator instead
        n = g.M * g.L * g.N
        A = np.zeros((n, n))
        for t in range(n):
            e = np.zeros(g.shape(), order="F")
            interior(e).flat[t] = 1.0
            A[:, t] = numpy_operator(g, e).ravel()
        u, sv, vt = np.linalg.svd(A)
        y = u[:, -1]
        b = interior(dp).ravel()
        interior(dp).flat[:] = b - (y @ b) * y
    phi = oracle.direct(g, dp, extended=extended)
    sol = dense_solution(g, dp)
    assert np.abs(interior(phi).ravel() - sol).max() <= 1e-12 * np.abs(sol).max()


@pytest.mark.parametrize("mp,np_", [(0, 1), (1, 0), (0, 0)])
def test_periodic_ghosts(mp, np_):
    g = periodic_grid(mp, np_, m=8, n=6, l=5)
    phi = oracle.direct(g, periodic_rhs(g, 4))
    if mp == 0:  # direct.f:1090-1093: DP(1)=DP(NX-1), DP(NX)=DP(2)
        np.testing.assert_array_equal(phi[0], phi[-2])
        np.testing.assert_array_equal(phi[-1], phi[1])
    else:
        np.testing.assert_array_equal(phi[0], phi[1])
    np.testing.assert_array_equal(phi[:, 0], phi[:, -2])
    assert not phi[:, :, 0].any() and not phi[:, :, -1].any()
    assert abs(interior(phi).mean()) < 1e-13 * np.abs(phi).max()


def test_lp1_ghost_rule():
    """LP=1 changes only DIRECT's y ghost fill (direct.f:1100-1106): copies of J=2 and NY-1."""
    g, dp = tiny(l=6)
    a = oracle.direct(g, dp)
    g.lp = 1
    b = oracle.direct(g, dp)
    np.testing.assert_array_equal(interior(a), interior(b))
    np.testing.assert_array_equal(b[:, 0], b[:, 1])
    np.testing.assert_array_equal(b[:, -1], b[:, -2])


def test_both_paths_agree_on_symmetric_periodic_x2():
    """The x2-eigenbasis path and the x1-eigenbasis path solve the same system."""
    g = periodic_grid(1, 0, m=10, n=8, l=6, zb=1.3)
    dp = periodic_rhs(g, 5)
    a = oracle.direct(g, dp, extended=True, path=1)
    b = oracle.direct(g, dp, extended=True, path=2)
    assert np.abs(interior(a) - interior(b)).max() <= 1e-15 * np.abs(interior(a)).max() * 100


def test_periodic_x2_uniform_eigenvalues_closed_form():
    """Uniform periodic x2: the oracle's x2 operator spectrum is -4/dz^2 sin^2(pi m/N), pinned
    through single-mode fields: the answer to b = cos(2 pi q k/N) is b / lambda_q (x1, y
    constant: the zero y-mode, AK = 0; x1 walls: the constant x1 mode)."""
    g = periodic_grid(1, 0, m=4, n=12, l=3)
    N = g.N
    dz = 1.0 / g.cp[1]
    for q in (1, 2, 5, 6):
        kk = np.arange(N)
        b = np.cos(2 * np.pi * q * kk / N)
        dp = np.zeros(g.shape(), order="F")
        interior(dp)[:] = b[None, None, :]
        phi = oracle.direct(g, dp, extended=True)
        lam = -4.0 / dz ** 2 * np.sin(np.pi * q / N) ** 2
        np.testing.assert_allclose(interior(phi), np.broadcast_to(b / lam, interior(phi).shape), atol=1e-13)
