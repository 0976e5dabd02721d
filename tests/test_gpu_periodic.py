"""The boundary flags DIRECT takes besides the walls (direct.f:344-378, 1090-1106): periodic x1
(MP=0, cyclic x1 lines), periodic x2 (NP=0, an x2 eigenbasis of the cyclic operator, or the x1
eigenbasis with cyclic x2 lines when CW(1) != CW(N+1)), LP=1 (DIRECT's other y ghost rule) and
IPRES=3 (FFT along a uniform periodic x2 of spacing sqrt(DELZSQ)). HIP path through the C-ABI
against the long-double oracle (1e-12 after the plain-mean removal, ghosts exact) and against
outputs of the reference's own DIRECT (tests/golden/direct_ref.npz)."""
import numpy as np
import pytest

import eddyproj
import oracle
from eddyproj import configs, grid
from helpers import golden_case, golden_cases, golden_rhs, interior, rel_maxdiff_nomean

pytestmark = pytest.mark.gpu

TOL = 1e-12


def say(msg):
    print(msg, flush=True)


def periodic_grid(m, n, l, mp, np_, xb=None, zb=None, cw_wrap=True):
    g = grid.make_grid(m, n, l, x_beta=xb if xb is not None else 1.5, z_beta=zb, x_stretch=xb is not None)
    g.mp, g.np = mp, np_
    if np_ == 0 and zb is not None and cw_wrap:
        g.cw[0] = g.cw[-2]  # a periodic stretched x2: CW(1) = CW(N+1)
    return g


def random_rhs(g, seed):
    rng = np.random.default_rng(seed)
    b = rng.uniform(-1.0, 1.0, (g.M, g.L, g.N))
    vol = g.cell_volume()
    b -= (vol * b).sum() / vol.sum()
    dp = np.zeros(g.shape(), order="F")
    interior(dp)[:] = b
    return dp


def check_ghost_rules(g, phi):
    if g.mp == 0:  # DP(1)=DP(NX-1), DP(NX)=DP(2)
        np.testing.assert_array_equal(phi[0, 1:-1, 1:-1], phi[-2, 1:-1, 1:-1])
        np.testing.assert_array_equal(phi[-1, 1:-1, 1:-1], phi[1, 1:-1, 1:-1])
    else:
        np.testing.assert_array_equal(phi[0, 1:-1, 1:-1], phi[1, 1:-1, 1:-1])
        np.testing.assert_array_equal(phi[-1, 1:-1, 1:-1], phi[-2, 1:-1, 1:-1])
    if g.lp == 1:  # DP(:,1)=DP(:,2), DP(:,NY)=DP(:,NY-1)
        np.testing.assert_array_equal(phi[:, 0, 1:-1], phi[:, 1, 1:-1])
        np.testing.assert_array_equal(phi[:, -1, 1:-1], phi[:, -2, 1:-1])
    else:
        np.testing.assert_array_equal(phi[:, 0, 1:-1], phi[:, -2, 1:-1])
        np.testing.assert_array_equal(phi[:, -1, 1:-1], phi[:, 1, 1:-1])
    assert not phi[:, :, 0].any() and not phi[:, :, -1].any()


def gpu_direct(torch, g, dp, tuning=None):
    plan = eddyproj.Plan(g, tuning=tuning)
    d = eddyproj.to_device(dp)
    plan.direct(d)
    torch.cuda.synchronize()
    out = eddyproj.to_host(d)
    plan.close()
    return out


# (m, n, l, MP, NP, x1 beta, x2 beta, CW(1) = CW(N+1)): x1 line lengths that take 1..64 lanes
# per line and padded chunks; uniform and stretched periodic x2; the x1 form (M < N, or no
# symmetric x2 form); both periodic
CASES = [
    (24, 16, 8, 0, 1, None, None, True),      # periodic x1 + DCT x2
    (100, 20, 6, 0, 1, None, 1.2, True),      # periodic x1 + dense x2 (padded lines)
    (300, 12, 4, 0, 1, None, None, True),     # 64-lane periodic lines
    (32, 24, 8, 1, 0, 1.5, None, True),       # periodic uniform x2 (dense eigenbasis)
    (40, 18, 6, 1, 0, 1.5, 1.3, True),        # periodic stretched x2, symmetric
    (20, 36, 6, 1, 0, 1.5, 1.3, False),       # CW(1) != CW(N+1): x1 eigenbasis, cyclic x2 lines
    (48, 160, 4, 1, 0, 1.7, None, True),      # x1 form (M < N, N >= 128) with cyclic x2 lines
    (36, 20, 10, 0, 0, None, None, True),     # doubly periodic
    (130, 64, 4, 0, 0, None, 1.2, True),      # doubly periodic: x2-form products, padded cyclic x1 lines
    (64, 130, 4, 0, 0, None, 1.2, True),      # doubly periodic: x1 form (cyclic x1 basis, cyclic x2 lines)
]


@pytest.mark.parametrize("m,n,l,mp,np_,xb,zb,wrap", CASES)
def test_periodic_direct_matches_oracle(torch_cuda, m, n, l, mp, np_, xb, zb, wrap):
    say("case periodic direct %dx%dx%d MP=%d NP=%d xb=%s zb=%s wrap=%s" % (m, n, l, mp, np_, xb, zb, wrap))
    g = periodic_grid(m, n, l, mp, np_, xb, zb, wrap)
    dp = random_rhs(g, m + n)
    if not wrap:  # only the x1 form solves it; compare on the same pinned system
        ref = oracle.direct(g, dp, extended=True, path=2)
    else:
        ref = oracle.direct(g, dp, extended=True)
    got = gpu_direct(torch_cuda, g, dp)
    err = rel_maxdiff_nomean(got, ref)
    say("  err %.2e" % err)
    assert err <= TOL
    check_ghost_rules(g, got)


@pytest.mark.parametrize("mp,np_", [(0, 1), (1, 0), (0, 0)])
def test_periodic_fp32_refine(torch_cuda, mp, np_):
    say("case periodic fp32 refine MP=%d NP=%d" % (mp, np_))
    g = periodic_grid(64, 48, 16, mp, np_, None if mp == 0 else 1.5, None)
    dp = random_rhs(g, 7)
    plan = eddyproj.Plan(g)
    d = eddyproj.to_device(dp)
    res = plan.direct_refined(d, sweeps=3)
    torch_cuda.cuda.synchronize()
    got = eddyproj.to_host(d)
    plan.close()
    err = rel_maxdiff_nomean(got, oracle.direct(g, dp, extended=True))
    say("  residual %.2e err %.2e" % (res, err))
    assert res <= 1e-10
    assert err <= 1e-10
    check_ghost_rules(g, got)


def test_lp1_ghost_rule(torch_cuda):
    say("case LP=1")
    g = grid.make_grid(40, 16, 12, x_beta=1.5)
    g.lp = 1
    dp = random_rhs(g, 3)
    got = gpu_direct(torch_cuda, g, dp)
    ref = oracle.direct(g, dp, extended=True)
    assert rel_maxdiff_nomean(got, ref) <= TOL
    check_ghost_rules(g, got)


@pytest.mark.parametrize("mp,np_", [(0, 1), (1, 0), (0, 0)])
def test_periodic_projection(torch_cuda, mp, np_):
    """DIVERGENCE -> DIRECT -> REFRESHBC (periodic z ghosts when NP=0) -> CORRECT, with the
    faces of setup.f:17-72 for the periodic flags, against the oracle's same sequence."""
    say("case periodic projection MP=%d NP=%d" % (mp, np_))
    torch = torch_cuda
    g = periodic_grid(48, 32, 16, mp, np_, None if mp == 0 else 1.5, None)
    us, vs, ws = configs.random_uniform_velocity(g, 5)
    # periodic ghost faces, as the velocity BCs of a periodic direction leave them
    if mp == 0:
        us[0] = us[-2]
    if np_ == 0:
        ws[:, :, 0] = ws[:, :, -2]
    dt = configs.DT
    plan = eddyproj.Plan(g)
    dv = [eddyproj.to_device(a) for a in (us, vs, ws)]
    outs = [t.clone() for t in dv]
    dpd = torch.zeros_like(dv[0])
    plan.project(dv[0], dv[1], dv[2], outs[0], outs[1], outs[2], dpd, dt, remove_mean=True)
    torch.cuda.synchronize()
    phi = eddyproj.to_host(dpd)
    # oracle: RHS with the volume-weighted mean removed, DIRECT, REFRESHBC, CORRECT
    div = interior(oracle.divergence(g, us, vs, ws)) / dt
    vol = g.cell_volume()
    rhs = np.zeros(g.shape(), order="F")
    interior(rhs)[:] = div - (vol * div).sum() / vol.sum()
    ref = oracle.direct(g, rhs, extended=True)
    assert rel_maxdiff_nomean(phi, ref) <= TOL
    if np_ == 0:  # REFRESHBC at one rank: periodic images in the z ghost planes
        np.testing.assert_array_equal(phi[:, :, 0], phi[:, :, -2])
        np.testing.assert_array_equal(phi[:, :, -1], phi[:, :, 1])
    uo, vo, wo = oracle.correct(g, us, vs, ws, phi, dt)
    for got, want in zip(outs, (uo, vo, wo)):
        got = eddyproj.to_host(got)
        # corrected faces (interior y rows; the periodic y ghost rows are refreshed on the GPU)
        a, b = got[:, 1:-1, :], want[:, 1:-1, :]
        assert np.abs(a - b).max() <= 1e-13 * max(np.abs(b).max(), 1.0)
    plan.close()


@pytest.mark.parametrize("name", golden_cases())
def test_reference_direct_fixtures(torch_cuda, name):
    """The drop-in DIRECT (Fortran signature, host arrays) on the inputs the reference's own
    DIRECT solved (IPRES as run there), against its answer: 1e-13 where x2 is uniform, and 50x
    the reference's own residual where x2 is stretched (pdc2dn's accuracy)."""
    g, ref_dp = golden_case(name)
    say("case reference fixture %s (IPRES=%d)" % (name, g.ipres))
    rhs = golden_rhs(name)
    dp = np.array(rhs, order="F", copy=True)
    eddyproj.direct(dp, g, ipres=g.ipres)
    b = interior(rhs)
    res = np.abs(interior(oracle.laplacian(g, ref_dp)) - b).max() / np.abs(b).max()
    uniform = bool(np.all(g.cp[1:-1] == g.cp[1]) and np.all(g.cw == g.cw[1]))
    tol = 1e-13 if uniform else max(1e-13, 50.0 * res)
    err = rel_maxdiff_nomean(dp, ref_dp)
    say("  err %.2e tol %.2e" % (err, tol))
    assert err <= tol
    check_ghost_rules(g, dp)
    # the GPU answer itself is solved to FP64 accuracy, whatever the reference's residual
    gres = np.abs(interior(oracle.laplacian(g, dp)) - b).max() / np.abs(b).max()
    assert gres <= 1e-12
