"""Synthetic workloads of the projection path (host side, numpy).

Each builder returns (grid, us, vs, ws) with velocity arrays in the reference's Fortran
layout (NX, NY, NZ): U (x1) lives on x1-faces I=1..NX-1, V (x3, periodic) on y-faces, W (x2)
on z-faces K=1..NZ-1. Periodic ghost faces in y are made consistent (V(:,1,:)=V(:,NY-1,:)
and likewise the ghost rows of U and W), since DIVERGENCE reads V(I,1,K) at J=2
(timeindeg.f:945).

Axis naming: the workloads are quoted as (x1, x2, x3) = (walls, walls, periodic); the
reference's DIRECT transforms its 2nd index Y, so x3 -> Y (DP 2nd index), x2 -> Z (3rd).
"""
import os

import numpy as np

from . import grid
from .grid import make_grid

DT = 1.0e-3


def _periodic_y(a):
    a[:, 0, :] = a[:, -2, :]
    a[:, -1, :] = a[:, 1, :]
    return a


def _fortran(shape_xyz, fn):
    nx, ny, nz = shape_xyz
    return np.asfortranarray(fn(nx, ny, nz))


def random_uniform_velocity(g, seed, amp=1.0):
    """u, v, w i.i.d. uniform(-amp, amp) on every face slot (seeded numpy PCG64)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = g.shape()
    out = []
    for _ in range(3):
        a = rng.uniform(-amp, amp, size=(nz, ny, nx)).T.copy(order="F")
        out.append(_periodic_y(a))
    return tuple(out)


def random_64x64x32():
    """64x64x32 (x1,x2,x3); x1 tanh beta=1.5; x2, x3 uniform; U(-1,1) velocity, seed 1."""
    g = make_grid(64, 64, 32, x_beta=1.5)
    return (g,) + random_uniform_velocity(g, 1)


def lamb_oseen_dipole(g, seed, sep=0.3, rc=0.1, gamma=1.0, sigma=1e-2):
    """Two Lamb-Oseen vortices with axes along x3 at (x1, x2) = (Lx1/2 -+ sep/2, Lx2/2),
    circulation +gamma / -gamma, plus N(0, sigma^2) noise (seed)."""
    nx, ny, nz = g.shape()
    x1c = 0.5 * (g.xu[0] + g.xu[nx - 2])
    x2c = 0.5 * (g.zw[0] + g.zw[nz - 2])

    def vel(px, pz):
        u1 = np.zeros(np.broadcast(px, pz).shape)
        u2 = np.zeros_like(u1)
        for xc, gam in ((x1c - 0.5 * sep, gamma), (x1c + 0.5 * sep, -gamma)):
            dx, dz = px - xc, pz - x2c
            r2 = dx * dx + dz * dz
            r2 = np.where(r2 == 0.0, 1e-300, r2)
            f = gam / (2.0 * np.pi * r2) * (1.0 - np.exp(-r2 / (rc * rc)))
            u1 += -f * dz
            u2 += f * dx
        return u1, u2

    # U on x1-faces (xu(I), zc(K)); W on x2-faces (xc(I), zw(K))
    u1, _ = vel(g.xu[:, None], g.zc[None, :])
    _, u2 = vel(g.xc[:, None], g.zw[None, :])
    rng = np.random.default_rng(seed)
    us = np.asfortranarray(np.repeat(u1[:, None, :], ny, axis=1))
    ws = np.asfortranarray(np.repeat(u2[:, None, :], ny, axis=1))
    vs = np.zeros((nx, ny, nz), order="F")
    for a in (us, vs, ws):
        a += rng.normal(0.0, sigma, size=(nz, ny, nx)).T
        _periodic_y(a)
    return us, vs, ws


def dipole_256x192x128():
    """256x192x128; x1 tanh beta=2.0; x2 two-sided tanh beta=1.2; Lamb-Oseen dipole, seed 2."""
    g = make_grid(256, 192, 128, x_beta=2.0, z_beta=1.2)
    return (g,) + lamb_oseen_dipole(g, 2)


def taylor_green(g, seed, k=2, noise=1e-3, lx=(4.0, 2.0, 2.0)):
    """u1 = sin(a x1) cos(b x2) cos(c x3), u2 = -cos(a x1) sin(b x2) cos(c x3), u3 = 0 with
    a = 2 pi k/Lx1, b = 2 pi k/Lx2, c = 2 pi k/Lx3, plus U(-noise, noise) (seed)."""
    nx, ny, nz = g.shape()
    a, b, c = (2.0 * np.pi * k / L for L in lx)
    yc = g.yc
    us = (np.sin(a * g.xu)[:, None, None] * np.cos(b * g.zc)[None, None, :] * np.cos(c * yc)[None, :, None])
    ws = (-np.cos(a * g.xc)[:, None, None] * np.sin(b * g.zw)[None, None, :] * np.cos(c * yc)[None, :, None])
    us = np.asfortranarray(us)
    ws = np.asfortranarray(ws)
    vs = np.zeros((nx, ny, nz), order="F")
    rng = np.random.default_rng(seed)
    for arr in (us, vs, ws):
        arr += rng.uniform(-noise, noise, size=(nz, ny, nx)).T
        _periodic_y(arr)
    return us, vs, ws


def taylor_green_512x256x256():
    """512x256x256; x1 tanh beta=1.5; x2, x3 uniform; Taylor-Green k=2 + U(-1e-3,1e-3), seed 3."""
    g = make_grid(512, 256, 256, x_beta=1.5)
    return (g,) + taylor_green(g, 3)


def wake(g, amp=0.3, width=0.2):
    """Streamwise (x3, the periodic axis) velocity 1 - amp*exp(-r^2/(2 width^2)), r measured
    from the (x1, x2) centre of the domain; u = w = 0. Divergence-free and compatible with the
    walls; the resident loop perturbs it with seeded noise every step."""
    nx, ny, nz = g.shape()
    x1c = 0.5 * (g.xu[0] + g.xu[nx - 2])
    x2c = 0.5 * (g.zw[0] + g.zw[nz - 2])
    r2 = (g.xc[:, None] - x1c) ** 2 + (g.zc[None, :] - x2c) ** 2
    v = 1.0 - amp * np.exp(-r2 / (2.0 * width * width))
    vs = np.asfortranarray(np.repeat(v[:, None, :], ny, axis=1))
    us = np.zeros((nx, ny, nz), order="F")
    ws = np.zeros((nx, ny, nz), order="F")
    return us, vs, ws


# Per-step noise of the resident loop: sigma of the N(0,1) draws, Philox key (seed).
WAKE_SIGMA = 1e-2
WAKE_SEED = 5


def wake_768x384x256():
    """768x384x256; x1 tanh beta=1.5; x2, x3 uniform; wake profile (noise comes per step,
    Philox4x32-10 with key WAKE_SEED)."""
    g = make_grid(768, 384, 256, x_beta=1.5)
    return (g,) + wake(g)


BATCH_NX1 = (64, 96, 128, 192, 256, 384)
BATCH_NX2 = (64, 128, 192, 256)
BATCH_NX3 = (32, 64, 96, 128, 256)


def batch48_sizes(seed=4, count=48):
    """Sizes and x1 stretching of the 48-problem batch (seeded)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        m = int(rng.choice(BATCH_NX1))
        n = int(rng.choice(BATCH_NX2))
        l = int(rng.choice(BATCH_NX3))
        beta = float(rng.uniform(1.0, 2.5))
        out.append((m, n, l, beta))
    return out


def divergence(g, us, vs, ws):
    """DIVERGENCE (timeindeg.f:944-946) on the interior, numpy; returns (M, L, N) array."""
    ap = g.ap[1:-1, None, None]
    rp = g.rp[1:-1, None, None]
    ru = g.ru
    d = ap * (ru[1:-1, None, None] * us[1:-1, 1:-1, 1:-1] - ru[:-2, None, None] * us[:-2, 1:-1, 1:-1]) / rp
    d += g.bp[1:-1][None, :, None] * (vs[1:-1, 1:-1, 1:-1] - vs[1:-1, :-2, 1:-1]) / rp
    d += g.cp[1:-1][None, None, :] * (ws[1:-1, 1:-1, 1:-1] - ws[1:-1, 1:-1, :-2])
    return d


def rhs_from_velocity(g, us, vs, ws, dt=DT):
    """RHS = DIV(u*)/dt with the volume-weighted mean removed (solvability of the Neumann
    problem); returned as a full DP(NX,NY,NZ) array with zero ghosts."""
    b = divergence(g, us, vs, ws) / dt
    vol = g.cell_volume()
    b -= (vol * b).sum() / vol.sum()
    dp = np.zeros(g.shape(), order="F")
    dp[1:-1, 1:-1, 1:-1] = b
    return dp


DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def production_grid(stride=1, l=128):
    """The reference's production case (run/eddy.input: icyl=1, nx=302, ny=130, nz=1026):
    radial x1 from its run/x1_grid_half.in (300 cells from the axis), stretched axial x2 from
    run/x3_grid.in (1024 cells; every stride-th face), l azimuthal cells over 2 pi. The two grid
    files are the reference's own data, kept in eddyproj/data."""
    xf = grid.read_grid_file(os.path.join(DATA, "x1_grid_half.in"))
    zf = grid.read_grid_file(os.path.join(DATA, "x3_grid.in"))[::stride]
    return grid.from_faces(xf, zf, l, 2.0 * np.pi, cyl=True)


def production_302x130x1026(seed=21):
    """Production grid at full size with U(-1,1) velocities (seed 21)."""
    g = production_grid()
    u, v, w = random_uniform_velocity(g, seed)
    return g, u, v, w
