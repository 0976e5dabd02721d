"""Staggered-grid metrics for the projection path (host side).

Mirrors the reference's metric construction in src/grid.f:
  * stretched wall-bounded axis from face coordinates: grid.f:45-58 (x) and :136-143 (z);
  * uniform axis: grid.f:66-82 (x), :99-111 (y), :159-172 (z);
  * Cartesian radii RP = RU = 1: grid.f:470-472; cylindrical RU = XU, RP = XC: grid.f:462-466.

Array conventions follow the reference: AP, AU, RP, RU have length NX (1-based I in Fortran,
index I-1 here); BP, BV length NY; CP, CW length NZ (= NZG on one rank). The pressure array
is DP(NX,NY,NZ) with one ghost layer on each side, so the interior counts are
M = NX-2 (x1, walls), L = NY-2 (x3, periodic, the FFT axis), N = NZ-2 (x2, walls).
"""
import numpy as np


def faces_uniform(n, length, origin=0.0):
    return origin + length * np.arange(n + 1, dtype=np.float64) / n


def faces_tanh_onesided(n, length, beta):
    """x_i = length * (1 + tanh(beta*(i/n - 1)) / tanh(beta)), i = 0..n (clustered at x=length)."""
    i = np.arange(n + 1, dtype=np.float64)
    return length * (1.0 + np.tanh(beta * (i / n - 1.0)) / np.tanh(beta))


def faces_tanh_twosided(n, length, beta):
    """x_i = (length/2) * (1 + tanh(beta*(2i/n - 1)) / tanh(beta)), clustered at both walls."""
    i = np.arange(n + 1, dtype=np.float64)
    return 0.5 * length * (1.0 + np.tanh(beta * (2.0 * i / n - 1.0)) / np.tanh(beta))


def metrics_stretched(faces):
    """AP, AU (and centre coordinates) from face coordinates xu(1..NX-1) (grid.f:45-58)."""
    xu_in = np.asarray(faces, dtype=np.float64)
    nx = xu_in.size + 1
    xu = np.zeros(nx + 1)
    xu[1:nx] = xu_in
    ix1, ix2 = 2, nx - 1
    xu[nx] = 2.0 * xu[ix2] - xu[ix2 - 1]
    au = np.zeros(nx + 1)
    ap = np.zeros(nx + 1)
    xc = np.zeros(nx + 1)
    xc[ix1:ix2 + 1] = 0.5 * (xu[ix1 - 1:ix2] + xu[ix1:ix2 + 1])
    xc[1] = 2.0 * xu[1] - xc[2]
    xc[nx] = 2.0 * xu[ix2] - xc[ix2]
    au[ix1:ix2 + 1] = 2.0 / (xu[ix1 + 1:ix2 + 2] - xu[ix1 - 1:ix2])
    au[1] = 2.0 / (-3.0 * xu[1] + 4.0 * xu[2] - xu[3])
    au[ix2] = 2.0 / (3.0 * xu[ix2] - 4.0 * xu[ix2 - 1] + xu[ix2 - 2])
    au[nx] = au[ix2]
    ap[ix1:ix2 + 1] = 1.0 / (xu[ix1:ix2 + 1] - xu[ix1 - 1:ix2])
    ap[1] = ap[ix1]
    ap[nx] = ap[ix2]
    return ap[1:].copy(), au[1:].copy(), xu[1:].copy(), xc[1:].copy()


def metrics_uniform(n_interior, length, origin=0.0):
    """Uniform axis (grid.f:66-82 / :159-172): every metric equals 1/delta."""
    nx = n_interior + 2
    d = length / float(nx - 2)
    r = 1.0 / d
    xu = origin + d * np.arange(nx, dtype=np.float64)
    xc = np.zeros(nx)
    xc[1:nx - 1] = 0.5 * (xu[0:nx - 2] + xu[1:nx - 1])
    xc[0] = 2.0 * xu[0] - xc[1]
    xc[nx - 1] = 2.0 * xu[nx - 2] - xc[nx - 2]
    return np.full(nx, r), np.full(nx, r), xu, xc


class Grid:
    """Metrics of one DP(NX,NY,NZ) problem (walls in x1/x2, periodic x3)."""

    def __init__(self, x_axis, z_axis, ny_interior, ly):
        self.ap, self.au, self.xu, self.xc = x_axis
        self.cp, self.cw, self.zw, self.zc = z_axis
        self.nx = self.ap.size
        self.nz = self.cp.size
        self.ny = ny_interior + 2
        dely = ly / float(self.ny - 2)
        self.dely = dely
        self.delysq = dely * dely
        self.bp = np.full(self.ny, 1.0 / dely)
        self.bv = self.bp.copy()
        self.yv = dely * np.arange(self.ny, dtype=np.float64)
        self.yc = self.yv - 0.5 * dely
        self.rp = np.ones(self.nx)
        self.ru = np.ones(self.nx)
        self.mp, self.lp, self.np = 1, 0, 1
        self.delzsq = (1.0 / self.cp[1]) ** 2

    @property
    def M(self):
        return self.nx - 2

    @property
    def L(self):
        return self.ny - 2

    @property
    def N(self):
        return self.nz - 2

    @property
    def cells(self):
        return self.M * self.L * self.N

    def shape(self):
        return (self.nx, self.ny, self.nz)

    def cell_volume(self):
        """RP/(AP*BP*CP) on the interior, shape (M, L, N) (solvability weights)."""
        v = (self.rp[1:-1] / self.ap[1:-1])[:, None, None] / self.bp[1:-1][None, :, None]
        return v / self.cp[1:-1][None, None, :]

    def z_uniform(self):
        c = self.cp[1:-1]
        w = self.cw[1:-1]
        return bool(np.all(c == c[0]) and np.all(w == c[0]))


def read_grid_file(path):
    """Face coordinates from a grid file of the reference (read by grid.f:38-40, 121-124): a
    first line with the count ix2 = NX-1, then ix2 lines "i x_i"."""
    with open(path) as f:
        n = int(f.readline().split()[0])
        vals = [float(line.split()[1]) for line in f if line.strip()]
    if len(vals) != n:
        raise ValueError("%s: %d coordinates, header says %d" % (path, len(vals), n))
    return np.array(vals, dtype=np.float64)


def from_faces(x_faces, z_faces, l, ly, cyl=False):
    """Grid from face coordinates of x1 (NX-1 values) and x2 (NZ-1 values), l periodic x3 cells
    of total length ly; cyl as in make_grid (x1 the radius, RU = XU, RP = XC)."""
    g = Grid(metrics_stretched(x_faces), metrics_stretched(z_faces), l, ly)
    if cyl:
        g.ru = g.xu.copy()
        g.rp = g.xc.copy()
    return g


def make_grid(m, n, l, lx1=4.0, lx2=2.0, lx3=2.0, x_beta=1.5, z_beta=None, x_stretch=True, cyl=False):
    """Grid of m x n x l interior cells in (x1, x2, x3) = (X, Z, Y) of the reference.

    x1 is one-sided tanh-stretched with `x_beta` (uniform if x_stretch is False); x2 is
    uniform unless `z_beta` is given (two-sided tanh); x3 is uniform and periodic. With
    cyl=True x1 is the radius from the axis (x1 = 0) and x3 the azimuth: RU = XU and RP = XC
    (grid.f:462-466), lx3 the angular period.
    """
    if x_stretch:
        xa = metrics_stretched(faces_tanh_onesided(m, lx1, x_beta))
    else:
        xa = metrics_uniform(m, lx1)
    if z_beta is None:
        za = metrics_uniform(n, lx2)
    else:
        za = metrics_stretched(faces_tanh_twosided(n, lx2, z_beta))
    g = Grid(xa, za, l, lx3)
    if cyl:
        g.ru = g.xu.copy()
        g.rp = g.xc.copy()
    return g
