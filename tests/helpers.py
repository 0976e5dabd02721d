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


# ------------------------------------------------------------------ reference DIRECT fixtures
_GOLDEN = None


class FixtureGrid:
    """Grid of one case of tests/golden/direct_ref.npz (the metric arrays the reference ran on)."""

    def __init__(self, d, name):
        nx, ny, nz, mp, lp, np_, ipres, nprocs = (int(v) for v in d[name + "/dims"])
        self.nx, self.ny, self.nz = nx, ny, nz
        self.mp, self.lp, self.np = mp, lp, np_
        self.ipres, self.nprocs = ipres, nprocs
        self.delysq = float(d[name + "/delysq"][0])
        for k in ("ap", "au", "rp", "ru", "cp", "cw"):
            setattr(self, k, np.array(d[name + "/" + k], dtype=np.float64))
        self.bp = np.full(ny, 1.0 / np.sqrt(self.delysq))
        self.bv = self.bp.copy()
        self.delzsq = (1.0 / self.cp[1]) ** 2  # the DELZSQ the reference's IPRES=3 cases ran with

    @property
    def M(self):
        return self.nx - 2

    @property
    def L(self):
        return self.ny - 2

    @property
    def N(self):
        return self.nz - 2

    def shape(self):
        return (self.nx, self.ny, self.nz)

    def cell_volume(self):
        v = (self.rp[1:-1] / self.ap[1:-1])[:, None, None] / self.bp[1:-1][None, :, None]
        return v / self.cp[1:-1][None, None, :]


def _golden():
    global _GOLDEN
    if _GOLDEN is None:
        _GOLDEN = np.load(golden_case.path)
    return _GOLDEN


def golden_cases(kind=None):
    """Names of the full-field cases; kind 'walls' (MP = NP = 1) or 'periodic' filters."""
    out = []
    for n in _golden()["cases"]:
        n = str(n)
        if n + "/dp" not in _golden():
            continue
        g = FixtureGrid(_golden(), n)
        if kind == "walls" and not (g.mp == 1 and g.np == 1):
            continue
        if kind == "periodic" and g.mp == 1 and g.np == 1:
            continue
        out.append(n)
    return out


def golden_case(name):
    """(grid, DP the reference returned) of a fixture case."""
    d = _golden()
    return FixtureGrid(d, name), np.array(d[name + "/dp"], order="F")


golden_case.path = __import__("os").path.join(__import__("os").path.dirname(__import__("os").path.abspath(__file__)),
                                              "golden", "direct_ref.npz")


def golden_rhs(name):
    """The RHS of a fixture case as a full DP(NX,NY,NZ) array (zero ghosts)."""
    d = _golden()
    g = FixtureGrid(d, name)
    rhs = np.zeros(g.shape(), order="F")
    rhs[1:-1, 1:-1, 1:-1] = d[name + "/rhs"]
    return rhs
