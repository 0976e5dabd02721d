"""Writes tests/golden/direct_ref.npz: outputs of the reference's own DIRECT.

The reference's src/direct.f (with pdc2dn.f, fftpack.f, numrec.f, genbun.f, gnbnaux.f,
comf.f, info.f) is compiled from its sources by `make -C oracle cpuref` and run under
`mpiexec -np P` by oracle/cpuref.py, called as src/Eddy6.f:2722 calls it (one z slab per rank).
Two of its solver branches run without the absent BLKTRI source:
  IPRES=4: pdc2dn (PSCR with LAPACK eigenpairs) for walls in x1 and x2 (MP = NP = 1), any
           periodic length L (fftpack's mixed radix 4, 2, 3, 5 + general odd factors);
  IPRES=3: an FFT along a uniform periodic x2 (spacing from DELZSQ) + TRIDAG per line
           (direct.f:834-869), i.e. NP = 0 with x1 walls.
Each case stores the inputs (dimensions, flags, metrics, RHS) and the DP the reference
returned. Full fields for the small grids; for the 64x64x32 workload (seed 1) the RHS is
rebuilt from its seed and only sampled values and norms of the answer are stored.

Run in this container only (needs /root/reference and /opt/conda):
    make -C oracle cpuref && python tests/golden/make_direct_ref_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "eddy-hybrid_amd"), os.path.join(ROOT, "oracle")]

import cpuref  # noqa: E402
from eddyproj import configs, grid  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "direct_ref.npz")


def random_rhs(g, seed):
    rng = np.random.default_rng(seed)
    b = rng.uniform(-1.0, 1.0, (g.M, g.L, g.N))
    vol = g.cell_volume()
    b -= (vol * b).sum() / vol.sum()
    rhs = np.zeros(g.shape(), order="F")
    rhs[1:-1, 1:-1, 1:-1] = b
    return rhs


def velocity_rhs(g, seed):
    us, vs, ws = configs.random_uniform_velocity(g, seed)
    return configs.rhs_from_velocity(g, us, vs, ws)


# name, grid builder, RHS builder, IPRES, ranks
CASES = [
    ("walls_16x8x8", lambda: grid.make_grid(16, 8, 8, x_beta=1.5), lambda g: velocity_rhs(g, 11), 4, 1),
    ("walls_z_stretched_24x16x12", lambda: grid.make_grid(24, 16, 12, x_beta=1.5, z_beta=1.2),
     lambda g: velocity_rhs(g, 12), 4, 2),
    ("cylindrical_20x12x8", lambda: grid.make_grid(20, 12, 8, x_beta=1.3, cyl=True, lx3=2.0 * np.pi),
     lambda g: velocity_rhs(g, 13), 4, 2),
    ("odd_L15_16x9x15", lambda: grid.make_grid(16, 9, 15, x_beta=1.5), lambda g: random_rhs(g, 14), 4, 3),
    ("L14_half7_12x6x14", lambda: grid.make_grid(12, 6, 14, x_beta=1.5), lambda g: random_rhs(g, 15), 4, 2),
    ("L22_half11_z_stretched_10x12x22", lambda: grid.make_grid(10, 12, 22, x_beta=1.5, z_beta=1.1),
     lambda g: random_rhs(g, 16), 4, 1),
    ("odd_L21_12x7x21", lambda: grid.make_grid(12, 7, 21, x_beta=2.0), lambda g: random_rhs(g, 17), 4, 1),
    ("periodic_x2_16x12x8", lambda: _periodic_z(grid.make_grid(16, 12, 8, x_beta=1.5)), lambda g: random_rhs(g, 18),
     3, 2),
    ("periodic_x2_odd_L_12x10x9", lambda: _periodic_z(grid.make_grid(12, 10, 9, x_beta=1.2)),
     lambda g: random_rhs(g, 19), 3, 1),
]


def _periodic_z(g):
    g.np = 0
    return g


def main():
    if not cpuref.available():
        raise SystemExit("build the reference DIRECT first: make -C oracle cpuref")
    data = {}
    names = []
    for name, mk, mkrhs, ipres, nprocs in CASES:
        g = mk()
        rhs = mkrhs(g)
        r = cpuref.run(g, nprocs, rhs=rhs, ipres=ipres, want_dp=True)
        names.append(name)
        data[name + "/dims"] = np.array([g.nx, g.ny, g.nz, g.mp, g.lp, g.np, ipres, nprocs], dtype=np.int32)
        data[name + "/delysq"] = np.array([g.delysq])
        for k in ("ap", "au", "rp", "ru", "cp", "cw"):
            data[name + "/" + k] = getattr(g, k)
        data[name + "/rhs"] = rhs[1:-1, 1:-1, 1:-1].copy()
        data[name + "/dp"] = r["dp"]
        print("%-34s ranks=%d first call %.3e s" % (name, nprocs, r["first_s"]), flush=True)
    # the 64x64x32 workload: U(-1,1) velocities, seed 1; RHS rebuilt from the seed by the tests
    g, us, vs, ws = configs.random_64x64x32()
    rhs = configs.rhs_from_velocity(g, us, vs, ws)
    r = cpuref.run(g, 4, rhs=rhs, ipres=4, want_dp=True)
    phi = r["dp"][1:-1, 1:-1, 1:-1]
    phi = phi - phi.mean()
    rng = np.random.default_rng(2024)
    flat = rng.choice(phi.size, 4096, replace=False)
    name = "random_64x64x32"
    names.append(name)
    data[name + "/sample_index"] = flat.astype(np.int64)  # into the (M, L, N) interior, C order
    data[name + "/sample_value"] = phi.reshape(-1)[flat]
    data[name + "/norms"] = np.array([np.sqrt((phi * phi).sum()), np.abs(phi).max()])
    data[name + "/ghosts"] = np.array([np.abs(r["dp"][0] - r["dp"][1]).max(), np.abs(r["dp"][:, 0] - r["dp"][:, -2]).max(),
                                       np.abs(r["dp"][:, :, 0]).max()])
    print("%-34s ranks=4 first call %.3e s" % (name, r["first_s"]), flush=True)
    data["cases"] = np.array(names)
    np.savez_compressed(OUT, **data)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
