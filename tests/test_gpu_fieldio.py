"""Field files between disk and HBM (the staged device paths of the native reader/writer)
against the host paths, which the CPU tests pin to Fortran-written fixtures. Fields larger
than one 32 MiB staging buffer exercise the double-buffered runs of records."""
import os

import numpy as np
import pytest

import eddyproj
from eddyproj import configs, fieldio

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fields")


def say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("shape", [(6, 5, 8), (514, 258, 70)])
def test_scalar_device_read_write(torch_cuda, tmp_path, shape):
    say("case ioscalar device %s" % (shape,))
    torch = torch_cuda
    rng = np.random.default_rng(11)
    p = np.asfortranarray(rng.standard_normal(shape))
    a, b = str(tmp_path / "host.res"), str(tmp_path / "dev.res")
    fieldio.write_scalar(a, p, nstep=7, time=0.5, dtm1=1e-3, grav=9.81)
    d, h = fieldio.read_scalar(a, device=True)
    torch.cuda.synchronize()
    assert h["nstep"] == 7
    np.testing.assert_array_equal(eddyproj.to_host(d), p)
    fieldio.write_scalar(b, d, nstep=7, time=0.5, dtm1=1e-3, grav=9.81)
    assert open(a, "rb").read() == open(b, "rb").read()
    # a rank slab of the same file (two ranks): planes from nz_local-2 on
    nzl = (shape[2] - 2) // 2 + 2
    d1, _ = fieldio.read_scalar(a, shape[0], shape[1], nzl, myrank=1, mysize=2, device=True)
    np.testing.assert_array_equal(eddyproj.to_host(d1), p[:, :, nzl - 2:2 * nzl - 2])


@pytest.mark.parametrize("shape", [(6, 5, 8), (514, 258, 40)])
def test_vector_device_read_write(torch_cuda, tmp_path, shape):
    say("case iovector device %s" % (shape,))
    torch = torch_cuda
    rng = np.random.default_rng(12)
    f = [np.asfortranarray(rng.standard_normal(shape)) for _ in range(3)]
    a, b = str(tmp_path / "host.res"), str(tmp_path / "dev.res")
    fieldio.write_vector(a, *f, time=2.0, per_rank=True)
    d, h = fieldio.read_vector(a, device=True)
    torch.cuda.synchronize()
    assert h["time"] == 2.0
    for got, want in zip(d, f):
        np.testing.assert_array_equal(eddyproj.to_host(got), want)
    fieldio.write_vector(b, *d, time=2.0, per_rank=True)
    assert open(a, "rb").read() == open(b, "rb").read()


def test_fortran_fixture_to_device(torch_cuda):
    say("case Fortran-written fixture into HBM")
    d, _ = fieldio.read_scalar(os.path.join(G, "ioscalar_dir1.res"), device=True)
    h, _ = fieldio.read_scalar(os.path.join(G, "ioscalar_dir1.res"))
    np.testing.assert_array_equal(eddyproj.to_host(d), h)


def test_projection_of_a_restart_set(torch_cuda, tmp_path):
    """Restart files written and read back into HBM give the same projection, bit for bit, as
    the fields they were written from (the ingest path of src/Eddy6.f:780-796)."""
    say("case restart set -> projection 64x32x48")
    torch = torch_cuda
    from eddyproj import grid
    g = grid.make_grid(64, 32, 48, x_beta=1.5)
    u, v, w = configs.random_uniform_velocity(g, 5)
    d = str(tmp_path) + os.sep
    fieldio.save_restart(d, 300, {"u": u, "v": v, "w": w}, nstep=300, time=0.3, dtm1=configs.DT)
    got, h = fieldio.load_restart(d, 300, names=("u", "v", "w"), device=True)
    assert h["nstep"] == 300
    plan = eddyproj.Plan(g)
    outs = []
    for src in ((got["u"], got["v"], got["w"]), tuple(eddyproj.to_device(a) for a in (u, v, w))):
        uo, vo, wo = (s.clone() for s in src)
        dp = torch.zeros_like(src[0])
        plan.project(*src, uo, vo, wo, dp, configs.DT)
        outs.append([eddyproj.to_host(t) for t in (uo, vo, wo, dp)])
    torch.cuda.synchronize()
    for x, y in zip(*outs):
        np.testing.assert_array_equal(x, y)
    plan.close()
