"""eddy_direct_noalloc_dev: FP64 DIRECT whose work buffers come from the library's fixed
device array (the call allocates nothing), against eddy_direct_dev on the same plans: two calls
in a row, the second on a smaller grid, and the size limit."""
import ctypes

import numpy as np
import pytest

import eddyproj
import oracle
from eddyproj import configs, grid
from helpers import rel_maxdiff_nomean

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def noalloc(plan, t):
    L = eddyproj.lib()
    L.eddy_direct_noalloc_dev.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    return L.eddy_direct_noalloc_dev(plan.handle, ctypes.c_void_p(t.data_ptr()))


def solve_both(torch, g, rhs, tuning=None):
    plan = eddyproj.Plan(g, tuning=tuning)
    a = eddyproj.to_device(rhs)
    b = eddyproj.to_device(rhs)
    plan.direct(a)
    rc = noalloc(plan, b)
    assert rc == 0, eddyproj.lib().eddy_last_error()
    torch.cuda.synchronize()
    plan.close()
    return eddyproj.to_host(a), eddyproj.to_host(b)


def test_noalloc_matches_direct_dev_twice(torch_cuda):
    torch = torch_cuda
    # a larger grid, then a smaller one: the second call reuses the array the first left dirty
    # (the stretched x2 with the fused tile kernel: the GEMM forms keep the plan's buffers)
    for (m, n, l, zb) in [(160, 96, 64, 1.2), (48, 40, 24, None)]:
        say("case noalloc %dx%dx%d" % (m, n, l))
        g = grid.make_grid(m, n, l, x_beta=1.5, z_beta=zb)
        us, vs, ws = configs.random_uniform_velocity(g, m + l)
        rhs = configs.rhs_from_velocity(g, us, vs, ws)
        ref, got = solve_both(torch, g, rhs, tuning={"zgemm": -1})
        np.testing.assert_array_equal(got, ref)  # same kernels and arithmetic
        assert rel_maxdiff_nomean(got, oracle.direct(g, rhs, extended=True)) <= 1e-12


def test_noalloc_rejects_beyond_scratch(torch_cuda):
    say("case noalloc size limit")
    g = grid.make_grid(256, 128, 264, x_beta=1.5)  # 8.65 M cells: above the 64 MiB array
    plan = eddyproj.Plan(g)
    d = torch_cuda.zeros((g.nz, g.ny, g.nx), dtype=torch_cuda.float64, device="cuda")
    rc = noalloc(plan, d)
    assert rc == -2  # EDDY_ERR_UNSUPPORTED
    assert b"64 MiB" in eddyproj.lib().eddy_last_error()
    plan.close()


def test_noalloc_rejects_gemm_forms(torch_cuda):
    say("case noalloc GEMM form")
    g = grid.make_grid(64, 160, 16, x_beta=1.5, z_beta=1.2)  # x1 form (M < N, N >= 128)
    plan = eddyproj.Plan(g)
    d = torch_cuda.zeros((g.nz, g.ny, g.nx), dtype=torch_cuda.float64, device="cuda")
    assert noalloc(plan, d) == -2
    assert b"GEMM" in eddyproj.lib().eddy_last_error()
    plan.close()
