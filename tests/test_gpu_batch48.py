"""The 48-problem batch workload in both batch modes, every problem against the long-double
oracle.

Workload: 48 problems with seeded sizes (seed 4) nx1 in {64, 96, 128, 192, 256, 384},
nx2 in {64, 128, 192, 256}, nx3 in {32, 64, 96, 128, 256}, x1 tanh-stretched with
beta ~ U(1.0, 2.5) per problem, x2 and x3 uniform; RHS = DIV(u)/dt of an i.i.d. U(-1, 1)
velocity field (seed 400 + problem index) with the volume-weighted mean removed. Each problem
is DIRECT's per-mode loop (direct.f:710-981) on its own grid.
"""
import os

import numpy as np
import pytest

import eddyproj
import oracle
from eddyproj import configs, grid
from helpers import interior, rel_maxdiff_nomean

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def batch48():
    say("case batch48: building 48 problems and their long-double oracle answers")
    oracle.set_threads(min(16, os.cpu_count() or 1))
    probs = []
    for n, (m, nz, l, beta) in enumerate(configs.batch48_sizes()):
        g = grid.make_grid(m, nz, l, x_beta=beta)
        us, vs, ws = configs.random_uniform_velocity(g, 400 + n)
        rhs = configs.rhs_from_velocity(g, us, vs, ws)
        del us, vs, ws
        probs.append((g, rhs, oracle.direct(g, rhs, extended=True)))
    say("  %d problems, %d cells" % (len(probs), sum(p[0].M * p[0].N * p[0].L for p in probs)))
    return probs


def _run(torch, probs, mode, precision, sweeps):
    b = eddyproj.Batch([p[0] for p in probs])
    b.set_mode(mode)
    dps = [eddyproj.to_device(p[1]) for p in probs]
    b.bind(dps)
    b.direct(precision, sweeps)
    torch.cuda.synchronize()
    out = [eddyproj.to_host(d) for d in dps]
    b.close()
    del dps
    return out


@pytest.mark.parametrize("mode", [eddyproj.BATCH_BRANCHES, eddyproj.BATCH_QUEUES])
def test_batch48_fp64_every_problem_matches_oracle(torch_cuda, batch48, mode):
    say("case batch48 fp64 mode %d" % mode)
    got = _run(torch_cuda, batch48, mode, eddyproj.EDDY_FP64, 0)
    worst = 0.0
    for (g, rhs, ref), phi in zip(batch48, got):
        err = rel_maxdiff_nomean(phi, ref)
        worst = max(worst, err)
        assert err <= 1e-12, (g.M, g.N, g.L, err)
    say("  fp64: worst rel max err vs oracle %.2e" % worst)


@pytest.mark.parametrize("mode", [eddyproj.BATCH_BRANCHES, eddyproj.BATCH_QUEUES])
def test_batch48_fp32_refine_residual_and_oracle_error(torch_cuda, batch48, mode):
    say("case batch48 fp32+2 mode %d" % mode)
    got = _run(torch_cuda, batch48, mode, eddyproj.EDDY_FP32_REFINE, 2)
    worst_res = worst_max = worst_l2 = 0.0
    for (g, rhs, ref), phi in zip(batch48, got):
        b = interior(rhs)
        r = b - interior(oracle.laplacian(g, phi))
        res = float(np.linalg.norm(r) / np.linalg.norm(b))
        a = interior(phi) - interior(phi).mean()
        e = interior(ref) - interior(ref).mean()
        emax = float(np.abs(a - e).max() / np.abs(e).max())
        el2 = float(np.linalg.norm(a - e) / np.linalg.norm(e))
        worst_res, worst_max, worst_l2 = max(worst_res, res), max(worst_max, emax), max(worst_l2, el2)
        assert res <= 1e-10, (g.M, g.N, g.L, res)
        assert emax <= 1e-9, (g.M, g.N, g.L, emax)
    say("  fp32+2: worst residual %.2e, worst error vs oracle: max %.2e, L2 %.2e" % (worst_res, worst_max, worst_l2))
