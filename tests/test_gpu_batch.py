"""Batched solves (one launch per stage over many uneven problems) against the
per-problem plans and the oracle."""
import numpy as np
import pytest

import eddyproj
import oracle
from eddyproj import configs, grid
from helpers import interior, rel_maxdiff_nomean

pytestmark = pytest.mark.gpu

# (M, N, L, x1 beta, x2 beta): uneven sizes, odd x2 lengths (dense x2 transform), and a
# stretched x2 of 128 cells, which the per-problem plans of the branch mode solve with the
# x2 transform as matrix-core products on the branch's stream
SHAPES = [(64, 32, 16, 1.3, None), (96, 64, 32, 2.1, None), (48, 40, 24, 1.0, 1.2), (128, 16, 64, 2.5, None),
          (32, 96, 12, 1.7, None), (20, 12, 10, 1.5, 1.1), (64, 128, 16, 1.4, 1.2)]


def say(msg):
    print(msg, flush=True)


def plan_tuning(g, mode):
    """Tuning of the single-problem plan that runs the batch's own arithmetic: the queue mode
    applies a stretched x2's eigenbasis with its tile kernel in every problem; the branch mode
    runs each problem's own default plan (matrix-core products once max(M, N) >= 128, the x1
    eigenbasis when M < N)."""
    return {"zgemm": -1} if mode == eddyproj.BATCH_QUEUES else None


def make(seed=21):
    probs = []
    for n, (m, nz, l, xb, zb) in enumerate(SHAPES):
        g = grid.make_grid(m, nz, l, x_beta=xb, z_beta=zb)
        us, vs, ws = configs.random_uniform_velocity(g, seed + n)
        probs.append((g, us, vs, ws, configs.rhs_from_velocity(g, us, vs, ws)))
    return probs


@pytest.mark.parametrize("mode", [eddyproj.BATCH_BRANCHES, eddyproj.BATCH_QUEUES])
def test_batch_fp64_direct_matches_plans_and_oracle(torch_cuda, mode):
    say("case batch fp64 direct")
    torch = torch_cuda
    probs = make()
    b = eddyproj.Batch([p[0] for p in probs])
    b.set_mode(mode)
    dps = [eddyproj.to_device(p[4]) for p in probs]
    b.bind(dps)
    b.direct()
    torch.cuda.synchronize()
    for (g, _, _, _, rhs), d in zip(probs, dps):
        plan = eddyproj.Plan(g, tuning=plan_tuning(g, mode))
        e = eddyproj.to_device(rhs)
        plan.direct(e)
        torch.cuda.synchronize()
        got, ref = eddyproj.to_host(d), eddyproj.to_host(e)
        # same tiles and arithmetic, inlined into different kernels: agree to rounding
        assert rel_maxdiff_nomean(got, ref) <= 1e-14
        exact = oracle.direct(g, rhs, extended=True)
        err = rel_maxdiff_nomean(got, exact)
        # The x2 eigenbasis of a stretched x2 of 128 cells is at FP64's conditioning limit: a
        # plain FP64 solve by the same method (the FP64 oracle) lands 6e-13 from the exact
        # answer there. The bar is 1e-12, or twice that method's own FP64 error when larger.
        bar = max(1e-12, 2.0 * rel_maxdiff_nomean(oracle.direct(g, rhs, path=1), exact))
        say("  %dx%dx%d err %.2e (bar %.2e)" % (g.M, g.N, g.L, err, bar))
        assert err <= bar
        plan.close()
    b.close()


@pytest.mark.parametrize("mode", [eddyproj.BATCH_BRANCHES, eddyproj.BATCH_QUEUES])
def test_batch_autotuned_tiles(torch_cuda, mode):
    """The batch autotuner's choice is a valid tuning of every problem, and the batch built with
    it solves to the oracle like the default one."""
    say("case batch autotune")
    from eddyproj import autotune
    torch = torch_cuda
    probs = make(51)
    grids = [p[0] for p in probs]
    tuning, trace = autotune.autotune_batch(grids, mode, reps=2)
    timed = [t for t in trace if "ms" in t]
    assert timed and all(t["ms"] > 0 for t in timed)
    for param, values in autotune.BATCH_SEARCH:
        assert tuning[param] in values
    b = eddyproj.Batch(grids, tuning=tuning)
    b.set_mode(mode)
    dps = [eddyproj.to_device(p[4]) for p in probs]
    b.bind(dps)
    b.direct()
    torch.cuda.synchronize()
    for (g, _, _, _, rhs), d in zip(probs, dps):
        exact = oracle.direct(g, rhs, extended=True)
        bar = max(1e-12, 2.0 * rel_maxdiff_nomean(oracle.direct(g, rhs, path=1), exact))
        err = rel_maxdiff_nomean(eddyproj.to_host(d), exact)
        assert err <= bar, (tuning, g.M, g.N, g.L, err, bar)
    b.close()


@pytest.mark.parametrize("mode", [eddyproj.BATCH_BRANCHES, eddyproj.BATCH_QUEUES])
def test_batch_fp32_refine_and_repeat(torch_cuda, mode):
    say("case batch fp32 refine")
    torch = torch_cuda
    probs = make(31)
    b = eddyproj.Batch([p[0] for p in probs])
    b.set_mode(mode)
    dps = [eddyproj.to_device(p[4]) for p in probs]
    b.bind(dps)
    for rep in range(2):  # a second call must start from the RHS again
        for d, p in zip(dps, probs):
            d.copy_(eddyproj.to_device(p[4]))
        torch.cuda.synchronize()
        b.direct(eddyproj.EDDY_FP32_REFINE, sweeps=2)
        torch.cuda.synchronize()
        for (g, _, _, _, rhs), d in zip(probs, dps):
            err = rel_maxdiff_nomean(eddyproj.to_host(d), oracle.direct(g, rhs, extended=True))
            assert err <= 1e-9, (rep, g.M, err)
    b.close()


@pytest.mark.parametrize("mode", [eddyproj.BATCH_BRANCHES, eddyproj.BATCH_QUEUES])
def test_batch_projection_matches_plans(torch_cuda, mode):
    say("case batch projection")
    torch = torch_cuda
    probs = make(41)
    b = eddyproj.Batch([p[0] for p in probs])
    b.set_mode(mode)
    dev = [[eddyproj.to_device(a) for a in p[1:4]] for p in probs]
    outs = [[t.clone() for t in v] for v in dev]
    dps = [torch.zeros_like(v[0]) for v in dev]
    b.bind(dps, [v[0] for v in dev], [v[1] for v in dev], [v[2] for v in dev], [o[0] for o in outs],
           [o[1] for o in outs], [o[2] for o in outs])
    plans = [eddyproj.Plan(p[0], tuning=plan_tuning(p[0], mode)) for p in probs]
    # two time steps of different ALFXDT in a row: the batch reads ALFXDT on the device, so the
    # same captured graph / job arrays must serve both
    for dt in (configs.DT, 2.5 * configs.DT):
        b.project(dt)
        torch.cuda.synchronize()
        for (g, *_), plan, v, o, d in zip(probs, plans, dev, outs, dps):
            ro = [t.clone() for t in v]
            rd = torch.zeros_like(v[0])
            plan.project(v[0], v[1], v[2], ro[0], ro[1], ro[2], rd, dt, remove_mean=True)
            torch.cuda.synchronize()
            # The batch folds the volume-weighted mean over its own workgroup shapes, so the
            # sums round in a different order than the single-problem plan: agree to rounding.
            err = rel_maxdiff_nomean(eddyproj.to_host(d), eddyproj.to_host(rd))
            assert err <= 1e-13, (dt, g.M, err)
            for a, c in zip(o, ro):
                a, c = eddyproj.to_host(a), eddyproj.to_host(c)
                assert np.abs(a - c).max() <= 1e-13 * max(np.abs(c).max(), 1.0), (dt, g.M)
    for plan in plans:
        plan.close()
    b.close()
