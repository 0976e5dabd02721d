"""The resident time loop on the device (Philox noise -> projection -> divergence log, one
captured hipGraph per step) against the oracle's CPU loop."""
import numpy as np
import pytest

import eddyproj
import oracle
from eddyproj import configs, grid

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def test_perturb_matches_oracle(torch_cuda):
    say("case perturb 20x12x10 step 7 seed 5")
    torch = torch_cuda
    g = grid.make_grid(20, 12, 10, x_beta=1.5)
    u, v, w = configs.random_uniform_velocity(g, 3)
    plan = eddyproj.Plan(g)
    du, dv, dw = (eddyproj.to_device(a) for a in (u, v, w))
    out = [torch.full_like(du, np.nan) for _ in range(3)]
    step = torch.tensor([7], dtype=torch.int32, device="cuda")
    plan.perturb(du, dv, dw, *out, 1e-2, 5, step)
    torch.cuda.synchronize()
    ref = oracle.perturb(g, u, v, w, 1e-2, 5, 7)
    for name, got, want, base in zip("uvw", out, ref, (u, v, w)):
        got = eddyproj.to_host(got)
        copied = want == base
        np.testing.assert_array_equal(got[copied], want[copied])
        err = np.abs(got - want).max()
        say("  %s max |gpu - oracle| %.2e" % (name, err))
        # libm vs device log/sin/cos differ in the last bits of sigma*g (~1e-18); u + sigma*g
        # can then round to the neighbouring double of u's magnitude: one ulp of max |u|
        assert err <= np.spacing(np.abs(want).max())
    plan.close()


@pytest.mark.parametrize("precision,tol", [(eddyproj.EDDY_FP64, 1e-12), (eddyproj.EDDY_FP32_REFINE, 1e-11)])
def test_resident_loop_matches_oracle(torch_cuda, precision, tol):
    say("case resident 32x24x16 precision %d" % precision)
    torch = torch_cuda
    g = grid.make_grid(32, 24, 16, x_beta=1.5)
    u, v, w = configs.wake(g)
    nsteps = 4
    ru, rv, rw, rlog = oracle.resident_steps(g, u, v, w, nsteps, configs.WAKE_SIGMA, configs.WAKE_SEED, configs.DT,
                                             extended=True)
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    try:
        plan = eddyproj.Plan(g, stream=stream)
        du, dv, dw = (eddyproj.to_device(a) for a in (u, v, w))
        dp = torch.zeros_like(du)
        torch.cuda.synchronize()
        res = eddyproj.Resident(plan, du, dv, dw, dp, configs.DT, configs.WAKE_SIGMA, configs.WAKE_SEED,
                                precision=precision, sweeps=2, max_steps=nsteps)
        res.run(2)
        res.run(2)
        with pytest.raises(eddyproj.EddyError):
            res.run(1)  # past max_steps
        n, log = res.log()
        torch.cuda.synchronize()
        assert n == nsteps and log.shape == (nsteps, 2)
        for name, got, want in zip("uvw", (du, dv, dw), (ru, rv, rw)):
            got = eddyproj.to_host(got)
            err = np.abs(got - want).max() / np.abs(want).max()
            say("  %s rel err %.2e" % (name, err))
            assert err <= tol
        say("  div log gpu %s oracle %s" % (np.abs(log).max(axis=1), np.abs(rlog).max(axis=1)))
        assert np.abs(log).max() <= 1e-12
        # the logged range is the standalone divergence-range kernel's on the final field
        assert tuple(log[-1]) == plan.div_minmax(du, dv, dw)
        res.close()
        plan.close()
    finally:
        torch.cuda.set_stream(torch.cuda.default_stream())


@pytest.mark.parametrize("cyl", [False, True])
def test_resident_log_equals_standalone_divergence_range(torch_cuda, cyl):
    """After every step the logged divergence range equals, bit for bit, the standalone
    kernel's range of the corrected field left in (u, v, w), on a grid whose divergence is not
    roundoff (random velocities, ragged x1 blocks), Cartesian and cylindrical: the log belongs
    to the step just run and to the field the caller holds."""
    say("case resident divergence range 70x9x12 cyl=%s" % cyl)
    torch = torch_cuda
    g = grid.make_grid(70, 9, 12, x_beta=1.5, z_beta=1.2, cyl=cyl, lx3=2.0 * np.pi if cyl else 2.0)
    u, v, w = configs.random_uniform_velocity(g, 8)
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    try:
        plan = eddyproj.Plan(g, stream=stream)
        du, dv, dw = (eddyproj.to_device(a) for a in (u, v, w))
        dp = torch.zeros_like(du)
        torch.cuda.synchronize()
        res = eddyproj.Resident(plan, du, dv, dw, dp, configs.DT, 0.3, 11, max_steps=3)
        for s in range(3):
            res.run(1)
            n, log = res.log()
            want = plan.div_minmax(du, dv, dw)
            say("  step %d logged %s standalone %s" % (s, tuple(log[-1]), want))
            assert n == s + 1 and tuple(log[-1]) == want
        res.close()
        plan.close()
    finally:
        torch.cuda.set_stream(torch.cuda.default_stream())


def test_wake_768x384x256_resident_drift(torch_cuda):
    """Full size: 20 FP64 steps; the logged post-projection divergence stays at roundoff
    relative to the divergence the noise injects, and does not grow."""
    say("case resident wake_768x384x256 20 steps")
    torch = torch_cuda
    g, u, v, w = configs.wake_768x384x256()
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    try:
        plan = eddyproj.Plan(g, stream=stream)
        du, dv, dw = (eddyproj.to_device(a) for a in (u, v, w))
        dp = torch.zeros_like(du)
        star = [torch.zeros_like(du) for _ in range(3)]
        step = torch.zeros(1, dtype=torch.int32, device="cuda")
        plan.perturb(du, dv, dw, *star, configs.WAKE_SIGMA, configs.WAKE_SEED, step)
        d0 = plan.div_minmax(*star)
        torch.cuda.synchronize()
        del star
        res = eddyproj.Resident(plan, du, dv, dw, dp, configs.DT, configs.WAKE_SIGMA, configs.WAKE_SEED,
                                max_steps=20)
        res.run(20)
        n, log = res.log()
        scale = max(abs(d0[0]), abs(d0[1]))
        drift = np.abs(log).max(axis=1) / scale
        say("  injected divergence %.3e; relative post-projection divergence first %.2e last %.2e max %.2e"
            % (scale, drift[0], drift[-1], drift.max()))
        assert n == 20
        assert drift.max() <= 1e-12
        assert drift[-5:].max() <= 10 * max(drift[:5].max(), 1e-15)
        res.close()
        plan.close()
    finally:
        torch.cuda.set_stream(torch.cuda.default_stream())
