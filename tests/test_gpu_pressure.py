"""The pressure update after CORRECT (Eddy6.f:2772-2846) on the device against the oracle."""
import numpy as np
import pytest

import eddyproj
import oracle
from eddyproj import configs, grid

pytestmark = pytest.mark.gpu

CASES = [
    # (grid kwargs, update kwargs)
    (dict(m=24, n=10, l=16), dict()),
    (dict(m=24, n=10, l=16), dict(itype5=500, itype6=500)),
    (dict(m=33, n=12, l=8, z_beta=1.2), dict(implx=1, imply=1, ru1=0.02)),
    (dict(m=20, n=8, l=12, cyl=True, lx3=2 * np.pi), dict(itype1=300, implx=1, imply=1, ru1=0.01)),
]


@pytest.mark.parametrize("gk,uk", CASES)
def test_pressure_update_matches_oracle(torch_cuda, gk, uk):
    print("case pressure %s %s" % (gk, uk), flush=True)
    torch = torch_cuda
    m, n, l = gk.pop("m"), gk.pop("n"), gk.pop("l")
    g = grid.make_grid(m, n, l, **gk)
    rng = np.random.default_rng(7)
    p = np.asfortranarray(rng.standard_normal(g.shape()))
    us, vs, ws = configs.random_uniform_velocity(g, 8)
    dp = oracle.direct(g, configs.rhs_from_velocity(g, us, vs, ws))  # ghosts as DIRECT leaves them
    masks = {}
    if uk.get("implx"):
        masks["rvimplx"] = (rng.random(g.nx) > 0.3).astype(float)
    if uk.get("imply"):
        masks["rvimply"] = (rng.random(g.nx) > 0.3).astype(float)
    ref = oracle.pressure_update(g, p, dp, configs.DT, **uk, **masks)
    plan = eddyproj.Plan(g)
    plan.set_pressure_bc(**uk, **masks)
    dpd, pd = eddyproj.to_device(dp), eddyproj.to_device(p)
    plan.pressure_update(pd, dpd, configs.DT)
    torch.cuda.synchronize()
    got = eddyproj.to_host(pd)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("  rel err %.2e" % err, flush=True)
    # one rounding of PREF's sum order and of contracted multiply-adds
    assert err <= 1e-14
    plan.close()
