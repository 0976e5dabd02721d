"""Autotuner of the projection kernels.

Picks, per grid, the launch parameters the kernels expose (include/eddyproj.h, eddy_tuning):
  * yfft_pencils   -- x1 pencils per workgroup of the y-FFT (pencil width),
  * ztrans_pencils -- x1 lines per workgroup of the x2 transform,
  * xsolve_chunk   -- rows per thread of the x1 solve, i.e. where parallel cyclic reduction
                      hands over to Thomas elimination (P = M / chunk threads per line do PCR
                      on the 2P chunk ends, each thread runs Thomas on its chunk),
  * fft_radix      -- the largest radix of the FFT radix plan (8, 4 or 2).
Each parameter only changes its own kernels, so the search is coordinate-wise: every
candidate value is timed with the HIP-event kernel timers of the plan (median of several
solves), and the best value is kept before the next parameter is searched. Results are
cached in a JSON table keyed by the grid shape and x2 transform kind.
"""
import json
import os

import numpy as np

from . import Plan, EddyError, to_device

# yfft_pencils / ztrans_pencils 0: the plan's own choice (the smallest of 16, 32, 64 giving tiles
# of >= 4096 elements; 16 for the dense x2 transform)
DEFAULT = {"yfft_pencils": 0, "ztrans_pencils": 0, "xsolve_chunk": 12, "use_dense_z": 0, "fft_radix": 8,
           "yfft_gemm": 0}

SEARCH = (
    ("yfft_pencils", (4, 8, 16, 32, 64), ("yfft_fwd", "yfft_inv")),
    ("ztrans_pencils", (4, 8, 16, 32, 64), ("ztrans_fwd", "ztrans_inv")),
    ("xsolve_chunk", (4, 6, 8, 12, 16, 24, 32), ("xsolve",)),
    ("fft_radix", (8, 4, 2), ("yfft_fwd", "yfft_inv", "ztrans_fwd", "ztrans_inv")),
)

TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lib", "tuning.json")


def key(g):
    return "%dx%dx%d:%s" % (g.M, g.L, g.N, "dct" if g.z_uniform() else "dense")


def _time(g, tuning, rhs, reps):
    import torch
    plan = Plan(g, tuning=tuning, stream=torch.cuda.current_stream())  # ordered with the copies below
    effective = plan.tuning()
    d = rhs.clone()
    plan.direct(d)
    samples = {}
    for _ in range(reps):
        plan.enable_timing(True)
        plan.reset_timing()
        d.copy_(rhs)
        plan.direct(d)
        kt = plan.kernel_times()
        for k, (n, ms) in kt.items():
            samples.setdefault(k, []).append(ms / max(n, 1))
    plan.close()
    torch.cuda.synchronize()
    return {k: float(np.median(v)) for k, v in samples.items()}, effective


def autotune(g, reps=5, seed=0, log=None):
    """Return (best tuning dict, trace list) for grid g (needs the GPU)."""
    from . import configs
    us, vs, ws = configs.random_uniform_velocity(g, seed)
    rhs = to_device(configs.rhs_from_velocity(g, us, vs, ws))
    best = dict(DEFAULT)
    trace = []
    for param, values, kernels in SEARCH:
        scores = {}
        for v in values:
            t = dict(best)
            t[param] = v
            try:
                times, eff = _time(g, t, rhs, reps)
            except EddyError as e:  # candidate not valid for this grid (LDS, divisibility)
                trace.append({"param": param, "value": v, "error": str(e)})
                continue
            ev = eff.get(param, v)  # the plan clamps tile widths to what the kernels allow
            if ev in scores:
                trace.append({"param": param, "value": v, "same_as": ev})
                continue
            scores[ev] = sum(times[k] for k in kernels)
            trace.append({"param": param, "value": ev, "ms": round(scores[ev], 5)})
            if log:
                log("autotune %s=%s: %.4f ms" % (param, ev, scores[ev]))
        if scores:
            best[param] = min(scores, key=scores.get)
    return best, trace


# The batch search: one value of each parameter for every problem (0 = each plan's own choice).
BATCH_SEARCH = (
    ("ztrans_pencils", (0, 8, 16, 32, 64)),
    ("yfft_pencils", (0, 8, 16, 32, 64)),
    ("xsolve_chunk", (12, 4, 6, 8, 16)),
    ("fft_radix", (8, 4)),
)


def autotune_batch(grids, mode, reps=5, seed=0, log=None):
    """Return (tuning dict shared by the problems, trace list) for a batch of `grids` run in
    batch mode `mode` (needs the GPU).

    The objective is the batched FP64 solve time of the whole batch, not of each problem: the
    problems share the GPU, and the tiles that are fastest for a problem alone (autotune())
    are not the fastest among the others (the small problems' best tiles alone favour latency,
    in a batch throughput counts). The search is coordinate-wise as in autotune(), over one
    value per parameter for all problems; a value that some problem rejects is skipped."""
    import time
    import torch
    from . import Batch, EDDY_FP64
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    dps = [torch.rand((g.nz, g.ny, g.nx), generator=gen, device="cuda", dtype=torch.float64) for g in grids]
    stream = torch.cuda.current_stream()

    def timed(tuning):
        b = Batch(grids, tuning=tuning, stream=stream)
        try:
            b.set_mode(mode)
            b.bind(dps)
            for _ in range(2):
                b.direct(EDDY_FP64)
            samples = []
            for _ in range(reps):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                b.direct(EDDY_FP64)
                torch.cuda.synchronize()
                samples.append(time.perf_counter() - t0)
            return float(np.median(samples)) * 1e3
        finally:
            b.close()

    best = dict(DEFAULT)
    trace = []
    for param, values in BATCH_SEARCH:
        scores = {}
        for v in values:
            t = dict(best)
            t[param] = v
            try:
                scores[v] = timed(t)
            except EddyError as e:
                trace.append({"param": param, "value": v, "error": str(e)})
                continue
            trace.append({"param": param, "value": v, "ms": round(scores[v], 4)})
            if log:
                log("autotune batch %s=%s: %.4f ms" % (param, v, scores[v]))
        if scores:
            best[param] = min(scores, key=scores.get)
    return best, trace


def load_table(path=TABLE):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def tuned(g, path=TABLE, reps=5, log=None, retune=False):
    """Tuning for g from the table, autotuning (and recording) it when absent."""
    table = load_table(path)
    k = key(g)
    if k in table and not retune:
        return table[k]
    best, _ = autotune(g, reps=reps, log=log)
    table[k] = best
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    return best
