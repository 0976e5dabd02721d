"""Benchmark of the pressure projection on one MI355X.

Workload: the 512x256x256 (x1, x2, x3) staggered grid with x1 tanh-stretched (beta 1.5),
x2/x3 uniform, Lx = (4, 2, 2), Taylor-Green k=2 velocity + U(-1e-3, 1e-3) noise (seed 3).
A step is one projection: DIVERGENCE -> RHS/dt - volume-weighted mean -> Poisson solve ->
CORRECT, with every field resident in HBM. Two solve variants are timed:
  * plain FP64 (the drop-in DIRECT; matches the oracle to 1e-12),
  * FP32 transforms/eliminations + FP64 defect correction, with the fewest sweeps that reach
    ||b - L phi||_2 / ||b||_2 <= 1e-10 (checked on the device before timing).
`value` is the FP64 projection rate, each timed step one launch of the projection captured
as a hipGraph (the eager per-kernel launch rate is reported beside it); the refined rate,
solves/s and cells/s are reported beside it too. `roofline` covers the kernel with the largest time per step (HIP events on the
plan's stream); `cpu_baseline` times the oracle (C, OpenMP) on the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "eddy-hybrid_amd"))

METRIC = "Poisson projections/s at 512x256x256: FP64 vs FP32+FP64-refine at 1e-10 resid"
HBM_PEAK_GBS = 8000.0

# algorithmic HBM bytes per cell of each kernel (FP64 field = 8 B/cell)
BYTES_PER_CELL = {
    # the projection fuses the divergence into the forward y-FFT: reads u, v, w, writes S
    "yfft_fwd": 32.0, "ztrans_fwd": 16.0, "xsolve": 16.0, "ztrans_inv": 16.0, "yfft_inv": 16.0,
    "divergence": 32.0, "correct": 56.0,
}


# kernel names in the committed HBM-traffic profile (rocprofv3 PMC, calibrated bytes/launch)
PROFILE_NAMES = {"yfft_fwd": "k_yfft_fwd<double, 3,", "ztrans_fwd": "k_zdct_fwd<double>", "xsolve": "k_xsolve<double,",
                 "ztrans_inv": "k_zdct_inv<double>", "yfft_inv": "k_yfft_inv<double, 0,", "correct": "k_correct"}
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_taylor_green_512x256x256.md")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed PMC summary (None if absent)."""
    name = PROFILE_NAMES.get(kernel)
    if not name or not os.path.exists(TRAFFIC_PROFILE):
        return None
    for line in open(TRAFFIC_PROFILE):
        cols = [c.strip() for c in line.strip().strip("|").split("|")]
        kname = cols[0].strip("`")
        if kname.startswith("void "):  # template kernels carry their return type
            kname = kname[5:]
        if len(cols) == 5 and kname.startswith(name):
            try:
                return float(cols[4])
            except ValueError:
                return None
    return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def bench_batch(torch, eddyproj, configs, args, log):
    """48 independent problems of uneven sizes (seeded sizes, x1 stretching beta ~ U(1, 2.5)),
    velocities U(-1, 1); timed in both batch modes: per-problem kernels on parallel graph
    branches, and one launch per stage with a workgroup per work item of all problems; the tiles
    of each mode autotuned on the whole batch (autotune.autotune_batch)."""
    from eddyproj import grid
    sizes = configs.batch48_sizes()
    grids = [grid.make_grid(m, n, l, x_beta=b) for (m, n, l, b) in sizes]
    gen = torch.Generator(device="cuda")
    gen.manual_seed(4)
    vel = []
    for g in grids:
        shp = (g.nz, g.ny, g.nx)
        vel.append([torch.rand(shp, generator=gen, device="cuda", dtype=torch.float64) * 2 - 1 for _ in range(3)])
    outs = [[t.clone() for t in v] for v in vel]
    dps = [torch.zeros_like(v[0]) for v in vel]
    b = eddyproj.Batch(grids, stream=torch.cuda.current_stream())
    b.bind(dps, [v[0] for v in vel], [v[1] for v in vel], [v[2] for v in vel], [o[0] for o in outs],
           [o[1] for o in outs], [o[2] for o in outs])
    cells, items = b.info()
    log("batch48: %d problems, %d cells, %d items in the largest stage" % (len(grids), cells, items))
    steps = max(3, min(20, args.steps // 3))

    def timed(fn):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    out = {"problems": len(grids), "cells": cells, "largest_stage_items": items, "steps": steps}
    # the RHS of the solve timings: each problem's divergence of its U(-1, 1) velocities with the
    # volume-weighted mean removed (the batch's own projection forms it: its DP after one step
    # holds the solution, so the RHS is rebuilt here on the host from the same velocities)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle
    oracle.set_threads(min(16, os.cpu_count() or 1))
    rhs_h = []
    for g, v in zip(grids, vel):
        us, vs, ws = (eddyproj.to_host(t) for t in v)
        rhs_h.append(configs.rhs_from_velocity(g, us, vs, ws))
    rhs = [eddyproj.to_device(r) for r in rhs_h]
    work = [torch.zeros_like(d) for d in dps]

    def residuals():
        """max over the problems of ||b - L phi||_2 / ||b||_2 (the oracle's operator, host)."""
        worst = 0.0
        for g, r, w in zip(grids, rhs_h, work):
            phi = eddyproj.to_host(w)
            bb = r[1:-1, 1:-1, 1:-1]
            res = oracle.laplacian(g, phi)[1:-1, 1:-1, 1:-1] - bb
            worst = max(worst, float(np.linalg.norm(res) / np.linalg.norm(bb)))
        return worst

    b.close()
    from eddyproj import autotune
    for mode, name in ((eddyproj.BATCH_BRANCHES, "branches"), (eddyproj.BATCH_QUEUES, "queues")):
        # tiles tuned for this mode on the whole batch's FP64 solve time (the first candidate
        # timed is every plan's default); a profiling pass keeps the defaults
        tuning, trace = ((dict(autotune.DEFAULT), []) if args.profile else autotune.autotune_batch(grids, mode))
        b = eddyproj.Batch(grids, tuning=tuning, stream=torch.cuda.current_stream())
        b.set_mode(mode)
        b.bind(dps, [v[0] for v in vel], [v[1] for v in vel], [v[2] for v in vel], [o[0] for o in outs],
               [o[1] for o in outs], [o[2] for o in outs])
        t_proj = timed(lambda: b.project(configs.DT))
        b.bind(work)
        t64 = timed(lambda: b.direct(eddyproj.EDDY_FP64))
        t32 = timed(lambda: b.direct(eddyproj.EDDY_FP32_REFINE, 2))
        # correctness of both precisions on the workload's own RHS
        checks = {}
        for prec, sw, key in ((eddyproj.EDDY_FP64, 0, "fp64"), (eddyproj.EDDY_FP32_REFINE, 2, "fp32_refine2")):
            for w, r in zip(work, rhs):
                w.copy_(r)
            b.direct(prec, sw)
            torch.cuda.synchronize()
            checks[key + "_max_rel_residual"] = residuals()
        out[name] = {"fp64_projection_ms": round(t_proj * 1e3, 4), "fp64_solve_ms": round(t64 * 1e3, 4),
                     "fp32_refine2_solve_ms": round(t32 * 1e3, 4), "fp64_batch_solves_per_s": round(1.0 / t64, 2),
                     "fp64_cells_per_s": round(cells / t64, 1),
                     "fp64_projection_batches_per_s": round(1.0 / t_proj, 2), **checks,
                     "tuning": {k: tuning[k] for k in ("yfft_pencils", "ztrans_pencils", "xsolve_chunk", "fft_radix")},
                     "fp64_solve_ms_default_tiles": trace[0].get("ms") if trace else None}
        log("batch48 %s: projection %.3f ms, FP64 solve %.3f ms (default tiles %s), FP32+2 %.3f ms, tuning %s, "
            "residuals %s" % (name, t_proj * 1e3, t64 * 1e3, trace[0].get("ms") if trace else None, t32 * 1e3,
                              out[name]["tuning"], checks))
        b.close()
    return out


def minimal_pass(cells, solve_ms, proj_ms):
    """Fractions of the minimal-pass HBM roofline: a solve moves 48 B/cell (three sweeps of
    8 B read + 8 B written), a projection 136 B/cell (divergence 32 + solve 48 + CORRECT 56)."""
    out = {}
    for key, bpc, ms in (("solve", 48.0, solve_ms), ("projection", 136.0, proj_ms)):
        if ms:
            gbs = bpc * cells / (ms * 1e-3) / 1e9
            out[key] = {"bytes_per_cell": bpc, "ms": round(ms, 4), "GBs": round(gbs, 1),
                        "frac_of_8TBs": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def refine_error(eddyproj, plan, rhs_dev, exact, tol=1e-10, max_sweeps=5):
    """FP32 + FP64 defect correction with the fewest sweeps reaching `tol`: sweeps, residual and
    the error against the long-double oracle's answer `exact` after the plain-mean removal, in
    the max and L2 norms (relative)."""
    import numpy as np
    e = exact[1:-1, 1:-1, 1:-1]
    e = e - e.mean()
    for s in range(max_sweeps + 1):
        t = rhs_dev.clone()
        r = plan.direct_refined(t, sweeps=s, residual=True)
        if r <= tol:
            got = eddyproj.to_host(t)[1:-1, 1:-1, 1:-1]
            got = got - got.mean()
            return {"sweeps": s, "residual": r, "err_max": float(np.abs(got - e).max() / np.abs(e).max()),
                    "err_l2": float(np.linalg.norm(got - e) / np.linalg.norm(e))}
    return {"sweeps": None, "residual": r}


def bench_config(torch, eddyproj, configs, name, build, reps, log, with_oracle=True):
    """One workload grid besides the headline: FP64 solve and projection rates (the solve as
    eager launches timed with HIP events, the projection as hipGraph launches), their
    minimal-pass roofline fractions, the FP64 error against the long-double oracle and the
    FP32 + FP64-refine error at the fewest sweeps reaching 1e-10."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    g, us, vs, ws = build()
    cells = g.cells
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    plan = eddyproj.Plan(g, stream=stream)
    rhs_h = configs.rhs_from_velocity(g, us, vs, ws)
    rhs = eddyproj.to_device(rhs_h)
    dus, dvs, dws = (eddyproj.to_device(a) for a in (us, vs, ws))
    uo, vo, wo = dus.clone(), dvs.clone(), dws.clone()
    dp = torch.zeros_like(dus)
    d = rhs.clone()
    for _ in range(5):
        d.copy_(rhs)
        plan.direct(d)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        plan.direct(d)  # the solve's cost does not depend on the values it runs on
    e1.record(stream)
    e1.synchronize()
    t_solve = e0.elapsed_time(e1) / reps
    gr = eddyproj.Graph(plan, dus, dvs, dws, uo, vo, wo, dp, configs.DT)
    for _ in range(5):
        gr.launch()
    e0.record(stream)
    for _ in range(reps):
        gr.launch()
    e1.record(stream)
    e1.synchronize()
    t_proj = e0.elapsed_time(e1) / reps
    gr.close()
    out = {"cells": cells, "grid_x1_x2_x3": [g.M, g.N, g.L], "reps": reps,
           "fp64_solve_ms": round(t_solve, 4), "fp64_solves_per_s": round(1e3 / t_solve, 1),
           "fp64_projection_ms": round(t_proj, 4), "fp64_projections_per_s": round(1e3 / t_proj, 1),
           "minimal_pass_roofline": minimal_pass(cells, t_solve, t_proj), "tuning": plan.tuning()}
    if with_oracle:
        oracle.set_threads(min(16, os.cpu_count() or 1))
        exact = oracle.direct(g, rhs_h, extended=True)
        d.copy_(rhs)
        plan.direct(d)
        got = eddyproj.to_host(d)[1:-1, 1:-1, 1:-1]
        e = exact[1:-1, 1:-1, 1:-1]
        got, e = got - got.mean(), e - e.mean()
        out["fp64_err_max"] = float(np.abs(got - e).max() / np.abs(e).max())
        out["fp32_refine"] = refine_error(eddyproj, plan, rhs, exact)
    log("%s: solve %.4f ms, projection %.4f ms, %s" % (name, t_solve, t_proj,
                                                        {k: v for k, v in out.items() if "err" in k or k == "fp32_refine"}))
    plan.close()
    return out


def bench_dft_gemm(torch, eddyproj, configs, plan, g, tuning, stream, dus, dvs, dws, uo, vo, wo, dp, rhs, sweeps,
                   args, log):
    """FP32+refine projections with the periodic-axis transforms as radix FFTs (the plan) and
    as dense FP32 MFMA products (a second plan with yfft_gemm): step time, residual, and the
    average FP32 y-transform kernel times."""
    gplan = eddyproj.Plan(g, tuning=dict(tuning, yfft_gemm=1), stream=stream)
    out = {"length": g.L, "sweeps": sweeps}
    for name, p in (("fft", plan), ("dft_gemm", gplan)):
        t = rhs.clone()
        res = p.direct_refined(t, sweeps=sweeps, residual=True)
        for _ in range(args.warmup):
            p.project(dus, dvs, dws, uo, vo, wo, dp, configs.DT, precision=eddyproj.EDDY_FP32_REFINE, sweeps=sweeps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            p.project(dus, dvs, dws, uo, vo, wo, dp, configs.DT, precision=eddyproj.EDDY_FP32_REFINE, sweeps=sweeps)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / args.steps * 1e3
        p.enable_timing(True)
        p.reset_timing()
        for _ in range(3):
            p.project(dus, dvs, dws, uo, vo, wo, dp, configs.DT, precision=eddyproj.EDDY_FP32_REFINE, sweeps=sweeps)
        kt = p.kernel_times()
        p.enable_timing(False)
        yf = kt["yfft_fwd"][1] / max(1, kt["yfft_fwd"][0])
        yi = kt["yfft_inv"][1] / max(1, kt["yfft_inv"][0])
        out[name] = {"ms_per_projection": round(ms, 4), "residual": res, "y_fwd_avg_ms": round(yf, 5),
                     "y_inv_avg_ms": round(yi, 5)}
        log("fp32 y transform %s: %.3f ms/projection, fwd %.4f ms inv %.4f ms, residual %.2e"
            % (name, ms, yf, yi, res))
    # achieved FP32 matrix-core rate of the forward product: 2 L flops per cell
    out["dft_gemm"]["fwd_tflops"] = round(2.0 * g.L * g.cells / (out["dft_gemm"]["y_fwd_avg_ms"] * 1e-3) / 1e12, 2)
    out["dft_gemm"]["fwd_frac_of_fp32_peak"] = round(out["dft_gemm"]["fwd_tflops"] / 157.3, 4)
    gplan.close()
    return out


def bench_production(torch, eddyproj, configs, log, reps=20):
    """The reference's production grid (302x130x1026, cylindrical, stretched axial x2): the FP64
    projection with the stretched-x2 transform as batched FP64 products diagonalising x1 (per
    y-mode 300 x 300 bases, the default here), as products diagonalising x2 (one 1024 x 1024
    basis) and as the fused tile kernel, and the products' rate against the FP64 matrix-core
    peak; the products on the library's MFMA kernel (default) and, for comparison, on rocBLAS."""
    g, us, vs, ws = configs.production_302x130x1026()
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    out = {"workload": "production_302x130x1026", "grid_x1_x2_x3": [g.M, g.N, g.L], "cells": g.cells,
           "metrics": "cylindrical (x1 radius from the axis, run/x1_grid_half.in), stretched x2 "
                      "(run/x3_grid.in), 128 periodic azimuthal cells"}
    dus, dvs, dws = (eddyproj.to_device(a) for a in (us, vs, ws))
    uo, vo, wo = dus.clone(), dvs.clone(), dws.clone()
    dp = torch.zeros_like(dus)
    for name, tun in (("gemm_x1", {}), ("gemm_x2", {"zgemm": 1}),
                      ("gemm_x1_rocblas", {"zgemm": 2, "zgemm_lib": 1}), ("gemm_x2_rocblas", {"zgemm": 1, "zgemm_lib": 1}),
                      ("tile_kernel_x2", {"zgemm": -1, "ztrans_pencils": 8})):
        t0 = time.perf_counter()
        plan = eddyproj.Plan(g, tuning=tun, stream=stream)
        tplan = time.perf_counter() - t0
        for _ in range(3):
            plan.project(dus, dvs, dws, uo, vo, wo, dp, configs.DT, remove_mean=True)
        plan.enable_timing(True)
        plan.reset_timing()
        for _ in range(5):
            plan.project(dus, dvs, dws, uo, vo, wo, dp, configs.DT, remove_mean=True)
        kt = plan.kernel_times()
        plan.enable_timing(False)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        stream.synchronize()
        e0.record(stream)
        for _ in range(reps):
            plan.project(dus, dvs, dws, uo, vo, wo, dp, configs.DT, remove_mean=True)
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        per = {k: round(v[1] / max(v[0], 1), 4) for k, v in kt.items() if v[0]}
        r = {"ms_per_projection": round(ms, 3), "projections_per_s": round(1e3 / ms, 2),
             "cells_per_s": round(g.cells / ms * 1e3, 1), "plan_create_s": round(tplan, 2), "kernels_ms": per}
        r["zgemm"] = plan.tuning()["zgemm"]
        if name.startswith("gemm"):
            # one batched product: L modes of (M x N)(N x N), or of (N x M)(M x M)
            flops = 2.0 * g.M * g.N * (g.N if name.startswith("gemm_x2") else g.M) * g.L
            for k in ("ztrans_fwd", "ztrans_inv"):
                if k in per:
                    tf = flops / (per[k] * 1e-3) / 1e12
                    r[k + "_tflops"] = round(tf, 2)
                    r[k + "_frac_of_fp64_peak"] = round(tf / 78.6, 3)
        out[name] = r
        log("production %s: %.3f ms/projection %s" % (name, ms, per))
        plan.close()
    return out


def bench_resident(torch, eddyproj, configs, args, log, tuning):
    """The resident loop at 768x384x256 (x1 tanh beta 1.5; streamwise = x3, the periodic axis):
    `resident_steps` replays of the captured step (Philox noise, sigma 1e-2, key 5 -> FP64 or
    FP32+refine projection -> divergence min/max logged on the device), timed as a whole. The
    FP32 variant runs with the length-256 periodic transform as radix FFT and as dense DFT on
    the FP32 matrix cores."""
    import numpy as np
    g, u, v, w = configs.wake_768x384x256()
    stream = torch.cuda.current_stream()
    plan = eddyproj.Plan(g, tuning=tuning, stream=stream)
    out = {"workload": "wake_768x384x256", "cells": g.cells, "steps": args.resident_steps,
           "noise": "Philox4x32-10 N(0,1) x 1e-2 per face per step, key 5, counter (cell, step)"}
    # divergence the noise injects into the undisturbed wake (scale of the drift figures)
    du, dv, dw = (eddyproj.to_device(a) for a in (u, v, w))
    star = [torch.zeros_like(du) for _ in range(3)]
    step = torch.zeros(1, dtype=torch.int32, device="cuda")
    plan.perturb(du, dv, dw, *star, configs.WAKE_SIGMA, configs.WAKE_SEED, step)
    lo, hi = plan.div_minmax(*star)
    scale = max(abs(lo), abs(hi))
    del star
    out["injected_divergence_max"] = scale
    gplan = eddyproj.Plan(g, tuning=dict(tuning, yfft_gemm=1), stream=stream)
    for name, prec, sw, pl in (("fp64", eddyproj.EDDY_FP64, 0, plan),
                               ("fp32_refine2", eddyproj.EDDY_FP32_REFINE, 2, plan),
                               ("fp32_refine2_dft_gemm", eddyproj.EDDY_FP32_REFINE, 2, gplan)):
        fu, fv, fw = (eddyproj.to_device(a) for a in (u, v, w))
        dp = torch.zeros_like(fu)
        n = args.resident_steps
        res = eddyproj.Resident(pl, fu, fv, fw, dp, configs.DT, configs.WAKE_SIGMA, configs.WAKE_SEED,
                                precision=prec, sweeps=sw, max_steps=n + 10)
        res.run(10)  # warm-up replays (steps 0..9)
        torch.cuda.synchronize()
        log("resident %s: %d graph replays" % (name, n))
        t0 = time.perf_counter()
        res.run(n)
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / n
        done, dl = res.log()
        drift = np.abs(dl).max(axis=1) / scale
        half = max(1, len(drift) // 2)
        out[name] = {"ms_per_step": round(t * 1e3, 4), "steps_per_s": round(1.0 / t, 2), "steps_done": done,
                     "sweeps": sw,
                     "rel_divergence_first": float(drift[0]), "rel_divergence_last": float(drift[-1]),
                     "rel_divergence_max": float(drift.max()),
                     "rel_divergence_mean_first_half": float(drift[:half].mean()),
                     "rel_divergence_mean_second_half": float(drift[half:].mean())}
        log("resident %s: %.3f ms/step, relative divergence max %.2e" % (name, t * 1e3, drift.max()))
        res.close()
        del fu, fv, fw, dp
    gplan.close()
    plan.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-batch", action="store_true", help="skip the 48-problem batch measurement")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the 64x64x32 and 256x192x128 workloads (rates, rooflines, errors)")
    ap.add_argument("--no-production", action="store_true", help="skip the production-grid measurement")
    ap.add_argument("--only-production", action="store_true",
                    help="run only the production-grid measurement (for kernel-trace profiles of it)")
    ap.add_argument("--only-batch", action="store_true",
                    help="run only the 48-problem batch measurement (with --profile: the plans' default tiles, "
                         "for kernel-trace profiles of it)")
    ap.add_argument("--only-dft-gemm", action="store_true",
                    help="run only the FP32 radix-FFT vs DFT-as-GEMM projections (for counter profiles of "
                         "the matrix-core transform)")
    ap.add_argument("--resident-steps", type=int, default=1000,
                    help="graph replays of the 768x384x256 resident loop (0 skips it)")
    ap.add_argument("--profile", action="store_true",
                    help="profiling pass: FP64 and FP32+refine projections only (no CPU leg, no sweep search), "
                         "plus one device copy of exactly 1 GiB that calibrates HBM byte counters")
    args = ap.parse_args()
    if args.profile or args.only_dft_gemm:
        args.no_cpu = True
        args.no_production = True
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the projection of one grid runs on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured path)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import eddyproj
    from eddyproj import configs

    torch.cuda.set_device(0)
    if args.only_production:
        print(json.dumps({"production_302x130x1026": bench_production(torch, eddyproj, configs, log, reps=5)}))
        return
    if args.only_batch:
        print(json.dumps({"batch48": bench_batch(torch, eddyproj, configs, args, log)}))
        return
    log("building workload taylor_green_512x256x256")
    g, us, vs, ws = configs.taylor_green_512x256x256()
    cells = g.cells
    stream = torch.cuda.Stream()
    # every torch operation below runs on the plan's stream, so library kernels and torch
    # copies are ordered without extra synchronisation
    torch.cuda.set_stream(stream)
    from eddyproj import autotune
    tuning, trace = ((autotune.DEFAULT, []) if args.profile or args.only_dft_gemm
                     else autotune.autotune(g, reps=5, log=log))
    log("tuning %s" % tuning)
    plan = eddyproj.Plan(g, tuning=tuning, stream=stream)
    dus, dvs, dws = (eddyproj.to_device(a) for a in (us, vs, ws))
    uo, vo, wo = dus.clone(), dvs.clone(), dws.clone()
    dp = torch.zeros_like(dus)
    torch.cuda.synchronize()

    if args.profile:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.neg(src, out=dst)  # a kernel reading 1 GiB and writing 1 GiB: counter calibration
        torch.cuda.synchronize()
        del src, dst
        run_steps = max(1, args.steps)
        for _ in range(run_steps):
            plan.project(dus, dvs, dws, uo, vo, wo, dp, configs.DT, remove_mean=True)
        for _ in range(run_steps):
            plan.project(dus, dvs, dws, uo, vo, wo, dp, configs.DT, remove_mean=True,
                         precision=eddyproj.EDDY_FP32_REFINE, sweeps=2)
        torch.cuda.synchronize()
        print(json.dumps({"profile_pass": True, "steps": run_steps}), flush=True)
        return

    if args.only_dft_gemm:
        rhs = eddyproj.to_device(configs.rhs_from_velocity(g, us, vs, ws))
        out = bench_dft_gemm(torch, eddyproj, configs, plan, g, tuning, stream, dus, dvs, dws, uo, vo, wo, dp, rhs, 2,
                             args, log)
        print(json.dumps({"fp32_y_transform_fft_vs_dft_gemm": out}), flush=True)
        return

    # fewest refinement sweeps reaching 1e-10 on this RHS
    rhs = eddyproj.to_device(configs.rhs_from_velocity(g, us, vs, ws))
    sweeps, resid = None, None
    for s in range(0, 6):
        t = rhs.clone()
        r = plan.direct_refined(t, sweeps=s, residual=True)
        log("refine sweeps=%d residual=%.3e" % (s, r))
        if r <= 1e-10:
            sweeps, resid = s, r
            break
    del t
    fp64_check = rhs.clone()
    plan.direct(fp64_check)
    lap = torch.zeros_like(rhs)
    plan.laplacian(fp64_check, lap)
    torch.cuda.synchronize()
    b = rhs[1:-1, 1:-1, 1:-1]
    fp64_resid = float(torch.linalg.vector_norm(lap[1:-1, 1:-1, 1:-1] - b) / torch.linalg.vector_norm(b))
    log("fp64 residual %.3e" % fp64_resid)
    del lap, fp64_check

    def run(n, precision, sw):
        for _ in range(n):
            plan.project(dus, dvs, dws, uo, vo, wo, dp, configs.DT, remove_mean=True, precision=precision,
                         sweeps=sw)

    def timed(step):
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / args.steps

    def timed_graph(precision, sw):
        # the projection captured once as a hipGraph (every kernel of the step re-runs on
        # every replay), then launched once per step; before the W warm-up steps the GPU is kept
        # busy with replays for 0.3 s, so that the timed steps do not start from the idle clocks
        # the host-side set-up above left
        gr = eddyproj.Graph(plan, dus, dvs, dws, uo, vo, wo, dp, configs.DT, remove_mean=True, precision=precision,
                            sweeps=sw)
        try:
            t_end = time.perf_counter() + 0.3
            while time.perf_counter() < t_end:
                for _ in range(20):
                    gr.launch()
                torch.cuda.synchronize()
            return timed(gr.launch)
        finally:
            gr.close()

    log("timing fp64 projection, %d steps (graph launches, then eager launches)" % args.steps)
    t64 = timed_graph(eddyproj.EDDY_FP64, 0)
    t64_eager = timed(lambda: run(1, eddyproj.EDDY_FP64, 0))
    t32 = t32_eager = None
    if sweeps is not None:
        log("timing fp32+refine projection (%d sweeps), %d steps" % (sweeps, args.steps))
        t32 = timed_graph(eddyproj.EDDY_FP32_REFINE, sweeps)
        t32_eager = timed(lambda: run(1, eddyproj.EDDY_FP32_REFINE, sweeps))

    # the length-256 periodic transform of the FP32 solve: radix FFT vs dense DFT on the FP32
    # matrix cores (same sweeps; the kernel times are the FP32 y transforms of one projection)
    gemm = None
    if sweeps is not None:
        try:
            gemm = bench_dft_gemm(torch, eddyproj, configs, plan, g, tuning, stream, dus, dvs, dws, uo, vo, wo, dp,
                                  rhs, sweeps, args, log)
        except Exception as e:  # reported, never hides the main measurement
            gemm = {"error": repr(e)}

    # per-kernel durations (HIP events on the plan's stream), same workload
    plan.enable_timing(True)
    plan.reset_timing()
    run(args.steps, eddyproj.EDDY_FP64, 0)
    kt = plan.kernel_times()
    plan.enable_timing(False)
    per = {k: (v[1] / v[0] if v[0] else 0.0, v[0] / args.steps) for k, v in kt.items()}
    dom = max((k for k in BYTES_PER_CELL), key=lambda k: per[k][0] * per[k][1])
    avg_ms = per[dom][0]
    achieved = BYTES_PER_CELL[dom] * cells / (avg_ms * 1e-3) / 1e9
    kernels = {k: {"avg_ms": round(per[k][0], 5), "launches_per_step": per[k][1],
                   "GBs": round(BYTES_PER_CELL[k] * cells / (per[k][0] * 1e-3) / 1e9, 1) if per[k][0] else None}
               for k in BYTES_PER_CELL}
    # the solve kernels inside the projection (the forward y transform carries the divergence)
    solve_in_proj_ms = sum(per[k][0] * per[k][1] for k in ("yfft_fwd", "ztrans_fwd", "xsolve", "ztrans_inv", "yfft_inv"))
    # DIRECT alone on a stored RHS (the solves/s figure): HIP events around 100 in-place solves
    # on the plan's stream (each solves the previous answer: same work, bounded values)
    dsol = rhs.clone()
    for _ in range(10):
        plan.direct(dsol)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(100):
        plan.direct(dsol)
    ev1.record(stream)
    ev1.synchronize()
    solve_ms = ev0.elapsed_time(ev1) / 100
    del dsol

    # the FP32+refine answer against the long-double oracle at full size (host, OpenMP)
    refine_vs_oracle = None
    if not args.no_cpu:
        try:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle
            oracle.set_threads(min(16, os.cpu_count() or 1))
            log("long-double oracle DIRECT of the headline RHS")
            exact = oracle.direct(g, configs.rhs_from_velocity(g, us, vs, ws), extended=True)
            refine_vs_oracle = refine_error(eddyproj, plan, rhs, exact)
            d = rhs.clone()
            plan.direct(d)
            got = eddyproj.to_host(d)[1:-1, 1:-1, 1:-1]
            e = exact[1:-1, 1:-1, 1:-1]
            got, e = got - got.mean(), e - e.mean()
            refine_vs_oracle["fp64_err_max"] = float(np.abs(got - e).max() / np.abs(e).max())
            del exact, got, e, d
            log("refine vs oracle: %s" % refine_vs_oracle)
        except Exception as e:  # reported, never hides the main measurement
            refine_vs_oracle = {"error": repr(e)}

    other_configs = {}
    if not args.no_configs:
        for name, build, reps in (("random_64x64x32", configs.random_64x64x32, 400),
                                  ("dipole_256x192x128", configs.dipole_256x192x128, 100)):
            try:
                other_configs[name] = bench_config(torch, eddyproj, configs, name, build, reps, log,
                                                   with_oracle=not args.no_cpu)
            except Exception as e:  # reported, never hides the main measurement
                other_configs[name] = {"error": repr(e)}
        torch.cuda.set_stream(stream)

    batch = None
    if not args.no_batch:
        try:
            batch = bench_batch(torch, eddyproj, configs, args, log)
        except Exception as e:  # reported, never hides the main measurement
            batch = {"error": repr(e)}

    # per-step distribution over captured-graph replays of the same projection
    graph = eddyproj.Graph(plan, dus, dvs, dws, uo, vo, wo, dp, configs.DT)
    for _ in range(10):
        graph.launch()
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(101)]
    evs[0].record(stream)
    for i in range(100):
        graph.launch()
        evs[i + 1].record(stream)
    torch.cuda.synchronize()
    reps = np.array([evs[i].elapsed_time(evs[i + 1]) for i in range(100)])
    graph.close()
    graph_stats = {"replays": 100, "median_ms": round(float(np.median(reps)), 4),
                   "p10_ms": round(float(np.percentile(reps, 10)), 4),
                   "p90_ms": round(float(np.percentile(reps, 90)), 4)}
    log("graph replays: %s" % graph_stats)

    # the drop-in boundary: DIRECT with the Fortran signature on host arrays (PCIe both ways)
    host_entry = None
    try:
        rhs_h = configs.rhs_from_velocity(g, us, vs, ws)
        eddyproj.direct(rhs_h.copy(order="F"), g)  # builds and caches the plan and the staging
        reps = 5
        arrays = [rhs_h.copy(order="F") for _ in range(reps)]  # DIRECT overwrites its argument
        t0 = time.perf_counter()
        for a in arrays:
            eddyproj.direct(a, g)
        th = (time.perf_counter() - t0) / reps
        del arrays
        eddyproj.release_direct_cache()
        host_entry = {"ms_per_direct": round(th * 1e3, 3), "solves_per_s": round(1.0 / th, 2),
                      "note": "host DP in and out through eddy_direct (pageable numpy arrays through pinned "
                              "staging, RHS check included), the PCIe-inclusive rate of the drop-in boundary"}
        log("host entry DIRECT: %.2f ms" % (th * 1e3))
    except Exception as e:
        host_entry = {"error": repr(e)}

    # pressure update after CORRECT (P = P + DP - PREF, ghosts; implicit y correction on) and
    # the field files of the restart set between HBM and a file, at the same size
    tail = {}
    try:
        p = torch.zeros_like(dp)
        mask = np.ones(g.nx)
        plan.set_pressure_bc(imply=1, ru1=1e-3, rvimply=mask)
        for _ in range(3):
            plan.pressure_update(p, dp, configs.DT)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(20):
            plan.pressure_update(p, dp, configs.DT)
        e1.record(stream)
        e1.synchronize()
        tpu = e0.elapsed_time(e1) / 20
        b = 3 * 8 * cells  # read P and DP, write P (the y neighbours of DP hit in cache)
        tail["pressure_update"] = {"ms": round(tpu, 4), "algorithmic_bytes": b, "GBs": round(b / tpu / 1e6, 1),
                                   "note": "P += DP - 0.5 ALFXDT RU1 d2(DP)/dy2 (IMPLY), PREF, ghost shell"}
        log("pressure update: %.4f ms" % tpu)
        import tempfile
        from eddyproj import fieldio
        with tempfile.TemporaryDirectory(dir="/tmp") as td:
            f = os.path.join(td, "p_00000001.res")
            nbytes = 8 * g.nx * g.ny * g.nz
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fieldio.write_scalar(f, dp, nstep=1, time=0.0, dtm1=configs.DT)
            tw = time.perf_counter() - t0
            t0 = time.perf_counter()
            q, _ = fieldio.read_scalar(f, device=True)
            tr = time.perf_counter() - t0
            same = bool(torch.equal(q, dp))
        tail["field_io"] = {"bytes": nbytes, "write_ms": round(tw * 1e3, 2), "write_GBs": round(nbytes / tw / 1e9, 2),
                            "read_ms": round(tr * 1e3, 2), "read_GBs": round(nbytes / tr / 1e9, 2),
                            "round_trip_exact": same,
                            "note": "IOSCALAR file (DIR=1) of the DP field: HBM -> pinned staging -> file, and "
                                    "back, through the host page cache of /tmp"}
        log("field io: write %.1f ms, read %.1f ms" % (tw * 1e3, tr * 1e3))
    except Exception as e:  # reported, never hides the main measurement
        tail["error"] = repr(e)

    production = None
    if not args.no_production:
        try:
            production = bench_production(torch, eddyproj, configs, log)
        except Exception as e:  # reported, never hides the main measurement
            production = {"error": repr(e)}
        torch.cuda.set_stream(stream)

    resident = None
    if args.resident_steps > 0:
        try:
            resident = bench_resident(torch, eddyproj, configs, args, log, tuning)
        except Exception as e:  # reported, never hides the main measurement
            resident = {"error": repr(e)}

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        oracle.set_threads(threads)
        log("cpu baseline: oracle projection on %d threads" % threads)
        reps = 0
        t0 = time.perf_counter()
        while True:
            div = oracle.divergence(g, us, vs, ws)
            b = div[1:-1, 1:-1, 1:-1] / configs.DT
            vol = g.cell_volume()
            div[1:-1, 1:-1, 1:-1] = b - (vol * b).sum() / vol.sum()
            phi = oracle.direct(g, div)
            oracle.correct(g, us, vs, ws, phi, configs.DT)
            reps += 1
            if time.perf_counter() - t0 > 10.0 or reps >= 5:
                break
        tc = (time.perf_counter() - t0) / reps
        # one projection on one core: the scalar rate of the same restatement
        oracle.set_threads(1)
        log("cpu baseline: one oracle projection on 1 thread")
        t1 = time.perf_counter()
        div = oracle.divergence(g, us, vs, ws)
        b = div[1:-1, 1:-1, 1:-1] / configs.DT
        div[1:-1, 1:-1, 1:-1] = b - (vol * b).sum() / vol.sum()
        oracle.correct(g, us, vs, ws, oracle.direct(g, div), configs.DT)
        t1 = time.perf_counter() - t1
        oracle.set_threads(threads)
        model = ""
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
        cpu = {"value": round(1.0 / tc, 4), "unit": "projections/s", "cores": threads, "kind": "port",
               "sample": "%d full FP64 projections (DIVERGENCE + DIRECT + CORRECT) of the same 512x256x256 "
                         "workload by the C oracle (OpenMP), %.2f s each" % (reps, tc),
               "single_core": {"value": round(1.0 / t1, 4), "unit": "projections/s", "cores": 1,
                               "sample": "1 projection of the same workload on one thread, %.2f s" % t1},
               "host": {"cpu_model": model, "nproc": os.cpu_count()}}

    out = {
        "metric": METRIC,
        "value": round(1.0 / t64, 3),
        "unit": "projections/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t64 * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: Taylor-Green k=2 velocity + U(-1e-3,1e-3) noise (seed 3) on the stretched grid",
        "config": {"workload": "taylor_green_512x256x256", "grid_x1_x2_x3": [512, 256, 256],
                   "x1_stretch": "tanh beta=1.5", "domain": [4.0, 2.0, 2.0], "bc": "walls x1/x2, periodic x3",
                   "step": "divergence + FP64 Poisson solve + correction"},
        "solves_per_s_fp64": round(1e3 / solve_ms, 2) if solve_ms else None,
        "fp64_solve_ms": round(solve_ms, 4),
        "fp64_solve_kernels_in_projection_ms": round(solve_in_proj_ms, 4),
        "minimal_pass_roofline": minimal_pass(cells, solve_ms, t64 * 1e3),
        "cells_per_s_fp64_projection": round(cells / t64, 1),
        "timing": "each step one launch of the projection captured as a hipGraph; eager per-kernel launches "
                  "of the same step: %.2f projections/s (FP64)" % (1.0 / t64_eager),
        "fp64_solve_residual": fp64_resid,
        "fp32_refine": None if t32 is None else {
            "projections_per_s": round(1.0 / t32, 3), "ms_per_step": round(t32 * 1e3, 4), "sweeps": sweeps,
            "eager_projections_per_s": round(1.0 / t32_eager, 3),
            "residual": resid, "speedup_vs_fp64": round(t64 / t32, 3), "vs_long_double_oracle": refine_vs_oracle},
        "other_configs": other_configs,
        "fp32_y_transform_fft_vs_dft_gemm": gemm,
        "kernels": kernels,
        "tuning": plan.tuning(),
        "autotune_trace": trace,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom),
                     "algorithmic_bytes_per_launch": BYTES_PER_CELL[dom] * cells},
        "cpu_baseline": cpu,
        "fp64_graph_replays": graph_stats,
        "host_entry_direct": host_entry,
        "batch48": batch,
        "pressure_update_and_field_io": tail,
        "production_302x130x1026": production,
        "resident_768x384x256": resident,
    }
    print(json.dumps(out), flush=True)
    plan.close()


if __name__ == "__main__":
    main()
