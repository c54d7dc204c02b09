"""Benchmark of the PDOT hot path on one MI355X.

Headline workload: the 10000 x 10000 Gaussian-shift transport problem
(source N(0,I2), target N((2,0),I2), squared Euclidean cost / max, U(0,1)
masses, seeded synthetic data), solved by restarted adaptive PDHG.

A step is one PDHG iteration (one take_step! pass over C, X and the running
average).  The timed region runs K steps of the FP64 form (the reference's
precision) on inputs resident in HBM, with HIP events recorded on the
library's stream around every launch of the streaming step kernel, so the
same run gives the iteration rate and the step kernel's average duration.
It starts --start-iteration (2048) iterations into the solve (untimed, with
the solver's own evaluations and restarts) and W untimed passes later: the
passes skip the iterates' entries recorded as +0.0 (zero maps), so their cost
depends on how many there are, and the zero start has only zeros.  Of the
points sampled along the 10k solve, iteration 2048 has the fewest.  The
roofline counts the bytes the passes must move there (zero shares measured
before and after the timed passes); the dense figure is reported beside it.

Also reported (extra fields): the same measurement for the FP32-storage and
point-cloud forms, full solves to 1e-4 relative KKT (time from after the
upload to the terminating evaluation, objectives against each other), the
CPU baseline (the oracle's literal restatement of the reference, a bounded
number of iterations on the host cores), and one entry per other workload,
each with the roofline of its dominant kernel:
  c1_cpu_1000        1000 x 1000 uniform points, FP64, solved to 1e-4 on the
                     host cores by the literal restatement (and on the GPU);
  c2_f32_1e-6        the headline instance in FP32 storage solved to 1e-6;
  c2_precision       FP32 storage against the FP64 oracle: the FP64 solve to
                     1e-8 and the FP32-storage solve to 1.5e-8 on the same
                     family at 1000 x 1000 (the 10k pair takes 15 minutes:
                     --workload precision), and the bound that storing C in
                     FP32 puts on the 10k optimum's change;
  c3_l1_mixture_30k  30000 x 30000 L1 mixture, dense FP32 C / FP64 duals, and
                     the FP64 form for the objective error, solved to 1e-4;
  c4_batched_512     512 instances of n, m in [64, 2048], FP32, to 1e-4, at
                     the reference's initial primal weight and the calibrated one;
  c5_warm_start_4096 held-out 4096 x 4096 Gaussian-mixture instances: zero
                     start against the learned starting states.
Together about four and a half minutes on one MI355X; --no-extra skips them.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-solve] [--no-cpu] [--no-extra]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "pdot.jl_amd"))

METRIC = "time-to-1e-4 rel KKT at 10k×10k, FP32/learned vs FP64; iters/s; HBM GB/s %"
PEAK_HBM_GBS = 8000.0


def algorithmic_bytes_per_step(n: int, m: int, elem: int, matrix_streams: int = 4) -> int:
    """Bytes one PDHG iteration moves on average, plus the row/column vectors
    (u, v, r, c and their averages read and written, a and b read: about 10
    doubles per row and per column).  A pass reads C and X and writes X+ (3
    n*m element streams); the running average is read and written (2 more)
    by every second pass only, which also applies the update the pass before
    it deferred: 4 streams per iteration on average (3 without a stored C,
    the point-cloud form).  With deferred averaging off, 5."""
    return matrix_streams * n * m * elem + 10 * 8 * (n + m)


def measured_traffic(n: int, m: int, form: str):
    """HBM bytes per step-kernel launch measured with rocprofv3 FETCH_SIZE /
    WRITE_SIZE (scaled by a copy kernel of known size) for this workload, as
    recorded in profiles/hbm_traffic_step_kernel_gaussian_shift_10k.md; None
    for other sizes."""
    path = os.path.join(REPO, "profiles", "hbm_traffic_step_kernel_gaussian_shift_10k.md")
    if (n, m) != (10000, 10000) or not os.path.exists(path):
        return None
    text = open(path).read()
    block = text.split("```json", 1)[1].split("```", 1)[0]
    return json.loads(block)[form]["traffic_bytes"]


def measured_batch_traffic():
    """HBM bytes per batched step-kernel launch (every instance active) of the
    512-instance workload, from profiles/hbm_traffic_bstep_kernel_batched_uniform_512.md."""
    path = os.path.join(REPO, "profiles", "hbm_traffic_bstep_kernel_batched_uniform_512.md")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    return json.loads(text.split("```json", 1)[1].split("```", 1)[0])["traffic_bytes"]


def measure_steps(dev, steps: int, warmup: int, prob=None, storage=None, start: int = 0):
    """Brings the device to `start` iterations into the solve of `prob` (the
    reference's defaults, tolerance 1e-4: its evaluations and restarts
    included, untimed), or to a fresh zero start; then `warmup` untimed and
    `steps` timed passes.  The zero shares of the current point and of the
    average's slot are read before and after the timed passes."""
    import pdot_amd as P
    from pdot_amd import _lib
    reached = 0
    if prob is not None and start > 0:
        params = P.default_parameters(1e-4, iteration_limit=start, verbosity=0, record_iteration_stats=False)
        reached = P.optimize(params, prob, storage=storage, solver=dev).iteration_count
    else:
        dev.set_start(policy=_lib.POLICY_ADAPTIVE, step_size=1.0, primal_weight=1.0)
    if warmup:
        dev.time_steps(warmup)

    def shares():
        zc = dev.zero_share(_lib.POINT_CURRENT)
        try:
            za = dev.zero_share(_lib.POINT_AVERAGE)
        except _lib.PdotError:  # empty average
            za = zc
        return zc, za

    z0 = shares()
    ks, tt = dev.time_steps(steps)
    z1 = shares()
    st = dev.get_state()
    return ks, tt, st, {"start_iteration": reached, "zero_share_current": 0.5 * (z0[0] + z1[0]),
                        "zero_share_average": 0.5 * (z0[1] + z1[1])}


def cpu_baseline(prob, iters: int):
    """The reference algorithm restated on host arrays (oracle literal mode:
    stores G and its average like the reference, 8 n*m FP64 arrays),
    OpenMP across the host cores, timed inside take_step! only."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    import pdot_amd as P
    params = P.default_parameters(1e-4, iteration_limit=iters + 1)
    t0 = time.time()
    r = O.optimize(params, prob, mode=O.LITERAL)
    wall = time.time() - t0
    st = r["stats"][-1]
    steps = int(r["take_step_calls"])
    basic = st["time_spent_doing_basic_algorithm"]
    return {"value": steps / basic, "unit": "iters/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"{steps} take_step! iterations of the literal FP64 restatement on the same "
                      f"{prob.num_row}x{prob.num_col} instance (evaluations excluded; wall {wall:.1f}s incl. "
                      f"{len(r['stats'])} evaluations)"}


def solve_to_tol(prob, storage, tol, dev=None, do_round=True, primal_importance=None):
    from dataclasses import replace
    import pdot_amd as P
    prof = P.SolveProfile()
    params = P.default_parameters(tol, verbosity=0, record_iteration_stats=False)
    if primal_importance is not None:
        params = replace(params, primal_importance=float(primal_importance))
    out = P.optimize(params, prob, storage=storage, solver=dev, profile=prof)
    ci = out.iteration_stats[-1].convergence_information[0]
    res = {"seconds": prof.solve_sec, "output_copy_sec": prof.output_sec, "iterations": out.iteration_count,
           "termination": out.termination_string, "objective": ci.primal_objective,
           "rel_kkt": P.relative_kkt_error(ci), "evaluations": prof.evaluations, "restarts": prof.restarts}
    if dev is not None and do_round:
        # output post-processing: feasibility rounding of the resident solution on the device
        # (four reads of X, one FP64 write) plus the device-to-host copy of the rounded plan
        t0 = time.time()
        Xr = dev.round(dev.output_point)
        res["rounding_sec_incl_copy"] = time.time() - t0
        if hasattr(prob, "cost_matrix"):
            res["rounded_objective"] = float((prob.cost_matrix * Xr).sum())
        del Xr
    return res


def roofline(achieved_gbs: float, bytes_per_launch: int, avg_launch_us: float, kernel: str, traffic=None) -> dict:
    return {"bound": "hbm", "achieved": achieved_gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": achieved_gbs / PEAK_HBM_GBS, "traffic": traffic, "kernel": kernel,
            "bytes_per_launch": bytes_per_launch, "avg_launch_us": avg_launch_us}


def sparse_bytes_per_step(launch: dict, n: int, m: int, elem: int, stored_c: bool, z_cur: float, z_avg: float) -> int:
    """Bytes one PDHG iteration moves on average with zero maps: C in full (the
    stored form); X read and X+ written only in the 16-byte units not recorded
    as +0.0 (share z_cur); the average, read and written by every second pass,
    only in its nonzero units (share z_avg); the maps themselves (one 64-bit
    word per wave and column: the current point's read and the new point's
    written every pass, the average's two every second pass); the row and
    column vectors.  With no zeros this is the dense figure plus the maps."""
    L = launch
    vec = 16 // elem
    waves = L["rows_per_block"] // (64 * vec)
    row_blocks = L["blocks"] // L["col_blocks"]
    map_bytes = 8 * row_blocks * waves * L["col_blocks"] * L["cols_per_block"]
    mats = n * m * elem * ((1 if stored_c else 0) + 2 * (1 - z_cur) + (1 - z_avg))
    return int(mats + 3 * map_bytes + 10 * 8 * (n + m))


def step_measurement(dev, n: int, m: int, elem: int, steps: int, warmup: int, matrix_streams: int = 4,
                     prob=None, storage=None, start: int = 0) -> dict:
    """K timed take_step! calls (after `start` solve iterations of `prob` and
    `warmup` untimed passes): iteration rate and the step kernel's HIP-event
    average against the bytes it had to move (zero maps: the units recorded
    as +0.0 are neither read nor written), and the dense figure beside it."""
    ks, tt, st, z = measure_steps(dev, steps, warmup, prob, storage, start)
    launch = dev.get_launch()
    alg = sparse_bytes_per_step(launch, n, m, elem, matrix_streams == 4, z["zero_share_current"], z["zero_share_average"])
    dense = algorithmic_bytes_per_step(n, m, elem, matrix_streams)
    k_avg = ks / steps * 1e-3
    return dict(z, launch=launch, ms_per_step=tt / steps, iters_per_s=steps / (tt * 1e-3), step_kernel_us=k_avg * 1e6,
                step_kernel_GBs=alg / k_avg / 1e9, bytes_per_step=alg, dense_bytes_per_step=dense,
                dense_equivalent_GBs=dense / k_avg / 1e9, numerical_error=int(st.numerical_error))


def l1_mixture(n: int, steps: int, warmup: int, solve: bool = True, point_cloud: bool = True, start: int = 0) -> dict:
    """n x n L1-cost mixture (c3 at n = 30000): dense FP32 storage (C, X and
    X_avg 3.6 GB each at 30000) with FP64 duals, and the FP64 form on the same
    instance for the objective error of the FP32-storage result."""
    from pdot_amd import _lib, instances
    m = n
    t0 = time.time()
    prob = instances.l1_mixture_3d(n)
    gen = time.time() - t0
    out = {"metric": f"{n}x{n} L1 mixture: FP32-storage vs FP64 solve to 1e-4", "unit": "s",
           "higher_is_better": False, "data": "synthetic", "n_gpus": 1,
           "config": {"workload": f"l1_mixture_3d {n}x{m} (source U[0,1]^3, 4-component mixture target, L1 cost, "
                                  f"lognormal masses, seed 3000)", "n": n, "m": m}, "generate_sec": gen}
    for storage, name, elem in ((_lib.STORAGE_F32, "f32", 4), (_lib.STORAGE_F64, "f64", 8)):
        dev = _lib.DeviceSolver(n, m, storage=storage)
        t0 = time.time()
        dev.upload(prob.cost_matrix, prob.source_distribution, prob.target_distribution)
        res = {"upload_sec": time.time() - t0, "device_GB": dev.device_bytes() / 1e9}
        res.update(step_measurement(dev, n, m, elem, steps, warmup, prob=prob, storage=storage, start=start))
        if solve:
            res["solve"] = solve_to_tol(prob, storage, 1e-4, dev=dev, do_round=False)
        out[name] = res
        dev.close()
    del prob
    if point_cloud:
        # point-cloud cost form, FP32 storage: C never stored (3 n*m streams per step on average)
        pts = instances.l1_mixture_3d_points(n)
        dev = _lib.DeviceSolver(n, m, storage=_lib.STORAGE_F32)
        t0 = time.time()
        dev.upload_points(pts)
        res = {"upload_sec": time.time() - t0, "device_GB": dev.device_bytes() / 1e9}
        res.update(step_measurement(dev, n, m, 4, steps, warmup, matrix_streams=3, prob=pts,
                                    storage=_lib.STORAGE_F32, start=start))
        if solve:
            res["solve"] = solve_to_tol(pts, _lib.STORAGE_F32, 1e-4, dev=dev, do_round=False)
        out["f32_point_cloud"] = res
        dev.close()
    f32 = out["f32"]
    out["roofline"] = roofline(f32["step_kernel_GBs"], f32["bytes_per_step"], f32["step_kernel_us"],
                               "step_kernel<float> (dense FP32 storage)")
    if solve:
        out["value"] = f32["solve"]["seconds"]
        out["speedup_f32_vs_f64"] = out["f64"]["solve"]["seconds"] / f32["solve"]["seconds"]
        out["objective_rel_err_f32_vs_f64"] = (abs(f32["solve"]["objective"] - out["f64"]["solve"]["objective"])
                                               / abs(out["f64"]["solve"]["objective"]))
    return out


def bench_l1_mixture(args) -> int:
    n = 30000 if args.n == 10000 else args.n
    print(json.dumps(l1_mixture(n, args.steps, args.warmup, solve=not args.no_solve, start=args.start_iteration)))
    return 0


def batched(steps: int, solve: bool = True) -> dict:
    """512 instances, n and m uniform in [64, 2048], FP32 storage, every
    instance solved to 1e-4 relative KKT in one batch, at the reference's
    initial primal weight (primal_importance 1.0) and at the calibrated one."""
    import numpy as np
    import pdot_amd as P
    from pdot_amd import _lib, instances
    from pdot_amd.batch import BatchProfile, autotune, optimize_batch
    t0 = time.time()
    probs = instances.batched_uniform(512)
    gen = time.time() - t0
    ns, ms = [p.num_row for p in probs], [p.num_col for p in probs]
    b = _lib.BatchSolver(ns, ms, storage=_lib.STORAGE_F32)
    t0 = time.time()
    for k, p in enumerate(probs):
        b.upload(k, p.cost_matrix, p.source_distribution, p.target_distribution)
    up = time.time() - t0
    # the pass timing below starts from zeros, which the zero maps would skip almost
    # entirely: time the dense passes (every element read and written) for the roofline
    b.set_sparse(False)
    tune = autotune(b, steps=4)
    b.set_start(_lib.POLICY_ADAPTIVE, 2.0, 1.0, 1.0, 1.0)
    b.time_steps(1)  # the first pass after a start has an empty average (untimed)
    # an even count: the passes that defer the averages' update alternate with those applying it
    steps = max(2, steps // 8 * 2)
    ks, tt = b.time_steps(steps)
    total_elems = sum(n * m for n, m in zip(ns, ms))
    alg = 4 * 4 * total_elems + 80 * (sum(ns) + sum(ms))  # 3 and 5 element streams alternately
    res = {"metric": "512-instance batch (n, m in [64, 2048]) solved to 1e-4, FP32 storage", "unit": "s",
           "higher_is_better": False, "data": "synthetic", "n_gpus": 1,
           "config": {"workload": "batched_uniform 512 x [64..2048]^2 (U[0,1]^2 points, squared Euclidean, "
                                  "U(0,1) masses, seed 4000)", "instances": 512, "elements": total_elems,
                      "storage": "f32"},
           "generate_sec": gen, "upload_sec": up, "autotune": tune["best"],
           "pass_ms_all_active": tt / steps, "step_kernel_GBs": alg / (ks / steps * 1e-3) / 1e9,
           "bytes_per_pass": alg}
    res["roofline"] = roofline(res["step_kernel_GBs"], alg, ks / steps * 1e3,
                               "bstep_kernel<float> (every instance active)", measured_batch_traffic())
    b.set_sparse(True)  # the solves run with zero maps, the default
    res["zero_maps"] = True
    if solve:
        from dataclasses import replace
        from pdot_amd import learned
        runs = [("reference_weight", 1.0)]
        w0 = learned.load_initial_primal_weight()
        if w0 is not None:
            runs.append(("calibrated_weight", w0))
        for name, w in runs:
            prof = BatchProfile()
            prm = replace(P.default_parameters(1e-4, record_iteration_stats=False), primal_importance=float(w))
            outs = optimize_batch(prm, probs, batch=b, use_graphs=True, profile=prof)
            iters = np.array([o.iteration_count for o in outs])
            res[name] = {"primal_importance": w, "solve_sec": prof.solve_sec,
                         "optimal": int(sum(o.termination_reason == P.TerminationReason.TERMINATION_REASON_OPTIMAL
                                            for o in outs)),
                         "instances_per_s": 512 / prof.solve_sec, "iterations_median": float(np.median(iters)),
                         "iterations_max": int(iters.max()), "passes_launched": prof.passes_launched,
                         "evaluations": prof.evaluations,
                         "objectives": [o.iteration_stats[-1].convergence_information[0].primal_objective
                                        for o in outs]}
        res["value"] = res["reference_weight"]["solve_sec"]
        if "calibrated_weight" in res:
            o1 = np.array(res["reference_weight"]["objectives"])
            o2 = np.array(res["calibrated_weight"]["objectives"])
            res["objective_rel_diff_max"] = float(np.max(np.abs(o2 - o1) / np.abs(o1)))
        for name, _ in runs:
            del res[name]["objectives"]
    b.close()
    return res


def bench_batched(args) -> int:
    print(json.dumps(batched(args.steps, solve=not args.no_solve)))
    return 0


def cpu_c1(tol: float = 1e-4, start: int = 0) -> dict:
    """Config c1: 1000 x 1000 uniform points, squared Euclidean, FP64, solved
    to 1e-4 by the reference algorithm restated on host arrays (the oracle's
    literal mode, OpenMP over the host cores), and by the GPU on the same
    instance for comparison."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O
    import pdot_amd as P
    from pdot_amd import _lib, instances
    prob = instances.uniform_sqeuclid(1000, seed=1000)
    params = P.default_parameters(tol, record_iteration_stats=False)
    t0 = time.time()
    r = O.optimize(params, prob, mode=O.LITERAL)
    wall = time.time() - t0
    st = r["stats"][-1]
    out = {"metric": "1000x1000 FP64 solve to 1e-4 on the host cores (reference algorithm, host arrays)",
           "unit": "s", "higher_is_better": False, "data": "synthetic",
           "config": {"workload": "uniform_sqeuclid 1000x1000 (U[0,1]^2 points, squared Euclidean, U(0,1) masses, "
                                  "seed 1000)", "n": 1000, "m": 1000, "storage": "f64"},
           "value": st["cumulative_time_sec"], "cores": O.num_threads(), "kind": "port",
           "iterations": r["iteration_count"], "termination_reason": r["termination_reason"],
           "iters_per_s": int(r["take_step_calls"]) / st["cumulative_time_sec"], "wall_sec": wall,
           "objective": st["primal_objective"],
           "note": "the 8 MB arrays are cache-resident on the host; time from after setup to the terminating "
                   "evaluation, as the reference's cumulative_time_sec"}
    dev = _lib.DeviceSolver(1000, 1000, storage=_lib.STORAGE_F64)
    dev.upload(prob.cost_matrix, prob.source_distribution, prob.target_distribution)
    g = step_measurement(dev, 1000, 1000, 8, 200, 20, prob=prob, storage=_lib.STORAGE_F64, start=start)
    g["solve"] = solve_to_tol(prob, _lib.STORAGE_F64, tol, dev=dev, do_round=False)
    dev.close()
    out["gpu_f64"] = g
    out["speedup_gpu_vs_cpu"] = out["value"] / g["solve"]["seconds"]
    out["roofline"] = roofline(g["step_kernel_GBs"], g["bytes_per_step"], g["step_kernel_us"],
                               "step_kernel<double> at 1000x1000 (24 MB per pass: cache-resident, launch-bound)")
    return out


def tight_solve(prob, storage, tol: float) -> dict:
    """One solve of the headline instance to a tighter tolerance."""
    from pdot_amd import _lib
    dev = _lib.DeviceSolver(prob.num_row, prob.num_col, storage=storage)
    dev.upload(prob.cost_matrix, prob.source_distribution, prob.target_distribution)
    res = solve_to_tol(prob, storage, tol, dev=dev, do_round=False)
    dev.close()
    return res


def warm_start_table(count: int, steps: int = 100, warmup: int = 10, limit_factor: int = 8, start: int = 0) -> dict:
    """Held-out 4096 x 4096 Gaussian-mixture instances (FP32 storage), time
    and iterations to 1e-4 from each starting state:
      default_weight      zero start, the reference's initial primal weight 1.0;
      calibrated_weight   zero start, the initial primal weight calibrated on
                          solves of the family (a constant);
      network_weight      zero start, the network's per-instance initial weight
                          (when the weights file carries that head);
      network_duals       the network's predicted duals as the starting iterate
                          (forced, no acceptance test) with the calibrated weight;
      exact_duals         the instance's own duals from the calibrated-weight
                          solve as the starting iterate, with the same weight:
                          what a perfect dual predictor would give (the time
                          of the solve that found them is not counted);
      learned_start       learned.solve_warm: everything the learned state offers,
                          behind its KKT acceptance test and fallback.
    Prediction time (and for learned_start the KKT evaluation of the predicted
    duals) is included in the learned rows; every row's solves share one
    uploaded problem, and the output copy is not counted.  A row's solve is capped
    at limit_factor times the zero start's iterations."""
    import numpy as np
    import pdot_amd as P
    from pdot_amd import _lib, instances, learned
    net = learned.WarmStartNet(learned.load_params())
    w_cal = learned.load_initial_primal_weight()
    has_head = learned.has_weight_head()
    rows = []
    for k in range(count):
        inst = instances.gmm_dirichlet(4096, seed=900200 + k)
        p = inst.problem
        n, m = p.cost_matrix.shape
        dev = _lib.DeviceSolver(n, m, storage=_lib.STORAGE_F32)
        dev.upload(p.cost_matrix, p.source_distribution, p.target_distribution)
        row = {"seed": 900200 + k}

        def run(name, weight, duals=None, extra_sec=0.0, cap=None, duals_out=None):
            from dataclasses import replace
            prm = replace(P.default_parameters(1e-4, record_iteration_stats=False), primal_importance=float(weight))
            if cap is not None:
                prm.termination_criteria.iteration_limit = int(cap)
            r = solve_to_tol_params(p, _lib.STORAGE_F32, prm, dev, warm_start=duals, duals_out=duals_out)
            r["seconds"] += extra_sec
            r["primal_importance"] = float(weight)
            row[name] = r
            return r

        z = run("default_weight", 1.0)
        cap = limit_factor * max(z["iterations"], 128)
        found = []
        if w_cal is not None:
            run("calibrated_weight", w_cal, cap=cap, duals_out=found)
        t0 = time.time()
        u0, v0 = net.predict_duals(inst.x, inst.y, p.source_distribution, p.target_distribution)
        t_pred = time.time() - t0
        if has_head:
            t0 = time.time()
            w_net = net.predict_weight(inst.x, inst.y, p.source_distribution, p.target_distribution)
            run("network_weight", w_net, extra_sec=time.time() - t0, cap=cap)
        if w_cal is not None:
            run("network_duals", w_cal, duals=(u0, v0), extra_sec=t_pred, cap=cap)
            run("exact_duals", w_cal, duals=found[0], cap=cap)
        # the learned start on the same uploaded problem: its dual prediction, the
        # KKT evaluation of the predicted duals and the solve, timed as the rows above
        prof = P.SolveProfile()
        out, info = learned.solve_warm(P.default_parameters(1e-4, record_iteration_stats=False), p, inst.x, inst.y,
                                       net, storage=_lib.STORAGE_F32, solver=dev, profile=prof)
        dev.close()
        ci = out.iteration_stats[-1].convergence_information[0]
        row["learned_start"] = {"seconds": info.predict_sec + info.probe_sec + prof.solve_sec,
                                "predict_sec": info.predict_sec, "probe_sec": info.probe_sec,
                                "iterations": out.iteration_count,
                                "termination": out.termination_string, "objective": ci.primal_objective,
                                "duals_used": info.used_duals, "initial_primal_weight": info.primal_weight,
                                "fell_back": info.fell_back}
        rows.append(row)
        print(f"# c5 instance {k}: " + ", ".join(f"{key} {row[key]['iterations']} it" for key in row
                                                 if key != "seed"), file=sys.stderr, flush=True)
    net.close()
    names = [k for k in rows[0] if k != "seed"]
    summary = {}
    for name in names:
        sec = sum(r[name]["seconds"] for r in rows)
        its = [r[name]["iterations"] for r in rows]
        summary[name] = {"seconds_total": sec, "iterations_total": int(sum(its)),
                         "optimal": int(sum(r[name]["termination"] == "OPTIMAL" for r in rows)),
                         "speedup_vs_default": sum(r["default_weight"]["seconds"] for r in rows) / sec,
                         "iteration_ratio_median_vs_default": float(np.median(
                             [r["default_weight"]["iterations"] / max(1, r[name]["iterations"]) for r in rows]))}
        if "calibrated_weight" in rows[0]:
            summary[name]["iteration_ratio_median_vs_calibrated"] = float(np.median(
                [r["calibrated_weight"]["iterations"] / max(1, r[name]["iterations"]) for r in rows]))
    # roofline of the dominant kernel on this workload (FP32 step kernel at 4096 x 4096; its 201 MB
    # working set fits in the 256 MB Infinity Cache, so the rate can exceed the HBM figure)
    inst = instances.gmm_dirichlet(4096, seed=900200)
    dev = _lib.DeviceSolver(4096, 4096, storage=_lib.STORAGE_F32)
    dev.upload(inst.problem.cost_matrix, inst.problem.source_distribution, inst.problem.target_distribution)
    g = step_measurement(dev, 4096, 4096, 4, steps, warmup, prob=inst.problem, storage=_lib.STORAGE_F32, start=start)
    dev.close()
    return {"metric": "learned warm start: time-to-1e-4 on held-out 4096x4096 GMM instances (FP32 storage)",
            "unit": "s", "higher_is_better": False, "data": "synthetic", "n_gpus": 1,
            "config": {"workload": "gmm_dirichlet 4096x4096 (3-8 component 2-D mixtures, Dirichlet masses, "
                                   "squared Euclidean), held-out seeds 900200+", "instances": len(rows),
                       "storage": "f32"},
            "value": summary.get("learned_start", {}).get("seconds_total"), "summary": summary,
            "calibrated_weight": w_cal, "step_kernel": g,
            "roofline": roofline(g["step_kernel_GBs"], g["bytes_per_step"], g["step_kernel_us"],
                                 "step_kernel<float> at 4096x4096 (working set fits in the Infinity Cache)"),
            "rows": rows}


def solve_to_tol_params(prob, storage, params, dev, warm_start=None, duals_out=None) -> dict:
    """One solve; duals_out (a list), if given, receives the returned duals (u, v)."""
    import pdot_amd as P
    prof = P.SolveProfile()
    out = P.optimize(params, prob, storage=storage, solver=dev, profile=prof, warm_start=warm_start)
    ci = out.iteration_stats[-1].convergence_information[0]
    if duals_out is not None:
        duals_out.append((out.dual_source_solution.copy(), out.dual_target_solution.copy()))
    return {"seconds": prof.solve_sec, "iterations": out.iteration_count, "termination": out.termination_string,
            "objective": ci.primal_objective, "rel_kkt": P.relative_kkt_error(ci)}


def bench_warm_start(args) -> int:
    print(json.dumps(warm_start_table(args.instances, start=args.start_iteration)))
    return 0


def bench_mlp(args) -> int:
    """The warm-start network's kernels at the c5 scale (4096 source + 4096
    target points): forward passes (dual prediction), training steps (loss,
    gradient and Adam) and weight-head predictions, timed on the host around
    the synchronous calls.  The MFMA products' rates come from the kernel
    trace (rocprofv3) of the same command; this line reports the algorithmic
    flops per call."""
    import numpy as np
    from pdot_amd import instances, learned
    inst = instances.gmm_dirichlet(4096, seed=900200)
    p = inst.problem
    feats = learned.features(inst.x, inst.y, p.source_distribution, p.target_distribution)
    ns = nt = 4096
    N = ns + nt
    net = learned.WarmStartNet(learned.load_params())
    target = np.zeros(N, dtype=np.float32)
    for _ in range(max(1, args.warmup)):
        net.forward(feats, ns, nt)
        net.loss_grad(feats, ns, nt, target)
        net.adam_step(0.0)
    t0 = time.time()
    for _ in range(args.steps):
        net.forward(feats, ns, nt)
    fwd = (time.time() - t0) / args.steps
    t0 = time.time()
    for _ in range(args.steps):
        net.loss_grad(feats, ns, nt, target)
        net.adam_step(0.0)
    step = (time.time() - t0) / args.steps
    t0 = time.time()
    for _ in range(args.steps):
        net.log_weight(feats, ns, nt)
    head = (time.time() - t0) / args.steps
    net.close()
    W, F = learned.HIDDEN, 16  # the first layer's inputs padded to the MFMA k-step
    fwd_flops = 2 * N * (F * W + 2 * W * W + W * 32)  # layer 1, layers 2 and 3, the output (32-column tile)
    print(json.dumps({"metric": "warm-start network at 4096 + 4096 points: forward and training step",
                      "unit": "ms", "higher_is_better": False, "data": "synthetic", "n_gpus": 1,
                      "config": {"workload": "gmm_dirichlet 4096x4096 seed 900200 features", "points": N,
                                 "hidden": W},
                      "value": step * 1e3, "forward_ms": fwd * 1e3, "train_step_ms": step * 1e3,
                      "weight_head_ms": head * 1e3, "forward_mfma_flops": fwd_flops,
                      "largest_product_flops": 2 * N * W * W,
                      "forward_TFLOPs": fwd_flops / fwd / 1e12}))
    return 0


def tight_pair(n: int, verbosity: int = 0) -> dict:
    """FP32 storage against the FP64 oracle on the n x n Gaussian-shift
    instance: the FP64 solve tightened to 1e-8 and the FP32-storage solve to
    1.5e-8; value = the relative difference of the two objectives <C, X>."""
    import numpy as np
    import pdot_amd as P
    from pdot_amd import _lib, instances
    prob = instances.gaussian_shift(n)
    out = {"metric": "objective of FP32 storage solved to 1.5e-8 vs the FP64 solve to 1e-8 (relative)",
           "unit": "relative", "higher_is_better": False, "data": "synthetic", "n_gpus": 1,
           "config": {"workload": f"gaussian_shift {n}x{n}", "n": n, "m": n}}
    for name, storage, tol in (("f64_oracle", _lib.STORAGE_F64, 1e-8), ("f32", _lib.STORAGE_F32, 1.5e-8)):
        params = P.default_parameters(tol, verbosity=verbosity, record_iteration_stats=verbosity > 0)
        t0 = time.time()
        sol = P.optimize(params, prob, storage=storage)
        wall = time.time() - t0
        ci = sol.iteration_stats[-1].convergence_information[0]
        out[name] = {"tolerance": tol, "termination": sol.termination_string, "iterations": sol.iteration_count,
                     "seconds": wall, "rel_kkt": P.relative_kkt_error(ci),
                     "objective": float(np.sum(prob.cost_matrix * sol.primal_solution))}
        sys.stdout.flush()
    out["value"] = abs(out["f32"]["objective"] - out["f64_oracle"]["objective"]) / abs(out["f64_oracle"]["objective"])
    return out


def storage_bound(prob, objective: float) -> dict:
    """How far storing C in FP32 can move the optimum: with masses summing to
    1, every feasible plan has sum(X) = 1, so |OPT(fl32(C)) - OPT(C)| <=
    max |C - fl32(C)|; relative to the FP64 objective given."""
    import numpy as np
    C = prob.cost_matrix
    d = 0.0
    for j0 in range(0, C.shape[1], 1000):
        blk = C[:, j0:j0 + 1000]
        d = max(d, float(np.abs(blk - blk.astype(np.float32).astype(np.float64)).max()))
    return {"max_abs_c_minus_fl32_c": d, "objective": objective, "relative_bound": d / abs(objective)}


def bench_precision(args) -> int:
    """FP32 storage against the FP64 oracle at 10000 x 10000 (gaussian shift),
    each solve printing its progress table every 50 evaluations.  About 15
    minutes on one MI355X."""
    print(json.dumps(tight_pair(args.n, verbosity=3)))
    return 0


def extras(args, prob, f32_roofline, f64_objective=None) -> dict:
    """The other workloads' entries of the default run, each printed to stderr
    as it finishes."""
    from pdot_amd import _lib
    out = {}

    def step(name, fn):
        t0 = time.time()
        out[name] = fn()
        out[name]["wall_sec"] = time.time() - t0
        print(f"# {name} done in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    if not args.no_cpu:
        step("c1_cpu_1000", lambda: cpu_c1(start=args.start_iteration))

    def c2():
        r = tight_solve(prob, _lib.STORAGE_F32, 1e-6)
        r.update({"metric": "headline instance, FP32 storage, time to 1e-6 relative KKT", "unit": "s",
                  "tolerance": 1e-6, "value": r["seconds"], "roofline": f32_roofline})
        return r
    step("c2_f32_1e-6", c2)

    def precision():
        # the 10k pair takes about 15 minutes (--workload precision); here the same family at
        # 1000 x 1000 (about 25 s), and the storage bound on the 10k instance itself
        r = tight_pair(1000)
        r["metric"] = ("FP32 storage vs the FP64 oracle: objective of the 1.5e-8 FP32-storage solve vs the "
                       "1e-8 FP64 solve at 1000x1000 (relative), and the bound on the 10k instance")
        r["storage_bound_10k"] = storage_bound(prob, f64_objective)
        return r
    if f64_objective is not None:
        step("c2_precision", precision)
    step("c3_l1_mixture_30k", lambda: l1_mixture(30000, max(10, args.steps // 2), args.warmup, point_cloud=False,
                                                  start=args.start_iteration))
    step("c4_batched_512", lambda: batched(args.steps))
    step("c5_warm_start_4096", lambda: warm_start_table(args.instances, start=args.start_iteration))
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--start-iteration", type=int, default=2048,
                    help="solve iterations (tolerance 1e-4, untimed) before the warmup and timed passes, so that "
                         "these see the iterates' zeros as they are at that point of a solve (the zero start "
                         "has no nonzero entries)")
    ap.add_argument("--n", type=int, default=10000)
    ap.add_argument("--no-solve", action="store_true", help="skip the time-to-tolerance solves (and the extras)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-extra", action="store_true", help="skip the c1, c2 1e-6, c2 precision, c3, c4 and c5 entries")
    ap.add_argument("--cpu-iters", type=int, default=80, help="iteration limit of the CPU baseline sample (about 15 s)")
    ap.add_argument("--instances", type=int, default=8, help="held-out instances of the warm-start workload")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run a plain device copy of this many MiB (counter calibration)")
    ap.add_argument("--workload", default="gaussian_shift",
                    choices=("gaussian_shift", "l1_mixture_3d", "batched_uniform", "gmm_warm_start", "precision", "mlp"),
                    help="gaussian_shift: the headline 10k x 10k measurement (plus the other workloads' entries); "
                         "the others measure one workload alone: the 30k x 30k FP32-storage stress case (with its "
                         "point-cloud form), the 512-instance batch, the learned warm start, the warm-start network's "
                         "kernels and (about 15 minutes) FP32 storage against the FP64 1e-8 solve at 10k x 10k")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the library has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        return 2
    if args.workload == "l1_mixture_3d":
        return bench_l1_mixture(args)
    if args.workload == "batched_uniform":
        return bench_batched(args)
    if args.workload == "precision":
        return bench_precision(args)
    if args.workload == "gmm_warm_start":
        return bench_warm_start(args)
    if args.workload == "mlp":
        return bench_mlp(args)

    from pdot_amd import _lib, instances

    n = m = args.n
    prob = instances.gaussian_shift(n)
    results = {}
    devs = {}
    for storage, name, elem in ((_lib.STORAGE_F64, "f64", 8), (_lib.STORAGE_F32, "f32", 4)):
        dev = _lib.DeviceSolver(n, m, storage=storage)
        t0 = time.time()
        dev.upload(prob.cost_matrix, prob.source_distribution, prob.target_distribution)
        up = time.time() - t0
        results[name] = step_measurement(dev, n, m, elem, args.steps, args.warmup, prob=prob, storage=storage,
                                         start=args.start_iteration)
        results[name].update({"iteration_GBs": results[name]["bytes_per_step"] / (results[name]["ms_per_step"] * 1e-3) / 1e9,
                              "launch": dev.get_launch(), "upload_sec": up, "upload_GBs_pcie": 8 * n * m / up / 1e9})
        devs[name] = dev

    # point-cloud cost form (C formed in registers, never stored): 3 n*m streams per step on average
    points = {}
    pts = instances.gaussian_shift_points(n, m)
    for storage, name, elem in ((_lib.STORAGE_F64, "f64", 8), (_lib.STORAGE_F32, "f32", 4)):
        dev = _lib.DeviceSolver(n, m, storage=storage)
        dev.upload_points(pts)
        points[name] = step_measurement(dev, n, m, elem, args.steps, args.warmup, matrix_streams=3, prob=pts,
                                        storage=storage, start=args.start_iteration)
        points[name]["device_GB"] = dev.device_bytes() / 1e9
        if not args.no_solve:
            points[name]["time_to_1e-4"] = solve_to_tol(pts, storage, 1e-4, dev=dev)
        dev.close()

    solves = {}
    if not args.no_solve:
        for name, storage in (("f64", _lib.STORAGE_F64), ("f32", _lib.STORAGE_F32)):
            solves[name] = solve_to_tol(prob, storage, 1e-4, dev=devs[name])
        solves["speedup_f32_vs_f64"] = solves["f64"]["seconds"] / solves["f32"]["seconds"]
        solves["objective_rel_diff_f32_vs_f64"] = (abs(solves["f32"]["objective"] - solves["f64"]["objective"])
                                                   / abs(solves["f64"]["objective"]))
        # the learned starting state's initial primal weight (calibrated on Gaussian-mixture
        # solves; the reference's default primal_importance is 1.0)
        from pdot_amd import learned
        w0 = learned.load_initial_primal_weight()
        if w0 is not None:
            for name, storage in (("f64", _lib.STORAGE_F64), ("f32", _lib.STORAGE_F32)):
                r = solve_to_tol(prob, storage, 1e-4, dev=devs[name], do_round=False, primal_importance=w0)
                r["primal_importance"] = w0
                solves[f"{name}_calibrated_primal_weight"] = r
            solves["speedup_calibrated_f64_vs_f64"] = solves["f64"]["seconds"] / solves["f64_calibrated_primal_weight"]["seconds"]
            solves["speedup_calibrated_f32_vs_f64"] = solves["f64"]["seconds"] / solves["f32_calibrated_primal_weight"]["seconds"]
    copy_probe = None
    if args.copy_probe_mib:
        nbytes = args.copy_probe_mib << 20
        ms = devs["f64"].copy_probe(nbytes, reps=5)
        copy_probe = {"bytes_per_copy": 2 * nbytes, "ms": ms, "GBs": 2 * nbytes / (ms * 1e-3) / 1e9,
                      "launches": 6}
    for d in devs.values():
        d.close()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(prob, args.cpu_iters)

    f64, f32 = results["f64"], results["f32"]
    f32_roof = roofline(f32["step_kernel_GBs"], f32["bytes_per_step"], f32["step_kernel_us"],
                        "step_kernel<float>", measured_traffic(n, m, "f32"))
    line = {
        "metric": METRIC,
        "value": f64["iters_per_s"],
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": f64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"gaussian_shift {n}x{m} FP64 (squared Euclidean / max, U(0,1) masses, seed 2000)",
                   "n": n, "m": m, "storage": "f64", "parallelism": "single GPU"},
        "roofline": roofline(f64["step_kernel_GBs"], f64["bytes_per_step"], f64["step_kernel_us"],
                             "step_kernel<double>", measured_traffic(n, m, "f64")),
        "cpu_baseline": cpu,
        "iteration_GBs": f64["iteration_GBs"],
        "f64_step": {k: f64[k] for k in ("start_iteration", "zero_share_current", "zero_share_average",
                                          "bytes_per_step", "dense_bytes_per_step", "dense_equivalent_GBs",
                                          "step_kernel_us", "launch")},
        "f32_storage": dict(f32, roofline=f32_roof),
        "time_to_1e-4": solves or None,
        "point_cloud_cost": points,
    }
    if copy_probe:
        line["copy_probe"] = copy_probe
    if not (args.no_solve or args.no_extra) and n == 10000:
        line["workloads"] = extras(args, prob, f32_roof, (solves.get("f64") or {}).get("objective"))
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
